"""Baked shard files (axicor_amd/baked.py) in the reference's layouts: the
paths file's header and 64-byte-aligned matrix offset
(genesis-core/src/layout.rs:157-191), geom and pos files, and the Night run
on the daemon's memory-mapped geometry equal to the in-memory one."""
import os
import struct

import numpy as np

from axicor_amd import baked as K
from axicor_amd import brain as B
from axicor_amd import night as NT
from test_night import small_brain


def test_paths_offsets():
    # 16-byte header + one length byte per axon, padded to a multiple of 64.
    assert K.paths_matrix_offset(100) == 128 and K.paths_file_size(100) == 128 + 100 * 1024
    assert K.paths_matrix_offset(48) == 64
    assert K.paths_matrix_offset(49) == 128
    assert K.paths_matrix_offset(0) == 64


def test_bake_round_trip(tmp_path):
    b = B.generate(B.spec_c1(), with_paths=True)
    d = str(tmp_path)
    K.write_bake(d, b)
    A = b.total_axons
    raw = open(os.path.join(d, "shard.paths"), "rb").read()
    assert raw[:16] == struct.pack("<IIII", 0x50415448, 1, A, 256)
    assert len(raw) == K.paths_file_size(A)
    assert (np.frombuffer(raw[16 + A:K.paths_matrix_offset(A)], np.uint8) == 0).all()
    assert open(os.path.join(d, "shard.state"), "rb").read() == b.state.tobytes()
    assert os.path.getsize(os.path.join(d, "shard.axons")) == 32 * A
    assert os.path.getsize(os.path.join(d, "shard.geom")) == 8 * A
    assert os.path.getsize(os.path.join(d, "shard.pos")) == 4 * b.padded_n
    g = K.load_geometry(d, b.padded_n, A, b.spec.grid[0], b.spec.grid[1], writable=False)
    ref = NT.geometry_of(b)
    for f in ("lengths", "paths", "tips", "dirs", "soma_positions"):
        np.testing.assert_array_equal(np.asarray(getattr(g, f)), getattr(ref, f), err_msg=f)


def test_night_on_mapped_geometry(tmp_path):
    """The sprouting pass on the daemon's read-write maps of shard.paths and
    shard.geom gives the in-memory result and leaves the files updated."""
    b = small_brain(11, n=300, slots=24)
    v = b.view()
    v.flags[:100] |= 0x02  # these somas' local axons are nudged (paths and tips grow in the files)
    d = str(tmp_path)
    K.write_bake(d, b)
    types = NT.sprout_types_from_lut(b.lut)
    g = K.load_geometry(d, b.padded_n, b.total_axons, b.spec.grid[0], b.spec.grid[1])
    t1, w1 = v.targets.copy(), v.weights.copy()
    r1 = NT.sprout_pass(t1, w1, v.flags.copy(), v.soma_to_axon.copy(), b.padded_n, g, types, 2, b.spec.seed)
    ref = NT.geometry_of(b)
    t2, w2 = v.targets.copy(), v.weights.copy()
    r2 = NT.sprout_pass(t2, w2, v.flags.copy(), v.soma_to_axon.copy(), b.padded_n, ref, types, 2, b.spec.seed)
    assert r1.new_synapses == r2.new_synapses > 0
    np.testing.assert_array_equal(t1, t2)
    np.testing.assert_array_equal(w1, w2)
    g.paths.flush()
    lengths, paths = K.read_paths(os.path.join(d, "shard.paths"))
    np.testing.assert_array_equal(lengths, ref.lengths)
    np.testing.assert_array_equal(paths, ref.paths)
    assert not np.array_equal(ref.lengths, b.axon_len)  # nudged axons grew
