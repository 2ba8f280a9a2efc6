"""The Night's shared-memory contract (axicor_amd/shm.py) on the CPU: the
64-byte header against the reference client's struct layout
(genesis-client/test_checkpointing.py:37-42) and ipc.rs's offsets, the
request and reply bytes, file-backed regions, and the daemon's side run on a
region equal to the sprouting pass run directly."""
import os
import struct

import numpy as np
import pytest

from axicor_amd import night as NT
from axicor_amd import shm as S
from test_night import small_brain


def test_header_layout_matches_the_reference_struct():
    n, zone = 10_000, 0xCAFEBABE
    h = S.new_header(zone, n, 12_345)
    w_off, t_off = 64, 64 + n * 256
    h_off = t_off + n * 512
    p_off = h_off + 10_000 * 20
    f_off = p_off + 10_000 * 8
    # The client test packs the same 17 fields in this order (values of ipc.rs's ShmHeader::new).
    want = struct.pack("<IBBHIIIIQIIIIIIII", 0x47454E53, 2, 0, 0, n, 128, w_off, t_off, 0, 12_345, h_off, 0, zone,
                       p_off, 0, 0, f_off)
    assert h.tobytes() == want
    assert S.shm_size(n) == f_off + n
    assert S.shm_file_path(zone) == "/dev/shm/genesis_shard_CAFEBABE"


def test_validate_messages():
    h = S.new_header(1, 64, 64)[0]
    assert S.validate(h) is None
    for field, bad, msg in (("magic", 0, "SHM magic mismatch"), ("version", 3, "SHM version mismatch"),
                            ("dendrite_slots", 64, "SHM dendrite_slots != 128")):
        g = h.copy()
        g[field] = bad
        assert S.validate(g) == msg
    with pytest.raises(ValueError):
        S.ShmRegion(np.zeros(S.shm_size(64), np.uint8))


def test_request_and_reply_bytes():
    req = S.bake_request(0xABCD, -7, 4, ghost_origins=[11, 22, 33], current_tick=500)
    assert len(req) == 16 + 12
    assert struct.unpack("<IIIhH", req[:16]) == (0x42414B45, 0xABCD, 500, -7, 4)
    acks = np.zeros(2, NT.ACK_DTYPE)
    acks["target_zone_hash"], acks["src_axon_id"], acks["dst_ghost_id"] = [5, 6], [7, 8], [9, 10]
    rep = S.bake_reply(acks)
    assert rep[:8] == struct.pack("<II", 0x424B4F4B, 2) and len(rep) == 8 + 24
    np.testing.assert_array_equal(S.parse_reply(rep), acks)
    assert len(S.parse_reply(S.bake_reply(np.zeros(0, NT.ACK_DTYPE)))) == 0
    with pytest.raises(ValueError):
        S.parse_reply(struct.pack("<II", 0xDEAD, 0))
    with pytest.raises(ValueError):
        S.parse_reply(rep[:-4])


def test_file_backed_region(tmp_path):
    path = os.path.join(tmp_path, S.shm_name(0x1234))
    r = S.ShmRegion.create(0x1234, 256, 300, path=path)
    r.weights[5, 7] = -42
    r.targets[127, 255] = 0xDEADBEEF
    r.flags[3] = 9
    r.header["epoch"] = 77
    r.set_state(S.ShmState.NIGHT_DONE)
    r._owner.flush()
    q = S.ShmRegion.open(path)
    assert q.weights[5, 7] == -42 and q.targets[127, 255] == 0xDEADBEEF and q.flags[3] == 9
    assert int(q.header["epoch"]) == 77 and int(q.header["state"]) == S.ShmState.NIGHT_DONE
    assert os.path.getsize(path) == S.shm_size(256)


def test_daemon_side_equals_the_sprouting_pass():
    """A region filled as the shard leaves it (flags before the prune, pruned
    matrices, incoming handovers) and a request: daemon_night changes the
    region exactly as the sprouting pass changes the same arrays."""
    b = small_brain(9, n=400, slots=24)
    v = b.view()
    n = b.padded_n
    types = NT.sprout_types_from_lut(b.lut)
    r = S.ShmRegion.create(0x77, n, b.total_axons)
    r.flags[:] = v.flags
    r.targets[:] = v.targets
    r.weights[:] = v.weights
    r.header["epoch"] = 3
    req = S.bake_request(0x77, 15, 3)
    reply = S.daemon_night(r, req, NT.geometry_of(b), v.soma_to_axon, types, b.spec.seed)
    acks = S.parse_reply(reply)
    t, w = v.targets.copy(), v.weights.copy()
    res = NT.sprout_pass(t, w, v.flags.copy(), v.soma_to_axon.copy(), n, NT.geometry_of(b), types, 3, b.spec.seed,
                         zone_hash=0x77, max_sprouts=3)
    assert res.new_synapses > 0
    np.testing.assert_array_equal(r.targets, t)
    np.testing.assert_array_equal(r.weights, w)
    assert int(r.header["state"]) == S.ShmState.NIGHT_DONE
    assert int(r.header["handovers_count"]) == len(res.handovers) and len(acks) == len(res.acks)
