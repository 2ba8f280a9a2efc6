"""The synthetic brain generator: determinism, layout contracts and the
distributions the workloads ask for."""
import numpy as np

from axicor_amd import brain as B
from axicor_amd._ffi import gen_lib
from oracle import oracle as O

SENT = 0x80000000


def test_entity_seed_same_as_oracle():
    for m, i in [(0, 0), (B.DEFAULT_MASTER_SEED, 7), (0x1234567890ABCDEF, 11), (2**64 - 1, 2**32 - 1)]:
        assert gen_lib().axc_gen_entity_seed(m, i) == O.lib().orc_entity_seed(m, i)


def test_deterministic_and_seed_sensitive():
    a = B.generate(B.spec_small(3000, 40, 100, seed=1))
    b = B.generate(B.spec_small(3000, 40, 100, seed=1))
    c = B.generate(B.spec_small(3000, 40, 100, seed=2))
    assert np.array_equal(a.state, b.state) and np.array_equal(a.heads, b.heads)
    assert not np.array_equal(a.state, c.state)


def test_c1_shape_and_contracts():
    b = B.generate(B.spec_c1())
    v = b.view()
    n = 4096
    assert b.padded_n == 4096 and b.total_axons == 4096 + 256
    assert ((v.targets[:64, :n] != 0).all()) and (v.targets[64:, :] == 0).all()
    types = v.flags[:n] >> 4
    assert abs((types == 1).mean() - 0.2) < 0.03
    lut = b.lut
    rest, thr = lut["rest_potential"][types], lut["threshold"][types]
    assert ((v.voltage[:n] >= rest) & (v.voltage[:n] < thr)).all()
    assert (v.soma_to_axon[:n] == np.arange(n)).all()
    axon = (v.targets[:64, :n] & 0xFFFFFF) - 1
    seg = v.targets[:64, :n] >> 24
    assert (seg < b.axon_len[axon]).all()
    w = v.weights[:64, :n].astype(np.int32)
    assert ((np.abs(w) >= 64) & (np.abs(w) <= 512)).all()
    inhib_src = (axon < n) & (lut["is_inhibitory"][(v.flags[np.minimum(axon, n - 1)] >> 4)] == 1)
    assert (w[inhib_src] < 0).all() and (w[~inhib_src] > 0).all()
    h = b.heads.reshape(-1, 8)
    assert (h[:, 1:] == SENT).all()
    assert (h[:n, 0] == (SENT - b.axon_len[:n].astype(np.uint32))).all()
    assert (b.axon_len[:n] >= 1).all() and (b.axon_len[:n] <= 32).all()


def test_targets_are_distance_biased():
    b = B.generate(B.spec_c1())
    v = b.view()
    n = 4096
    axon = ((v.targets[:64, :n] & 0xFFFFFF) - 1).astype(np.int64)
    local = axon < n
    pos = b.soma_pos
    def xyz(p):
        return np.stack([p & 0x3FF, (p >> 10) & 0x3FF, (p >> 20) & 0xFF], -1).astype(np.int64)
    me = xyz(np.broadcast_to(pos[:n], axon.shape)[local])
    them = xyz(pos[axon[local]])
    cheb = np.abs(me - them).max(-1)
    assert 1.5 < cheb.mean() < 3.5   # geometric law of mean 3, clamped at the grid edge


def test_inputs_bernoulli_rate_and_determinism():
    x = B.gen_inputs(5, 10_000, 50_000, 0, 50)
    bits = np.unpackbits(x.view(np.uint8), bitorder="little").reshape(50, -1)[:, :10_000]
    assert abs(bits.mean() - 0.05) < 0.003
    assert np.array_equal(x, B.gen_inputs(5, 10_000, 50_000, 0, 50))
    assert np.array_equal(x[10:], B.gen_inputs(5, 10_000, 50_000, 10, 40))


def test_c5_scale_point_and_its_subshard():
    full = B.spec_c5()
    assert (full.n_neurons, full.slots_filled, full.grid) == (12_000_000, 128, (256, 256, 128))
    assert full.n_inputs == 240_000 and full.input_ppm == 10_000 and full.axon_len == (1, 255)
    assert len(full.type_ppm) == 6 and sum(full.type_ppm) == 1_000_000
    assert (full.lut["threshold"][:6] > 0).all() and (full.lut["threshold"][6:] == 0).all()
    assert B.sizes(full) == (12_000_000, 12_000_000 + 240_000)
    sub = B.spec_c5(200_000)
    assert sub.grid == (64, 64, 32) and sub.n_inputs == 4_000 and sub.seed == full.seed
    b = B.generate(sub)
    v = b.view()
    n = 200_000
    assert (v.targets[:128, :n] != 0).all()
    types = v.flags[:n] >> 4
    assert types.max() == 5
    assert abs((types >= 3).mean() - 0.2) < 0.01  # 8% + 7% + 5% inhibitory
    axon = (v.targets[:128, :n] & 0xFFFFFF) - 1
    assert (axon < b.total_axons).all()
    assert ((v.targets[:128, :n] >> 24) < b.axon_len[axon]).all()
