"""The hand-traced tick cases of tests/test_oracle_kats.py run on the engine
instead of the oracle: the same tiny brains and the same hand-derived
expectations (leak toward rest, the adaptive leak, a GLIF spike with its reset, refractory
countdown and own-axon push, heartbeat spikes without a reset, the same-tick
snapshot rule, input pushes, propagation and the sentinel, a tick's GSOP
after a spike with and without dopamine), each case
through the per-tick entry point and through the batch graph."""
import numpy as np
import pytest

import test_oracle_kats as K
from axicor_amd.engine import IoDeviceBuffers, ShardEngine, VramState, upload_constant_memory

pytestmark = pytest.mark.gpu

CASES = {f.__name__[len("test_"):]: f for f in (
    K.test_leak_toward_rest_closed_form, K.test_adaptive_leak_mode_1, K.test_glif_spike_reset_refractory_and_push,
    K.test_heartbeat_spikes_without_reset, K.test_snapshot_rule_same_tick_push_invisible_to_integration,
    K.test_inputs_push_virtual_axons, K.test_propagation_reference_vectors)}
CASES["gsop_tick_after_a_spike_dop0"] = lambda: K.test_gsop_tick_after_a_spike(0)
CASES["gsop_tick_after_a_spike_dop100"] = lambda: K.test_gsop_tick_after_a_spike(100)


def engine_runner(use_graph: bool):
    """A stand-in for the KATs' oracle runner: uploads the tiny brain, runs T
    ticks on the engine and writes the state and heads back in place."""
    def run(state, pn, heads, A, lut, T=1, tick0=0, dop=0, bitmask=None, virt_off=0, n_virt=0):
        upload_constant_memory(lut)
        vram = VramState(pn, A)
        vram.upload_state(state)
        vram.upload_axon_heads(heads)
        eng = ShardEngine(vram)
        words = bitmask.shape[1] if bitmask is not None else 0
        io = IoDeviceBuffers(words, 0, 0, T)
        eng.step_day_phase_batch(T, io, bitmask, None, np.zeros(T, np.uint32), virt_off,
                                 n_virt if bitmask is not None else 0, None, 1, dop, tick0, use_graph=use_graph)
        eng.synchronize()
        state[:] = vram.download_state()
        heads[:] = np.asarray(vram.download_axons()).reshape(-1)
        return None, None
    return run


@pytest.mark.parametrize("use_graph", [False, True], ids=["per_tick", "graph"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_hand_traced_case_on_the_engine(gpu, monkeypatch, case, use_graph):
    monkeypatch.setattr(K, "run", engine_runner(use_graph))
    CASES[case]()
