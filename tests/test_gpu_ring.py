"""Ring-oscillator transient with the GPU loads against the same transient
with the oracle's loads (same driver, same linear algebra).

Bar: every node voltage at every time point within 1e-6 of the supply
(1e-6 relative), the same Newton iteration count, and the oscillation period
within 1e-6 relative.  The 3-stage ring is the small brute-force case; the
51-stage ring at 1 ps steps to 2 ns is the plain ring-oscillator workload.
"""
import numpy as np
import pytest

from ring_oracle import OracleLoader
from xyce_amd.ring import GpuLoader, Ring, crossings, period, transient

pytestmark = pytest.mark.gpu
WAVE_RTOL = 1e-6


def _compare(stages, t_stop, fused=False):
    ring = Ring(stages=stages)
    print(f"ring {stages} stages to {t_stop:g} s, fused={fused}", flush=True)
    tg, Vg, sg = transient(GpuLoader(ring, fused=fused), ring, t_stop=t_stop)
    tc, Vc, sc = transient(OracleLoader(ring), ring, t_stop=t_stop)
    err = np.max(np.abs(Vg - Vc)) / ring.vdd
    print(f"  newton {sg['newton_iterations']} vs {sc['newton_iterations']}, max |dV|/VDD {err:.3e}, "
          f"gpu load {sg['load_seconds'] / sg['loads'] * 1e6:.1f} us", flush=True)
    assert sg["newton_iterations"] == sc["newton_iterations"]
    assert err <= WAVE_RTOL
    return ring, tg, Vg, Vc


@pytest.mark.parametrize("fused", [False, True])
def test_ring3_brute_force(fused):
    _compare(3, 2e-10, fused)


@pytest.mark.parametrize("fused", [False, True])
def test_ring51_transient(fused):
    ring, t, Vg, Vc = _compare(51, 2e-9, fused)
    # the crossing times of every stage (a period needs two rising edges of
    # one node; in 2 ns each node crosses about twice)
    for k in range(ring.N):
        cg = crossings(t, Vg[:, k], ring.vdd / 2)
        cc = crossings(t, Vc[:, k], ring.vdd / 2)
        assert len(cg) == len(cc)
        np.testing.assert_allclose(cg, cc, rtol=WAVE_RTOL)
    pg, pc = period(t, Vg[:, 0], 0.5), period(t, Vc[:, 0], 0.5)
    assert pg is not None and abs(pg - pc) <= WAVE_RTOL * pc


def _cross_any(t, v, level):
    """Crossings of `level` in either direction, in time order (.MEASURE's
    CROSS= counting)."""
    return np.sort(np.concatenate([crossings(t, v, level, rising=True),
                                   crossings(t, v, level, rising=False)]))


def test_ring17_measured_period():
    """The shape of the reference's 17-stage ring example
    (utils/ADMS/examples/bsimcmg_107.0.0/benchmark_test/ringosc_17stg.sp):
    .IC on node 1 = 1 V, 1 ps steps, period = (7th - 1st crossing of
    v(1) = 0.5 V) / 3.  Here with BSIM4 default cards, run to 2 ns so seven
    crossings fall inside; GPU (fused loads) and oracle periods within 1e-6."""
    ring = Ring(stages=17, ic_node=0, ic_value=1.0)
    tg, Vg, _ = transient(GpuLoader(ring, fused=True), ring, t_stop=2e-9)
    tc, Vc, _ = transient(OracleLoader(ring), ring, t_stop=2e-9)
    cg, cc = _cross_any(tg, Vg[:, 0], 0.5), _cross_any(tc, Vc[:, 0], 0.5)
    assert len(cg) >= 7 and len(cg) == len(cc)
    pg, pc = (cg[6] - cg[0]) / 3, (cc[6] - cc[0]) / 3
    print(f"  17 stages: period {pg:.6e} s (oracle {pc:.6e}), {pg / 34:.3e} s per stage",
          flush=True)
    assert abs(pg - pc) <= WAVE_RTOL * pc


@pytest.mark.parametrize("fused", [False, True])
def test_ring_graph_replays_equal_eager(fused):
    """The plain loader's HIP-graph replay (one graph per solver-flag set,
    eager host Newton work and copies between the replays) gives the eager
    loads' waveforms bit for bit."""
    ring = Ring(stages=11)
    ta, Va, sa = transient(GpuLoader(ring, fused=fused), ring, t_stop=3e-10)
    loader = GpuLoader(ring, fused=fused, graph=True)
    tb, Vb, sb = transient(loader, ring, t_stop=3e-10)
    print(f"  graphs {len(loader.graphs)}, eager {sa['load_seconds'] / sa['loads'] * 1e6:.1f} us, "
          f"graph {sb['load_seconds'] / sb['loads'] * 1e6:.1f} us per load", flush=True)
    assert len(loader.graphs) >= 3
    np.testing.assert_array_equal(Va.view(np.int64), Vb.view(np.int64))
