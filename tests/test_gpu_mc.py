"""Batched Monte Carlo ring oscillators: the batch (device loads of every
ring together, device Newton solve per ring) against each ring run alone
through the plain drop-in (GPU loads, host Newton) and through the oracle.

Bar: every ring's node-0 waveform within 1e-6 of the supply of the plain
run and of the oracle run (the two Newton loops stop on the same 1e-9 V
update bound, the batch when every ring meets it), and the periods within
1e-6 relative.
"""
import numpy as np
import pytest

from ring_oracle import OracleLoader
from xyce_amd.mc import BatchTransient, RingBatch
from xyce_amd.ring import GpuLoader, period, transient

pytestmark = pytest.mark.gpu
WAVE_RTOL = 1e-6


def test_batch_matches_plain_and_oracle():
    batch = RingBatch(stages=[3, 5, 11, 7, 3, 9, 13, 5], seed=31)
    print(f"batch of {batch.n_rings} rings, stages {batch.stages.tolist()}", flush=True)
    bt = BatchTransient(batch)
    t, V0 = bt.run(t_stop=3e-10)
    print(f"  batched: {bt.newton_iterations} Newton iterations", flush=True)
    for r in range(batch.n_rings):
        ring = batch.ring(r)
        tg, Vg, _ = transient(GpuLoader(ring), ring, t_stop=3e-10)
        to, Vo, _ = transient(OracleLoader(ring), ring, t_stop=3e-10)
        eg = np.max(np.abs(V0[:, r] - Vg[:, ring.ic_node])) / ring.vdd
        eo = np.max(np.abs(V0[:, r] - Vo[:, ring.ic_node])) / ring.vdd
        print(f"  ring {r} ({ring.N} stages): vs plain {eg:.2e}, vs oracle {eo:.2e}", flush=True)
        assert eg <= WAVE_RTOL and eo <= WAVE_RTOL


def test_graph_replays_equal_eager():
    """The Newton iteration captured as HIP graphs (one per flag set, replayed
    every iteration with eager copies between) gives the eager run's bits."""
    batch = RingBatch(stages=[11, 5, 21, 7], seed=33)
    ta, Va = BatchTransient(batch, use_graph=False).run(t_stop=2e-10)
    bt = BatchTransient(batch, use_graph=True)
    tb, Vb = bt.run(t_stop=2e-10)
    print(f"  graphs captured: {len(bt.graphs)}", flush=True)
    assert len(bt.graphs) >= 4
    np.testing.assert_array_equal(Va.view(np.int64), Vb.view(np.int64))


def test_batch_periods_match_oracle():
    """Rings long enough to oscillate for two periods within the run."""
    batch = RingBatch(stages=[11, 15, 21], seed=32)
    bt = BatchTransient(batch)
    t, V0 = bt.run(t_stop=1.6e-9)
    for r in range(batch.n_rings):
        ring = batch.ring(r)
        to, Vo, _ = transient(OracleLoader(ring), ring, t_stop=1.6e-9)
        pb = period(t, V0[:, r], ring.vdd / 2)
        po = period(to, Vo[:, ring.ic_node], ring.vdd / 2)
        print(f"  ring {r} ({ring.N} stages): period {pb:.6e} vs oracle {po:.6e}", flush=True)
        assert pb is not None and po is not None
        assert abs(pb - po) <= WAVE_RTOL * po


def test_trapezoid_batch_matches_plain_and_oracle():
    """The one-step integrator's trapezoid rule (two backward-Euler steps,
    then order 2) in xa_batch_newton_step_ex against the same rings through
    the plain drop-in and the oracle with the same integrator; replayed from
    HIP graphs, and the periods of rings that oscillate twice."""
    batch = RingBatch(stages=[3, 5, 11, 7, 15], seed=34)
    print(f"trapezoid batch of {batch.n_rings} rings, stages {batch.stages.tolist()}", flush=True)
    bt = BatchTransient(batch, use_graph=True)
    t, V0 = bt.run(t_stop=1.2e-9, method="trap")
    print(f"  batched: {bt.newton_iterations} Newton iterations, {len(bt.graphs)} graphs",
          flush=True)
    tb, Vbe = BatchTransient(batch).run(t_stop=1.2e-9)
    for r in range(batch.n_rings):
        ring = batch.ring(r)
        tg, Vg, _ = transient(GpuLoader(ring, fused=True, graph=True), ring, t_stop=1.2e-9,
                              method="trap")
        to, Vo, _ = transient(OracleLoader(ring), ring, t_stop=1.2e-9, method="trap")
        eg = np.max(np.abs(V0[:, r] - Vg[:, ring.ic_node])) / ring.vdd
        eo = np.max(np.abs(V0[:, r] - Vo[:, ring.ic_node])) / ring.vdd
        # the integrators differ by far more than the parity bar
        d_be = np.max(np.abs(V0[:, r] - Vbe[:, r])) / ring.vdd
        print(f"  ring {r} ({ring.N} stages): vs plain {eg:.2e}, vs oracle {eo:.2e}, "
              f"trap vs BE {d_be:.2e}", flush=True)
        assert eg <= WAVE_RTOL and eo <= WAVE_RTOL
        assert d_be > 1e3 * WAVE_RTOL
        pb = period(t, V0[:, r], ring.vdd / 2)
        po = period(to, Vo[:, ring.ic_node], ring.vdd / 2)
        if po is not None:
            assert pb is not None and abs(pb - po) <= WAVE_RTOL * po


def test_trapezoid_needs_previous_f():
    """The entry point refuses a trapezoid step without F_prev (host check)."""
    import ctypes as C
    from xyce_amd import Context
    from xyce_amd._abi import XA_ERR_ARG, XA_STEP_TRAP, lib
    ctx = Context(0)
    one = C.c_void_p(8)   # never dereferenced: the call fails its argument checks
    args = [ctx.h, 1, 4, 1, 1] + [one] * 11 + [None] + [one, one, C.c_double(1e-12),
                                                        XA_STEP_TRAP, one, one, one]
    assert lib().xa_batch_newton_step_ex(*args) == XA_ERR_ARG
