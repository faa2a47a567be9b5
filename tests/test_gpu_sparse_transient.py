"""The device sparse transient (xyce_amd.sparse_transient): the flat logic
netlist read by the netlist reader, every Newton iteration's loads on the
GPU and its system solved by GMRES with ILU(0) on the GPU, against the
oracle's transient of the same netlist (oracle loads, the same integrator,
a direct sparse LU).  Bar: the waveforms of every node within 1e-6 of the
largest node voltage (the transient-waveform bar), over a DC operating
point, two backward-Euler steps and trapezoid steps through the first
input edges."""
import numpy as np
import pytest

from sparse_transient_oracle import OracleTransient

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_transistors,t_stop", [(300, 100e-12), (3000, 60e-12)])
def test_logic_transient_matches_oracle(n_transistors, t_stop):
    import torch
    import xyce_amd
    from xyce_amd import netlists
    from xyce_amd.netlist_reader import Circuit, Netlist
    from xyce_amd.sparse_transient import SparseTransient
    text, volts = netlists.logic_netlist(n_transistors)
    nl = Netlist(text)
    ctx = xyce_amd.Context(0)
    tr = SparseTransient(Circuit(ctx, nl))
    x0 = np.zeros(tr.n)
    for k, name in enumerate(nl.node_names):
        x0[k] = volts[int(name[1:])]
    rec = np.arange(nl.n_nodes)
    print(f"logic netlist {n_transistors}: {tr.n} unknowns, circuit nnz {tr.c.nnz}", flush=True)
    t, V, nw = tr.run(torch.from_numpy(x0).cuda(), t_stop, 1e-12, record=rec, progress=20)
    print(f"  device: Newton {nw.sum()} (DC {nw[0]}), GMRES {tr.stats['krylov']}", flush=True)
    ot = OracleTransient(nl)
    t_o, V_o, nw_o = ot.run(x0, t_stop, 1e-12, record=rec)
    err = np.max(np.abs(V - V_o)) / np.max(np.abs(V_o))
    print(f"  oracle: Newton {nw_o.sum()}; max waveform difference {err:.2e} of max |V|",
          flush=True)
    assert err <= 1e-6
    # the inputs switch inside the window, and the netlist responds
    assert np.max(np.abs(V_o[-1] - V_o[0])) > 0.1


def test_ilu_level_solve_equals_csrsv():
    """The level-scheduled triangular solves (launched, or replayed as a HIP
    graph) and rocSPARSE's csrsv apply the same ILU(0) factors: equal to
    rounding on the logic netlist's system, the graph bit-equal to the
    launched kernels on repeated replays."""
    import torch
    import xyce_amd
    from xyce_amd import netlists
    from xyce_amd.netlist_reader import Circuit, Netlist
    from xyce_amd.sparse_transient import SparseTransient
    text, volts = netlists.logic_netlist(3000)
    nl = Netlist(text)
    tr = SparseTransient(Circuit(ctx := xyce_amd.Context(0), nl))
    x0 = np.zeros(tr.n)
    for k, name in enumerate(nl.node_names):
        x0[k] = volts[int(name[1:])]
    tr.run(torch.from_numpy(x0).cuda(), 2e-12, 1e-12)
    levels = tr.ilu.config()
    print(f"level counts (L, U): {levels}", flush=True)
    b = torch.from_numpy(np.random.default_rng(5).standard_normal(tr.n)).cuda()
    tr.ilu.config(1)
    x1 = tr.ilu.solve(b).cpu().numpy()
    tr.ilu.config(2)
    x2 = tr.ilu.solve(b).cpu().numpy()
    x2b = tr.ilu.solve(b).cpu().numpy()
    tr.ilu.config(0)
    x0_ = tr.ilu.solve(b).cpu().numpy()
    tr.ilu.config(1)
    np.testing.assert_allclose(x1, x0_, rtol=1e-11, atol=1e-12 * np.max(np.abs(x0_)))
    # the graph replays the same kernels: bit for bit, on every replay
    np.testing.assert_array_equal(x2, x1)
    np.testing.assert_array_equal(x2b, x1)
    del ctx


def test_rc_step_follows_the_integrator_recurrences():
    """The device transient on a linear RC step: the closed-form backward
    Euler and trapezoid recurrences (the same pin as the oracle's)."""
    import torch
    import xyce_amd
    from test_sparse_transient_oracle import TEXT, _expected
    from xyce_amd.netlist_reader import Circuit, Netlist
    from xyce_amd.sparse_transient import SparseTransient
    nl = Netlist(TEXT)
    ctx = xyce_amd.Context(0)
    tr = SparseTransient(Circuit(ctx, nl), vtol=1e-13, krylov_tol=1e-13)
    out = nl.node("out")
    t, V, nw = tr.run(torch.zeros(tr.n, dtype=torch.float64, device="cuda"), 40e-12, 1e-12,
                      record=[out])
    np.testing.assert_allclose(V[:, 0], _expected(40, 1e-12, 1e-9), rtol=0, atol=1e-13)
