"""Central-difference Jacobian check of an oracle group (the TESTJAC check of
NumericalJacobian::testDAEMatrices, N_DEV_NumericalJacobian.C:175-582, with
central differences).  Used by the oracle pin tests."""
import numpy as np


def fd_check(p, x0, flags, rtol=1e-5, h_rel=1e-6, vecs=("f", "q"), exclude=None):
    """Perturb every instance's k-th local node at once (private nodes) and
    compare column sums J*2h with F(x+h) - F(x-h).  Returns the list of
    (vec, k, row, fd, analytic) failures."""
    p.set_inputs(nextSol=x0)
    base, _ = p.load(**flags)
    J = {"f": base["dFdx"], "q": base["dQdx"]}
    Jmag = {"f": p.mag["dFdx"].copy(), "q": p.mag["dQdx"].copy()}
    rp, ci = p.topo.row_ptr, p.topo.col_idx
    rows_of = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    nodes_per = p.counts[0] + p.counts[1]
    offs = np.concatenate([[0], np.cumsum(nodes_per)])
    bad = []
    for k in range(int(nodes_per.max())):
        has = nodes_per > k
        cols = p.topo.node_lids[offs[:-1][has] + k]
        h = h_rel * (1 + np.abs(x0[cols]))
        hx = np.zeros(p.topo.n_nodes)
        hx[cols] = h
        hmax = float(h.max())
        p.set_inputs(nextSol=x0 + hx)
        rp_, _ = p.load(**flags)
        mp_ = {v: p.mag[v].copy() for v in vecs}
        p.set_inputs(nextSol=x0 - hx)
        rm_, _ = p.load(**flags)
        mm_ = {v: p.mag[v].copy() for v in vecs}
        for vec in vecs:
            fd = rp_[vec] - rm_[vec]
            ana = np.zeros(p.topo.n_nodes)
            np.add.at(ana, rows_of, J[vec] * 2 * hx[ci])
            scale = np.maximum(np.abs(fd), np.abs(ana))
            # rounding allowance: F is a sum of terms that can cancel, so its
            # rounding error scales with the sum of the terms' magnitudes
            fmag = np.maximum(mp_[vec], mm_[vec])
            # the stamped entries carry rounding of their own terms' sizes
            # (e.g. a 1e3 S contact conductance stamped into a collapsed node)
            # and of the row's other entries: a conductance stamped as a
            # difference of nearly equal terms keeps only their rounding
            amag = np.zeros(p.topo.n_nodes)
            np.add.at(amag, rows_of, Jmag[vec] * 2 * hmax)
            ok = np.abs(fd - ana) <= rtol * scale + 1e-13 * fmag + 1e-12 * amag + 1e-30
            if exclude is not None:
                ok |= exclude
            for r in np.nonzero(~ok)[0]:
                bad.append((vec, k, int(r), float(fd[r]), float(ana[r])))
    p.set_inputs(nextSol=x0)
    return bad
