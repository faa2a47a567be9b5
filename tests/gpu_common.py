"""Shared drivers of the GPU parity tests (imported by tests/test_gpu_*.py)."""
import numpy as np

from harness import Pair, assert_parity

PDT = 2.0e9   # 1/dt of a 0.5 ns step
# Newton iterations of a DC operating point (junction initialisation, then a
# plain iteration) and of a transient step (first iteration of the first
# step, then later iterations).
ITERS = [dict(newtonIter=0, dcopFlag=1, initJctFlag=1),
         dict(newtonIter=1, dcopFlag=1),
         dict(newtonIter=0, transientFlag=1, initTranFlag=1, pdt=PDT),
         dict(newtonIter=1, transientFlag=1, pdt=PDT),
         dict(newtonIter=2, transientFlag=1, pdt=PDT)]
NOISE_RUNS = 6


def run_parity(p, rng, iters=ITERS, fused=False):
    """Load `p` on both back ends over `iters` with fresh random voltages and
    check every output with assert_parity; prints a per-output summary."""
    p.set_inputs(currSto=rng.uniform(-0.5, 1.5, p.topo.n_sto),
                 currSta=rng.uniform(-1e-15, 1e-15, p.topo.n_sta),
                 lastSta=rng.uniform(-1e-15, 1e-15, p.topo.n_sta))
    summary = {}
    carry = {}   # noise of state entries a later call leaves unwritten
    for it, flags in enumerate(iters):
        print(f"  iteration {it}: {flags}", flush=True)
        p.set_inputs(nextSol=rng.uniform(-0.5, 1.5, p.topo.n_nodes))
        ro, rg = p.load(fused=fused, **flags)
        for k in ro:
            noise = p.noise.get(k)
            if k in ("currSta", "lastSta") and noise is not None:
                noise = np.maximum(noise, carry.get(k, 0.0))
                carry[k] = noise
            st = assert_parity(rg[k], ro[k], what=f"iter {it} {flags} {k}",
                               scale=p.mag.get(k), noise=noise)
            agg = summary.setdefault(k, [0, 0, 0, 0.0])
            agg[0] += st["n"]
            agg[1] += int(round(st["bitexact"] * st["n"]))
            agg[2] += st["outliers"]
            agg[3] = max(agg[3], st["max_rel"])
    for k, (n, exact, outl, mx) in summary.items():
        print(f"  {k:12s} entries {n:8d} bit-exact {exact / max(n, 1):.4f} "
              f"outliers {outl} max rel {mx:.2e}")
    return summary


def twin_drive(model, spec, seed, rng, fused_b, topology="private", lead=None, iters=ITERS,
               prepare_b=None, ground_frac=0.05, options=None):
    """Two groups built from the same circuit and driven with the same inputs:
    the first through the four calls, the second fused (fused_b) or also
    through the four calls; their GPU outputs must be equal bit for bit.
    prepare_b(pair) runs on the second group before the inputs are set."""
    pa = Pair(model, spec, topology=topology, seed=seed, ground_frac=ground_frac, lead=lead,
              options=options)
    pb = Pair(model, spec, topology=topology, seed=seed, ground_frac=ground_frac, lead=lead,
              options=options)
    if prepare_b is not None:
        prepare_b(pb)
    cs = rng.uniform(-0.5, 1.5, pa.topo.n_sto)
    ca = rng.uniform(-1e-15, 1e-15, pa.topo.n_sta)
    for p in (pa, pb):
        p.set_inputs(currSto=cs, currSta=ca)
    for flags in iters:
        x = rng.uniform(-0.5, 1.5, pa.topo.n_nodes)
        pa.set_inputs(nextSol=x)
        pb.set_inputs(nextSol=x)
        _, ra = pa.load(fused=False, **flags)
        _, rb = pb.load(fused=fused_b, **flags)
        for k in ra:
            np.testing.assert_array_equal(ra[k].view(np.int64), rb[k].view(np.int64),
                                          err_msg=f"{flags} {k}")
