"""Device-load benchmark: instance loads/s of one load pass
(updateState + updateSecondaryState + loadDAEVectors + loadDAEMatrices) over
all instances of one model type, on one MI355X.

Workload: the device-evaluation microbenchmark -- 2^20 instances of the
model, private nodes, terminal voltages i.i.d. U[-0.5, 1.5] V, W/L lognormal
(sigma 0.2) around 90/45 nm (AREA around 1 for diode/BJT), transient
newtonIter = 1 with voltage limiting.  Inputs are resident in HBM before
timing starts.  The default line also carries, under "mc", the batched Monte
Carlo measurement (4096 ring oscillators through one transient, its speedup
over the plain drop-in and its period error against the oracle; --no-mc
leaves it off), under "hb_solve" the harmonic-balance Newton-Krylov solve of
the 2000-transistor amplifier at 129 samples, and under "logic_transient"
the first 50 ps of the 200k-transistor logic netlist's transient (--no-extra
leaves those two off).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6     # MI355X FP64 vector peak


def cpu_threads():
    """Host cores this process may use, at most 16 (one GPU's share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(model, n_sample, seconds, threads=1):
    """The oracle's load pass on a bounded sample of the workload: updateState
    over `threads` OpenMP threads (the reference's _OMP instance loop), the
    loads serial, in the reference's order."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle
    from xyce_amd.workloads import CARDS, SEED, instance_params
    rng = np.random.default_rng(np.random.SeedSequence(SEED))
    m = oracle.Master(model)
    cards = CARDS[model]
    for level, card in cards:
        m.add_model(card, level=level)
    mi = rng.integers(0, len(cards), size=n_sample).astype(np.int32)
    m.add_instances(n_sample, mi, instance_params(model, n_sample, rng))
    m.setup()
    n_ext, n_int, n_sta, n_sto = m.counts()
    nn = int((n_ext + n_int).sum())
    m.register_lids(np.arange(nn, dtype=np.int32), np.arange(int(n_sta.sum()), dtype=np.int32),
                    np.arange(int(n_sto.sum()), dtype=np.int32))
    r, c = m.jac_pairs()
    key = np.unique(r.astype(np.int64) * nn + c)
    rr = key // nn
    rp = np.zeros(nn + 1, np.int64)
    np.add.at(rp, rr + 1, 1)
    m.register_csr(np.cumsum(rp).astype(np.int32), (key % nn).astype(np.int32))
    m.finalize()
    h = {
        "nextSol": rng.uniform(-0.5, 1.5, nn), "currSol": np.zeros(nn),
        "nextSto": rng.uniform(-0.5, 1.5, max(int(n_sto.sum()), 1)),
        "currSto": rng.uniform(-0.5, 1.5, max(int(n_sto.sum()), 1)),
        "nextSta": np.zeros(max(int(n_sta.sum()), 1)), "currSta": np.zeros(max(int(n_sta.sum()), 1)),
        "lastSta": np.zeros(max(int(n_sta.sum()), 1)), "nextStaDeriv": np.zeros(max(int(n_sta.sum()), 1)),
        "f": np.zeros(nn), "q": np.zeros(nn), "b": np.zeros(nn), "dFdxdVp": np.zeros(nn),
        "dQdxdVp": np.zeros(nn), "dFdx": np.zeros(len(key)), "dQdx": np.zeros(len(key)),
    }
    v = oracle.make_vectors(**h)
    ss = oracle.SolverState()
    ss.transientFlag = 1
    ss.newtonIter = 1
    ss.gainScale = 1.0
    ss.nltermScale = 1.0
    oracle.set_threads(threads)
    reps = 0
    t0 = time.perf_counter()
    while True:
        m.updateState(ss, v)
        m.updateSecondaryState(ss, v)
        m.loadDAEVectors(ss, v)
        m.loadDAEMatrices(ss, v)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    oracle.set_threads(1)
    return n_sample * reps / el, (f"{reps} load passes over {n_sample} instances in {el:.1f} s, "
                                  f"updateState on {threads} thread(s)")


def measured_traffic(model, kernel, n):
    """HBM bytes per launch of `kernel` as the committed rocprofv3 summary of
    this workload measured them (FETCH_SIZE + WRITE_SIZE scaled by the copy
    calibration), rescaled to n instances; None when there is no such row."""
    path = os.path.join(ROOT, "profiles", f"{model}_load_pass_1M.md")
    if not os.path.exists(path):
        return None
    in_table = False
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells and cells[0] == "kernel":
            in_table = len(cells) > 1 and cells[1].startswith("FETCH")
        if in_table and len(cells) == 6 and kernel in cells[0]:
            try:
                return float(cells[5]) * 1e6 * n / (1 << 20)
            except ValueError:
                return None
    return None


def measured_fp64_ops(model, kernel, n):
    """FP64 operations issued per launch of `kernel` (SQ_INSTS_VALU_FLOPS_FP64
    x 64 lanes) as the committed rocprofv3 summary of this workload measured
    them, rescaled to n instances; None when there is no such row."""
    path = os.path.join(ROOT, "profiles", f"{model}_load_pass_1M.md")
    if not os.path.exists(path):
        return None
    in_table = False
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells and cells[0] == "kernel":
            in_table = len(cells) > 1 and cells[1].startswith("FLOPS_FP64")
        if in_table and len(cells) == 4 and kernel in cells[0]:
            try:
                return float(cells[2]) * n / (1 << 20)
            except ValueError:
                return None
    return None


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}))


def ring_transient(args):
    """The plain ring-oscillator workload: one 51-stage BSIM4 ring, DC operating
    point then backward Euler at 1 ps to 2 ns, every Newton iteration one load
    pass on the GPU (the one-call fused load: all 102 instances share one
    block, so the pass is one kernel).  The oracle runs the same transient as the CPU baseline
    and as the reference waveform for the period error."""
    import numpy as np
    from xyce_amd.ring import GpuLoader, Ring, period, transient
    ring = Ring()
    print("[bench] ring transient on the GPU", file=sys.stderr, flush=True)
    loader = GpuLoader(ring, fused=True, graph=not args.no_graph)
    if args.path:
        loader.master.set_path(args.path, args.block)
    t0 = time.perf_counter()
    t, V, st = transient(loader, ring, t_stop=2e-9, h=1e-12)
    wall = time.perf_counter() - t0
    per_gpu = period(t, V[:, 0], ring.vdd / 2)
    out = {
        "metric": "BSIM4 ring-oscillator transient: device load passes/s",
        "value": st["loads"] / st["load_seconds"],
        "unit": "load passes/s",
        "n_gpus": 1,
        "steps": st["steps"],
        "warmup": 0,
        "ms_per_step": wall / st["steps"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "51-stage CMOS ring oscillator, BSIM4 level 54 default cards, "
                               "L=45nm Wn=90nm Wp=180nm, 1fF/stage, VDD=1V, IC node1=0, "
                               "backward Euler 1ps to 2ns", "instances": ring.n_inst},
        "newton_iterations": st["newton_iterations"],
        "path": (f"{args.path} load_all, block {args.block}" if args.path else "load_all policy")
                + ("" if args.no_graph else ", replayed from HIP graphs"),
        "us_per_load": st["load_seconds"] / st["loads"] * 1e6,
        "period_s": per_gpu,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from ring_oracle import OracleLoader
        print("[bench] ring transient on the oracle", file=sys.stderr, flush=True)
        tc, Vc, sc = transient(OracleLoader(ring), ring, t_stop=2e-9, h=1e-12)
        per_cpu = period(tc, Vc[:, 0], ring.vdd / 2)
        out["cpu_baseline"] = {"value": sc["loads"] / sc["load_seconds"], "unit": "load passes/s",
                               "cores": 1, "kind": "port",
                               "sample": f"the same transient, {sc['loads']} load passes"}
        out["period_rel_err_vs_oracle"] = (abs(per_gpu - per_cpu) / per_cpu
                                           if per_gpu and per_cpu else None)
        out["max_wave_err_vs_oracle"] = float(np.max(np.abs(V - Vc)) / ring.vdd)
    print(json.dumps(out))


def mc_measure(args):
    """The batched Monte Carlo workload: n rings (odd stage counts 11..101,
    per-ring |VTH0| and TOXE samples) through one transient on the GPU -- one
    device load of all rings and one device Newton solve per iteration.  A
    seeded sample of the same rings runs one by one through the plain drop-in
    (GPU loads replayed from HIP graphs, host Newton) to price the plain form,
    and through the oracle for the period error and the one-core time.
    Returns the measurement fields."""
    import numpy as np
    import torch
    from xyce_amd.mc import BatchTransient, RingBatch
    from xyce_amd.ring import GpuLoader, period, transient
    batch = RingBatch(n_rings=args.mc_rings)
    h = 1e-12
    print(f"[bench] building {batch.n_rings} rings", file=sys.stderr, flush=True)
    bt = BatchTransient(batch, use_graph=True)
    print("[bench] batched transient", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, V0 = bt.run(t_stop=args.mc_tstop, h=h)
    batched = time.perf_counter() - t0
    batched_persistent = None
    if args.mc_persistent:
        # the same transient with the load's update as one persistent launch
        del bt
        bp = BatchTransient(batch, use_graph=True, persistent=True)
        print("[bench] batched transient, persistent update", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tp, Vp = bp.run(t_stop=args.mc_tstop, h=h)
        batched_persistent = time.perf_counter() - t0
        if not (np.array_equal(tp, t) and np.array_equal(Vp, V0)):
            raise RuntimeError("persistent update changed the waveforms")
        bt = bp
    rng = np.random.default_rng(np.random.SeedSequence(7))
    sample = sorted(rng.choice(batch.n_rings, size=min(args.mc_sample, batch.n_rings),
                               replace=False).tolist())
    plain_times, per_err = [], []
    for r in sample:
        ring = batch.ring(r)
        print(f"[bench] plain drop-in, ring {r} ({ring.N} stages)", file=sys.stderr, flush=True)
        t1 = time.perf_counter()
        transient(GpuLoader(ring, fused=True, graph=True), ring, t_stop=args.mc_tstop, h=h)
        plain_times.append(time.perf_counter() - t1)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from ring_oracle import OracleLoader
        cpu_t = 0.0
        for r in sample:
            ring = batch.ring(r)
            print(f"[bench] oracle, ring {r}", file=sys.stderr, flush=True)
            t1 = time.perf_counter()
            to, Vo, _ = transient(OracleLoader(ring), ring, t_stop=args.mc_tstop, h=h)
            cpu_t += time.perf_counter() - t1
            pb = period(t, V0[:, r], ring.vdd / 2)
            po = period(to, Vo[:, ring.ic_node], ring.vdd / 2)
            if pb is not None and po is not None:
                per_err.append(abs(pb - po) / po)
        cpu = {"value": len(sample) / cpu_t, "unit": "ring transients/s", "cores": 1,
               "kind": "port", "sample": f"{len(sample)} rings {sample} one by one on the oracle"}
    plain_total = float(np.mean(plain_times)) * batch.n_rings
    periods = [period(t, V0[:, r], 0.5) for r in range(batch.n_rings)]
    measured = [p for p in periods if p is not None]
    return {
        "speedup": plain_total / batched,
        "speedup_vs_oracle_1core": (batch.n_rings / cpu["value"]) / batched if cpu else None,
        "steps": int(round(args.mc_tstop / h)),
        "workload": f"{batch.n_rings} CMOS ring oscillators, stages uniform over odd "
                    "11..101, BSIM4 level 54 default cards with per-ring |VTH0| "
                    "N(1,0.03) (DELVTO) and TOXE N(1,0.02), 1fF/stage, VDD=1V, "
                    f"backward Euler 1ps to {args.mc_tstop:g}s",
        "rings": batch.n_rings, "instances": int(2 * batch.stages.sum()),
        "batched_seconds": batched,
        "batched_seconds_persistent_update": batched_persistent,
        "hip_graphs": len(bt.graphs),
        "newton_iterations": bt.newton_iterations,
        "plain_seconds_per_ring": plain_times,
        "plain_rings_sampled": sample,
        "plain_seconds_extrapolated": plain_total,
        "rings_with_period": len(measured),
        "period_rel_err_vs_oracle_max": max(per_err) if per_err else None,
        "cpu_baseline": cpu,
    }


def mc_transient(args):
    """The Monte Carlo measurement as its own bench line (speedup of the
    batched transient over the plain drop-in)."""
    m = mc_measure(args)
    out = {
        "metric": "batched-vs-plain MC transient speedup",
        "value": m["speedup"],
        "unit": "x",
        "n_gpus": 1,
        "steps": m["steps"],
        "warmup": 0,
        "ms_per_step": m["batched_seconds"] / m["steps"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": m["workload"], "rings": m["rings"], "instances": m["instances"]},
    }
    out.update({k: v for k, v in m.items() if k not in ("speedup", "steps", "workload", "rings",
                                                        "instances")})
    print(json.dumps(out))


def hb_measure(args):
    """The harmonic-balance workload: the 2000-transistor differential
    amplifier at N = 129 samples (64 harmonics, N = 2H + 1 as the reference's
    HB takes it): one HB load (inverse transform of the solution, the device
    load of every sample's instances in one launch, forward transforms and
    the j w k Q term), the block-Jacobi preconditioner's 65 FP32 MFMA LU
    factorizations (2408 rows, ld 2432) and one preconditioner solve with
    FP64 refinement of a random right-hand side.  Steps: HB loads."""
    import numpy as np
    import torch
    from xyce_amd import default_solver_state, netlists
    from xyce_amd.hb import HBPlan, HBPrecond
    from xyce_amd.workloads import Workload
    N, T = 129, 1e-9
    cards, mi, params, lids, n_nodes, volts = netlists.hb_diffamp(N, 2000)
    n = n_nodes // N
    print(f"[bench] HB: {N} samples x {n} unknowns, {len(mi)} BSIM4 instances", file=sys.stderr,
          flush=True)
    wl = Workload("bsim4", len(mi), cards=cards, model_index=mi, params=params,
                  netlist=(lids, n_nodes, volts))
    nnz1 = int(wl.row_ptr[n])
    plan = HBPlan(wl.ctx, N, n, T)
    xt = torch.from_numpy(volts).cuda()
    xf = torch.zeros((n, 2 * N), dtype=torch.float64, device="cuda")
    plan.dft(xt, xf)
    outs = [torch.zeros((n, 2 * N), dtype=torch.float64, device="cuda") for _ in range(4)]
    ss = default_solver_state(transientFlag=1, newtonIter=1, pdt=1e12)
    def load():
        plan.load(wl.master, ss, xf, wl.vectors, wl.nnz, *outs)
    for _ in range(args.warmup):
        load()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        load()
    torch.cuda.synchronize()
    t_load = (time.perf_counter() - t0) / args.steps
    # the gate-only nodes (bias, inputs) are held by sources outside this
    # BSIM4-only netlist: a 10 uS shunt on every diagonal stands for them
    rp, ci = wl.row_ptr[:n + 1], wl.col_idx[:nnz1]
    rows = np.repeat(np.arange(n), np.diff(rp))
    diag = torch.from_numpy(np.nonzero(ci == rows)[0]).cuda()
    dfdx = wl.vec["dFdx"][:N * nnz1].clone()
    dfdx.view(N, nnz1)[:, diag] += 1e-5
    dqdx = wl.vec["dQdx"][:N * nnz1].clone()
    pc = HBPrecond(plan, rp, ci)
    pc.factor(dfdx, dqdx)                      # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        info = pc.factor(dfdx, dqdx)
    torch.cuda.synchronize()
    t_factor = (time.perf_counter() - t0) / reps
    bf = torch.from_numpy(np.random.default_rng(5).standard_normal((n, 2 * N))).cuda()
    xs = torch.zeros_like(bf)
    t0 = time.perf_counter()
    r32, rfin, its = pc.solve(bf, xs, max_refine=30, tol=1e-13)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    ld = (2 * n + 63) // 64 * 64
    lu_flops = (N + 1) // 2 * (2.0 / 3.0) * float(ld) ** 3
    return {
        "samples": N, "harmonics": (N - 1) // 2, "unknowns_per_sample": n,
        "instances": len(mi), "hb_load_ms": t_load * 1e3,
        "instance_loads_per_s": len(mi) / t_load,
        "precond_blocks": (N + 1) // 2, "block_rows": 2 * n, "block_ld": ld,
        "factor_ms": t_factor * 1e3, "factor_tflops_fp32": lu_flops / t_factor / 1e12,
        "factor_info_nonzero": int((info != 0).sum()),
        "solve_ms": t_solve * 1e3, "fp32_rel_residual_max": float(np.max(r32)),
        "refined_rel_residual_max": float(np.max(rfin)),
        "refinement_steps_max": int(np.max(its)),
    }


def hb_transient(args):
    """The harmonic-balance measurement as its own bench line."""
    m = hb_measure(args)
    out = {
        "metric": "HB load passes/s (129 samples x 2000-transistor BSIM4 differential amplifier)",
        "value": 1e3 / m["hb_load_ms"], "unit": "HB load passes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["hb_load_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "harmonic balance of a 2000-transistor BSIM4 differential "
                               "amplifier, 64 harmonics (129 time samples), one HB load per step; "
                               "block-Jacobi FP32 MFMA LU + FP64 refinement",
                   "instances": m["instances"]},
        "roofline": {"bound": "mfma", "kernel": "block LU (factor, all kernels)",
                     "achieved": m["factor_tflops_fp32"], "peak": 157.3, "unit": "TFLOP/s",
                     "frac": m["factor_tflops_fp32"] / 157.3, "traffic": None},
        "hb": m, "cpu_baseline": None,
    }
    print(json.dumps(out))


def hb_solve_measure(args, cells_transistors=2000, N=129, T=1e-9, amp=0.01):
    """The harmonic-balance solve of the 2000-transistor differential
    amplifier with its RC filter and sources, driven differentially by a
    10 mV tone at 1 GHz, 64 harmonics (129 samples): the netlist is read,
    every sample's copy of it built in one set of groups, and the HB
    Newton-Krylov solve run to a 1e-9 V update from the generator's bias
    point.  Reports the Newton and Krylov iterations, the time, the final
    residual and the output tone."""
    import numpy as np
    import torch
    import xyce_amd
    from xyce_amd import netlists
    from xyce_amd.hb import HBSolver, sample_steps
    from xyce_amd.netlist_reader import Circuit, Netlist
    t0 = time.perf_counter()
    text, v0 = netlists.diffamp_netlist(cells_transistors)
    drive = lambda t: amp * np.sin(2 * np.pi * t / T)
    nl = Netlist(text + sample_steps({"VINP": lambda t: 0.5 + drive(t),
                                      "VINN": lambda t: 0.5 - drive(t)}, N, T))
    ctx = xyce_amd.Context(0)
    ckt = Circuit(ctx, nl, step=-2, options=xyce_amd.default_options(voltageLimiterFlag=0))
    hb = HBSolver(ckt, T)
    t_build = time.perf_counter() - t0
    print(f"[bench] HB solve: {N} samples x {hb.n} unknowns, {ckt.n_copies} copies built in "
          f"{t_build:.1f} s", file=sys.stderr, flush=True)
    x0 = np.zeros(hb.n)
    for k, name in enumerate(nl.node_names):
        x0[k] = v0[int(name[1:])]
    xt0 = torch.from_numpy(np.tile(x0, N)).cuda()
    xf0 = torch.zeros((hb.n, 2 * N), dtype=torch.float64, device="cuda")
    hb.plan.dft(xt0, xf0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xf, st = hb.solve(xf0, vtol=1e-9, krylov_tol=1e-10, verbose=True,
                      refactor_every=getattr(args, "hb_refactor", 2))
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    out = [nl.node(f"n{len(v0) - 2}"), nl.node(f"n{len(v0) - 1}")]
    x = xf.cpu().numpy()
    tone = [float(np.hypot(x[j, 2], x[j, 3]) * 2) for j in out]
    return {"samples": N, "harmonics": (N - 1) // 2, "unknowns_per_sample": hb.n,
            "unknowns": hb.n * N, "instances_bsim4": cells_transistors * N,
            "build_s": t_build, "solve_s": t_solve, "newton": st["newton"],
            "krylov_iterations": [h["krylov_iterations"] for h in st["history"]],
            "loads": st["loads"], "matvecs": st["matvecs"],
            "final_residual_max": st["final_residual_max"],
            "output_tone_amplitude_v": tone}


def logic_transient_measure(args, t_stop):
    """The flat logic netlist's FP64 transient (200k BSIM4 transistors in a
    random INV/NAND2/NOR2 DAG, wire RC, 20 PULSE inputs at 1 GHz, 1 ps
    steps): read from its netlist text, DC operating point then fixed steps
    (backward Euler twice, then trapezoid), every Newton iteration's loads
    and its GMRES + ILU(0) solve on the GPU."""
    import numpy as np
    import torch
    import xyce_amd
    from xyce_amd import netlists
    from xyce_amd.netlist_reader import Circuit, Netlist
    from xyce_amd.sparse_transient import SparseTransient
    t0 = time.perf_counter()
    text, volts = netlists.logic_netlist(args.n_logic, t_stop=t_stop)
    nl = Netlist(text)
    ctx = xyce_amd.Context(0)
    tr = SparseTransient(Circuit(ctx, nl), order=args.order, verbose=args.verbose)
    t_build = time.perf_counter() - t0
    print(f"[bench] logic transient: {args.n_logic} BSIM4, {tr.n} unknowns, circuit nnz "
          f"{tr.c.nnz}, built in {t_build:.1f} s", file=sys.stderr, flush=True)
    x0 = np.zeros(tr.n)
    for k, name in enumerate(nl.node_names):
        x0[k] = volts[int(name[1:])]
    rec = np.arange(1, 21)            # the 20 inputs' nodes
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, V, nw = tr.run(torch.from_numpy(x0).cuda(), t_stop, 1e-12, record=rec, progress=50)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = tr.stats
    return {"transistors": args.n_logic, "unknowns": tr.n, "jacobian_nnz": tr.J.nnz,
            "t_stop": t_stop, "steps": len(t) - 1, "wall_s": wall, "build_s": t_build,
            "steps_per_s": (len(t) - 1) / wall, "order": args.order,
            "order_info": getattr(tr, "order_info", None),
            "newton_total": int(nw.sum()), "newton_dcop": int(nw[0]),
            "newton_per_step_mean": float(nw[1:].mean()), "gmres_total": st["krylov"],
            "loads": st["loads"], "load_s": st["load_s"], "factor_s": st["factor_s"],
            "krylov_s": st["krylov_s"], "ilu_factorisations": st.get("factors")}


def logic_transient_line(args):
    m = logic_transient_measure(args, args.tran_stop)
    out = {
        "metric": "flat logic netlist FP64 transient: time steps/s (200k BSIM4, 1 ps steps)",
        "value": m["steps_per_s"], "unit": "time steps/s", "n_gpus": 1, "steps": m["steps"],
        "warmup": 0, "ms_per_step": m["wall_s"] / m["steps"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"flat logic netlist, {args.n_logic} BSIM4 transistors, wire RC, "
                               f"20 PULSE inputs at 1 GHz, .TRAN 1ps {args.tran_stop:g}; DC op "
                               "then BE x2 + trapezoid, GMRES + ILU(0) on the device",
                   "unknowns": m["unknowns"], "jacobian_nnz": m["jacobian_nnz"],
                   "order": m["order"], "order_info": m["order_info"]},
        "roofline": None, "cpu_baseline": None, "transient": m,
    }
    print(json.dumps(out))


def hb_solve_line(args):
    m = hb_solve_measure(args)
    out = {
        "metric": "HB solves/s (129 samples x 2000-transistor BSIM4 differential amplifier + RC "
                  "filter, Newton-Krylov to 1e-9 V)",
        "value": 1.0 / m["solve_s"], "unit": "HB solves/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": m["solve_s"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "harmonic balance of a 2000-transistor BSIM4 differential "
                               "amplifier with RC filter and sources, 1 GHz 10 mV differential "
                               "drive, 64 harmonics (129 samples): the full Newton-Krylov solve"},
        "roofline": None, "hb_solve": m, "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--model", default="bsim4", choices=["diode", "mosfet1", "bjt", "bsim4", "resistor", "capacitor"])
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--cpu-sample", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fused", action="store_true", help="time the one-call fused load")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the block shapes on the workload first, then time xa_load_all")
    ap.add_argument("--path", choices=["split", "fused", "pattern"],
                    help="pin xa_load_all's path (with --block) and time it")
    ap.add_argument("--block", type=int, default=128, help="block size for --path")
    ap.add_argument("--ring", action="store_true",
                    help="run the 51-stage BSIM4 ring-oscillator transient (1 ps steps to 2 ns) "
                         "instead of the microbenchmark")
    ap.add_argument("--no-graph", action="store_true",
                    help="--ring: launch every load's kernels instead of replaying HIP graphs")
    ap.add_argument("--netlist", choices=["logic", "hb"],
                    help="load passes over a shared-node netlist instead of private nodes: "
                         "'logic' = 200k-transistor random INV/NAND2/NOR2 DAG, 'hb' = the 128 "
                         "time samples of a 2k-transistor differential amplifier")
    ap.add_argument("--mc", action="store_true",
                    help="run the batched Monte Carlo ring-oscillator transient (4096 rings) and "
                         "its speedup over the plain drop-in")
    ap.add_argument("--no-extra", action="store_true",
                    help="leave the HB solve and the 50 ps logic transient off the default line")
    ap.add_argument("--no-mc", action="store_true",
                    help="leave the Monte Carlo measurement off the default line")
    ap.add_argument("--logic-transient", action="store_true",
                    help="the flat logic netlist's FP64 transient (read from its netlist, sparse "
                         "Newton with GMRES + ILU(0) on the device)")
    ap.add_argument("--n-logic", type=int, default=200_000,
                    help="--logic-transient: transistor count")
    ap.add_argument("--tran-stop", type=float, default=1e-9,
                    help="--logic-transient: end time (s)")
    ap.add_argument("--order", default="topo", choices=["topo", "rcm", "natural"],
                    help="--logic-transient: unknown order of the sparse solve")
    ap.add_argument("--verbose", action="store_true", help="print every Newton iteration")
    ap.add_argument("--hb-refactor", type=int, default=2,
                    help="--hb-solve: refactor the preconditioner every N Newton iterations")
    ap.add_argument("--hb-solve", action="store_true",
                    help="the full harmonic-balance Newton-Krylov solve of the 2000-transistor "
                         "differential amplifier with its filter and sources, 129 samples")
    ap.add_argument("--hb", action="store_true",
                    help="harmonic balance of the 2000-transistor differential amplifier at "
                         "129 samples: HB loads, the FP32 MFMA block LU and its refined solve")
    ap.add_argument("--mc-rings", type=int, default=4096)
    ap.add_argument("--mc-tstop", type=float, default=6e-9)
    ap.add_argument("--mc-persistent", action="store_true",
                    help="also run the batched transient with the persistent update kernel")
    ap.add_argument("--mc-sample", type=int, default=4,
                    help="rings run through the plain drop-in and the oracle")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.ring:
        ring_transient(args)
        return
    if args.mc:
        mc_transient(args)
        return
    if args.hb_solve:
        hb_solve_line(args)
        return
    if args.logic_transient:
        logic_transient_line(args)
        return
    if args.hb:
        hb_transient(args)
        return

    import torch
    from xyce_amd.workloads import Workload

    netlist_desc = None
    if args.netlist:
        from xyce_amd import netlists
        gen = netlists.logic_dag if args.netlist == "logic" else netlists.hb_diffamp
        cards, mi, params, lids, n_nodes, volts = gen()
        args.model, args.n = "bsim4", len(mi)
        print(f"[bench] building {args.netlist} netlist, {args.n} BSIM4 instances, "
              f"{n_nodes} nodes", file=sys.stderr, flush=True)
        wl = Workload("bsim4", args.n, cards=cards, model_index=mi, params=params,
                      netlist=(lids, n_nodes, volts))
        netlist_desc = {
            "logic": f"flat logic netlist: {args.n} BSIM4 transistors in a seeded random DAG "
                     "of INV/NAND2/NOR2 cells, fanout 1-4, 20 primary inputs, two shared cards "
                     "and sizes, one VDD rail; node voltages U[0,1]; one load pass per step",
            "hb": f"harmonic-balance samples: a {args.n // 128}-transistor BSIM4 differential "
                  "amplifier x 128 time samples of a 1.00/1.01 GHz 10 mV two-tone drive, all "
                  "samples' instances loaded together; one load pass per step"}[args.netlist]
    else:
        print(f"[bench] building {args.model} workload n={args.n}", file=sys.stderr, flush=True)
        wl = Workload(args.model, args.n)
    tuned = None
    if args.path:
        wl.master.set_path(args.path, args.block)
        tuned = {"route": args.path, "block": args.block, "pinned": True}
        args.fused = True
    elif args.autotune:
        route, block = wl.master.autotune(wl.ss, wl.vectors, reps=3)
        wl.zero_outputs()
        tuned = {"route": route, "block": block}
        args.fused = True
        print(f"[bench] autotune: {route}, block {block}", file=sys.stderr, flush=True)
    step = wl.load_fused if args.fused else wl.load_pass
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("[bench] timing", file=sys.stderr, flush=True)
    wl.master.enable_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_up, ms_vec, ms_mat = wl.master.last_times()
    wl.master.enable_timing(False)
    ms_step = (t1 - t0) / args.steps * 1e3
    value = args.n / (ms_step * 1e-3)
    bu, bv, bm = wl.master.traffic_bytes()
    # xa_load_all runs one fused kernel (all phases timed as update) or, where
    # the model is faster split at this size, the four calls
    one_kernel = args.fused and ms_vec == 0.0 and ms_mat == 0.0
    if one_kernel:
        dom_bytes, dom_ms, dom_name = bu + bv + bm, ms_up, "fused load"
    else:
        dom_bytes, dom_ms, dom_name = bu, ms_up, "update_state"
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0

    # both bounds of the dominant kernel (the survey's measurement rule: report
    # HBM bytes/s and FP64 operations/s against their peaks, headline the
    # binding one)
    kname = f"{args.model}::k_update<false"
    hbm = {"bound": "hbm", "kernel": dom_name, "achieved": achieved, "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "traffic": None if (one_kernel or args.netlist) else
           measured_traffic(args.model, kname, args.n),
           "algorithmic_bytes": dom_bytes}
    ops = None if (one_kernel or args.netlist) else measured_fp64_ops(args.model, kname, args.n)
    fp64 = None
    if ops and dom_ms > 0:
        tf = ops / (dom_ms * 1e-3) / 1e12
        fp64 = {"bound": "fp64", "kernel": dom_name, "achieved": tf, "peak": FP64_PEAK_TFS,
                "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFS, "fp64_ops": ops,
                "traffic": hbm["traffic"]}
    roofline = fp64 if fp64 and fp64["frac"] > hbm["frac"] else hbm

    cpu = None
    if not args.no_cpu and not args.netlist:
        n_sample = args.cpu_sample or {"diode": 1 << 17, "mosfet1": 1 << 15, "bjt": 1 << 15,
                                       "bsim4": 1 << 13, "resistor": 1 << 18,
                                       "capacitor": 1 << 18}[args.model]
        print(f"[bench] cpu baseline on {n_sample} instances", file=sys.stderr, flush=True)
        threads = cpu_threads()
        c1, d1 = cpu_baseline(args.model, n_sample, args.cpu_seconds / 2)
        cpu = {"value": c1, "unit": "instance loads/s", "cores": 1, "kind": "port", "sample": d1}
        if threads > 1:
            cn, dn = cpu_baseline(args.model, n_sample, args.cpu_seconds / 2, threads)
            cpu = {"value": cn, "unit": "instance loads/s", "cores": threads, "kind": "port",
                   "sample": dn, "single_core_value": c1, "single_core_sample": d1}
    out = {
        "metric": f"{args.model.upper()} instance loads/s (eval+stamp)",
        "value": value,
        "unit": "instance loads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": netlist_desc or
                   f"device-eval microbenchmark, {args.n} {args.model} instances, "
                   "private nodes, V~U[-0.5,1.5], one load pass per step",
                   "instances": args.n, "device": args.model,
                   "path": (f"{tuned['route']} load_all" if tuned and one_kernel else
                            "fused load_all") if one_kernel else
                           ("load_all (split at this size)" if args.fused else
                            "updateState+loadDAEVectors+loadDAEMatrices")},
        "phase_ms": {"update_state": ms_up, "load_vectors": ms_vec, "load_matrices": ms_mat},
        "autotune": tuned,
        "roofline": roofline,
        "roofline_hbm": hbm,
        "roofline_fp64": fp64,
        "cpu_baseline": cpu,
    }
    if not args.no_mc and not args.netlist and args.model == "bsim4" and args.n == 1 << 20:
        # the Monte Carlo half of the metric rides on the default line
        del wl, step
        torch.cuda.empty_cache()
        out["mc"] = mc_measure(args)
    if not args.no_extra and not args.netlist and args.model == "bsim4" and args.n == 1 << 20:
        # the other two workloads, short forms: the HB solve of the amplifier
        # and the first 50 ps of the logic netlist's transient
        torch.cuda.empty_cache()
        out["hb_solve"] = hb_solve_measure(args)
        torch.cuda.empty_cache()
        out["logic_transient"] = logic_transient_measure(args, 5e-11)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
