"""Benchmark: complete flow right-hand sides per second on one MI355X.

Workload (single RHS-evaluation bench): square lattice L=6 (16 irreducible sites),
nearest-neighbour Heisenberg J1=1, SU(2), two-loop RHS = compute_dΣ! + compute_dΓ_2l!
at Λ=1 with eval=10 and (atol, rtol) = (1e-8, 1e-4), meshes of the launch! recipe
with num_σ=num_Ω=num_ν=40 (41 points), and a seeded synthetic vertex
(N(0,1) amplitudes × 1/(1+|Ω|+|ν|+|ν'|), splitmix64 seed 20261203, 2 SU(2) components).

One step = one complete RHS on device-resident inputs.  The headline value is the
FP32-mixed path; the plain FP64 GPU path (the metric's denominator) is timed in the
same run, and the relative max-norm error of FP32-mixed against FP64 is reported.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "pffrgsolver.jl_amd"))

import numpy as np  # noqa: E402

from pffrg import action as A  # noqa: E402
from pffrg import flow as F  # noqa: E402
from pffrg.native import FP32_ELEMENT, FP32_MIXED, FP64  # noqa: E402

METRIC = "RHS evals/s + full-flow wall time vs plain FP64 GPU, at χ rel err ≤1e-4"
HBM_PEAK_GBS = 8000.0
SEED = 20261203


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_problem(L: int, N: int):
    return F.Problem.build("square", L, [[1.0]], initial=1.0, num_sigma=N, num_Omega=N, num_nu=N, num_chi=10)


# the autotuner's candidate list (pffrg_tune_shared): (waves per block, batch capacity in nodes of
# 32 site vectors, most nodes per batch)
TUNE_CANDIDATES = [(8, 8, 64), (8, 8, 32), (8, 8, 16), (8, 4, 64), (8, 4, 32),
                   (4, 8, 64), (4, 8, 32), (4, 4, 64), (4, 4, 32), (4, 4, 16)]


def time_mode(prob, a, mode: int, loops: int, steps: int, warmup: int, eval_: int, tune: dict | None = None):
    ns, m = prob.r.n_sites, prob.mesh
    ctx = prob.context(n_slots=3, precision=mode)
    ctx.upload(0, a)
    zero = A.Action.empty(ns, m)
    ctx.upload(1, zero)
    ctx.upload(2, zero)
    if tune is not None and tune.get("fixed"):
        ctx.set_shared_shape(*tune["fixed"])
        tune["shape"] = dict(zip(("waves", "batch_nodes", "max_nodes"), tune["fixed"]))
    elif tune is not None:
        # the shared-node pass's launch shape, picked by the autotuner on this problem's own
        # Katanin pass (slot 1 is an output and is overwritten below)
        shape, ms = ctx.tune_shared(1.0, 0, 1, eval_)
        tune.update({"shape": {"waves": shape[0], "batch_nodes": shape[1], "max_nodes": shape[2]},
                     "candidates_ms": {str(c): (t if t >= 0 else None) for c, t in zip(TUNE_CANDIDATES, ms)}})
        log(f"  autotuner: shape {shape}, Katanin pass ms {['%.1f' % t for t in ms]}")
    warm_ms = []
    for _ in range(warmup):
        tw = time.perf_counter()
        F.rhs(ctx, 1.0, loops, 0, 1, eval_, da_l=2)
        ctx.synchronize()
        warm_ms.append((time.perf_counter() - tw) * 1e3)
    ctx.synchronize()
    ctx.reset_stats()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for k in range(steps):
        F.rhs(ctx, 1.0, loops, 0, 1, eval_, da_l=2)
        log(f"  step {k + 1}/{steps} done")
    ctx.synchronize()
    dt = time.perf_counter() - t0
    st = ctx.stats()
    kms, klaunch = ctx.timing()
    out = ctx.download(1)
    ctx.close()
    return dt, st, kms, klaunch, out, warm_ms


L2_FOOTPRINT = 2 << 20  # bytes: half of one XCD's 4 MiB L2


def source_hash() -> str:
    """sha256 over the HIP sources (the library the measurement runs): a committed counter
    profile applies to the current kernels only if it was taken from the same sources."""
    import hashlib
    h = hashlib.sha256()
    for f in sorted((ROOT / "pffrgsolver.jl_amd" / "csrc").glob("*")):
        if f.suffix in (".h", ".hip", ".cpp"):
            h.update(f.name.encode())
            h.update(f.read_bytes())
    return h.hexdigest()[:16]


def committed_traffic(name: str):
    """HBM bytes per launch of the named kernel from the committed rocprofv3 counter profile,
    with its source, or (None, why) when the profile is missing or was taken from other sources."""
    tfile = ROOT / "profiles" / name
    if not tfile.exists():
        return None, f"no committed profile {name}"
    d = json.loads(tfile.read_text())
    if d.get("source_hash") != source_hash():
        return None, f"profiles/{name} was taken from other kernel sources ({d.get('source_hash')})"
    return d.get("hbm_bytes_per_launch"), f"profiles/{name} ({d.get('source', '')}, same sources)"


def per_pass_stats(prob, a, mode: int, eval_: int, shape=None) -> dict:
    """Nodes per element, gathered bytes per element-node and kernel ms of the Katanin (half) and
    left (full) passes of the 2-loop RHS: compute_dΓ_1l! alone, then compute_dΓ_2l!, whose
    difference is the left pass."""
    ns, m = prob.r.n_sites, prob.mesh
    ctx = prob.context(n_slots=4, precision=mode)
    if shape is not None:
        ctx.set_shared_shape(*shape)
    ctx.upload(0, a)
    zero = A.Action.empty(ns, m)
    for s in (1, 2, 3):
        ctx.upload(s, zero)
    # one untimed RHS first, so that both passes dispatch their units in measured order
    ctx.compute_dsigma(1.0, 0, 1)
    ctx.compute_dgamma_2l(1.0, 0, 1, 2, eval_)
    ctx.synchronize()
    ctx.set_timing(True)
    out = {}
    runs = []
    for loops in (1, 2):
        ctx.compute_dsigma(1.0, 0, 1)
        ctx.synchronize()
        ctx.reset_stats()
        if loops == 1:
            ctx.compute_dgamma_1l(1.0, 0, 1, eval_)
        else:
            ctx.compute_dgamma_2l(1.0, 0, 1, 2, eval_)
        ctx.synchronize()
        st = ctx.stats()
        ms, nl = ctx.timing()
        runs.append((st, ms))
    (s1, ms1), (s2, ms2) = runs
    left = {k: s2[k] - s1[k] for k in s1}
    for name, st, ms in (("katanin", s1, ms1), ("left", left, ms2 - ms1)):
        out[name] = {"elements": st["elements"], "element_nodes": st["nodes"],
                     "nodes_per_element": st["nodes"] / max(1, st["elements"]),
                     "gathered_bytes_per_element_node": st["gathered_bytes"] / max(1, st["nodes"]),
                     "kernel_ms": ms, "ns_per_element_node": ms * 1e6 / max(1, st["nodes"])}
    ctx.close()
    return out


def gather_ceilings(prob, a) -> dict:
    """Practical ceilings of the vertex pass's gathers on this GPU: random trilinear q3 lookups
    (8 corners x 3 channels, the pass's lane layout) on the same vertex's gather copies, with no
    other work (pffrg_gather_ceiling), capped at the vertex pass's 3 waves per SIMD and at the
    occupancy the hardware allows.  "mall": lookups over the whole q3 block (mostly served by
    the Infinity Cache); "l2": lookups confined to 2 MiB, which stays in each XCD's L2, as the
    vertex pass's gathers mostly do.  GB/s of gathered bytes, counted as the pass counts them."""
    ctx = prob.context(n_slots=1, precision=FP32_MIXED)
    ctx.upload(0, a)
    out = {}
    for name, mode in (("fp32", FP32_MIXED), ("fp64", FP64)):
        for where, fp in (("mall", 0), ("l2", L2_FOOTPRINT)):
            for w in (3, 0):
                gbps, ms = ctx.gather_ceiling(0, mode, w, fp, n_iter=256, n_rep=5)
                out[f"{name}_{where}_{'max_occupancy' if w == 0 else f'{w}_waves'}_GBps"] = gbps
                log(f"  gather ceiling {name} {where} waves/SIMD {'max' if w == 0 else w}: {gbps:.0f} GB/s "
                    f"({ms:.2f} ms/launch)")
    ctx.close()
    return out


def cpu_baseline(prob, a, da_gpu, loops: int, eval_: int, budget_s: float):
    """The FP64 oracle (C++/OpenMP restatement) on a strided sample of the same RHS's
    elements; extrapolated to one full RHS.  Returns (RHS/s, threads, sample text)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O  # checker-only import, cpu_baseline leg

    ns, m = prob.r.n_sites, prob.mesh
    orc = O.Oracle(prob.r, m)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    orc.set_threads(threads)
    ref = A.Action.empty(ns, m)
    t0 = time.perf_counter()
    orc.compute_dsigma(1.0, a, ref)
    t_sigma = time.perf_counter() - t0
    ref.Sigma[:] = da_gpu.Sigma
    corrs_kat = orc.compute_corrs(1.0, a, ref, True)
    # element lists of the two vertex passes, sampled with a stride
    nO, nn = m.num_Omega, m.num_nu
    half = [(ch, k, w1, w2, w3) for ch in range(3) for w1 in range(nO)
            for (k, w2, w3) in ([(1, -1, -1)] + [(2, w, -1) for w in range(nn)] + [(3, -1, w) for w in range(nn)]
                                + [(4, w2, w3) for w3 in range(nn) for w2 in range(w3, nn)])]
    full = [(ch, k, w1, w2, w3) for ch in range(3) for w1 in range(nO)
            for (k, w2, w3) in ([(1, -1, -1)] + [(2, w, -1) for w in range(nn)] + [(3, -1, w) for w in range(nn)]
                                + [(4, w2, w3) for w3 in range(nn) for w2 in range(nn)])]
    passes = [(O.P_KAT, half, a, ref, corrs_kat)]
    if loops >= 2:
        corrs = orc.compute_corrs(1.0, a, a, False)
        passes.append((O.P_LEFT, full, a, da_gpu, corrs))
    scratch = A.Action.empty(ns, m)
    stride = 4096
    total_est = t_sigma
    sampled = 0
    t_start = time.perf_counter()
    while True:
        est = t_sigma
        n_done = 0
        for (ps, elems, aa, dd, cc) in passes:
            orc.corrs[:] = cc
            idx = list(range(0, len(elems), stride))
            t1 = time.perf_counter()
            _sample_elements(orc, ps, aa, dd, elems, idx, scratch, eval_)
            el = time.perf_counter() - t1
            est += el * len(elems) / max(1, len(idx))
            n_done += len(idx)
        total_est, sampled = est, n_done
        spent = time.perf_counter() - t_start
        if spent > budget_s / 4 or stride <= 1:
            break
        stride = max(1, stride // 4)
    sample = (f"1 in {stride} elements of both vertex passes ({sampled} elements) + full dSigma, "
              f"FP64 oracle, {threads} OpenMP threads, extrapolated to one RHS ({total_est:.1f} s)")
    # the same on one thread, over a sparser sample: the per-core rate
    orc.set_threads(1)
    t0 = time.perf_counter()
    orc.compute_dsigma(1.0, a, ref)
    est1 = time.perf_counter() - t0
    stride1 = max(256, stride * 16)
    for (ps, elems, aa, dd, cc) in passes:
        orc.corrs[:] = cc
        idx = list(range(0, len(elems), stride1))
        t1 = time.perf_counter()
        _sample_elements(orc, ps, aa, dd, elems, idx, scratch, eval_, workers=1)
        est1 += (time.perf_counter() - t1) * len(elems) / max(1, len(idx))
    orc.set_threads(threads)
    extra = {"single_thread_value": 1.0 / est1,
             "single_thread_sample": f"1 in {stride1} elements of both vertex passes + full dSigma on 1 thread, "
                                     f"extrapolated to one RHS ({est1:.1f} s)",
             **host_cpu_info()}
    return 1.0 / total_est, threads, sample, extra


def host_cpu_info() -> dict:
    """The host CPU as lscpu names it, the CPUs this process may use and the thread settings.
    The reference's own Julia solver is not run (no Julia on the machine), so its thread count
    is recorded as not applicable."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if "Model name" not in info:
        try:
            info["Model name"] = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                                      if l.startswith("model name"))
        except (OSError, StopIteration):
            info["Model name"] = "unknown"
    return {"cpu_model": info.get("Model name"), "lscpu": info, "host_cpus": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0)), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "JULIA_NUM_THREADS": "n/a (reference Julia not run: no Julia on this machine)"}


def _sample_elements(orc, ps, aa, dd, elems, idx, out, eval_, workers=None):
    from concurrent.futures import ThreadPoolExecutor
    # the oracle parallelises a pass internally; for a sample run the elements as separate calls
    # on a thread pool (ctypes releases the GIL)
    n = workers or len(os.sched_getaffinity(0))
    with ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(lambda k: orc.compute_element(ps, 1.0, aa, dd, *elems[k], eval_, out), idx))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=6)
    ap.add_argument("--N", type=int, default=40)
    ap.add_argument("--loops", type=int, default=2)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ws-N", type=int, default=16, help="warmstart: mesh size")
    ap.add_argument("--ws-train", type=int, default=64, help="warmstart: training solutions")
    ap.add_argument("--ws-test", type=int, default=16, help="warmstart: held-out couplings")
    ap.add_argument("--ws-steps", type=int, default=20000, help="warmstart: Adam steps")
    ap.add_argument("--ws-verbose", action="store_true", help="warmstart: print every parquet iteration")
    ap.add_argument("--flow-config", choices=["square", "pyrochlore"], default="square",
                    help="flow: square L=3 1-loop N=24 50->0.3, or pyrochlore L=4 3-loop with self-energy "
                         "corrections N=32 50->0.1")
    ap.add_argument("--flow-modes", default="fp64,fp32", help="flow: precisions to run, in order")
    ap.add_argument("--flow-N", type=int, default=0, help="flow: mesh size instead of the configuration's")
    ap.add_argument("--flow-max-seconds", type=float, default=None, help="flow: wall-time budget per flow")
    ap.add_argument("--flow-verbose", action="store_true", help="flow: print every BS3 step")
    ap.add_argument("--batch-N", type=int, default=32, help="batch / batchflow: mesh size of every problem")
    ap.add_argument("--batch-skip-single", action="store_true",
                    help="batchflow: run the batched flows only (no one-at-a-time comparison)")
    ap.add_argument("--workload", choices=["single", "batch", "batchflow", "flow", "warmstart", "mlp"], default="single",
                    help="single: the 2-loop square L=6 RHS (headline); batch: the J1-J2 sweep + uneven "
                         "truncations as one batched 1-loop RHS")
    ap.add_argument("--shape", default="", metavar="W,K,M",
                    help="single: the shared-node pass's launch shape (waves per block, batch capacity, most nodes "
                         "per batch) instead of the autotuner's choice")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (3 times), the known-size reference "
                         "for scaling hardware byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is a design note, "
              f"not part of this build", file=sys.stderr)
        sys.exit(2)
    eval_ = 10
    if args.workload == "batch":
        return batch_workload(args)
    if args.workload == "batchflow":
        return batchflow_workload(args)
    if args.workload == "flow":
        return flow_workload(args)
    if args.workload == "warmstart":
        return warmstart_workload(args)
    if args.workload == "mlp":
        return mlp_workload(args)
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 x {args.copy_calibration} MiB device copies")
        del src, dst
    t0 = time.time()
    prob = build_problem(args.L, args.N)
    ns, m = prob.r.n_sites, prob.mesh
    a = A.synthetic_action(ns, m, prob.bare, seed=SEED)
    log(f"workload: square L={args.L} ns={ns} N={args.N} ({m.num_Omega} points) loops={args.loops} "
        f"(setup {time.time() - t0:.1f}s)")

    log("FP32-mixed leg (shared-node pass)")
    tune = {"fixed": tuple(int(x) for x in args.shape.split(","))} if args.shape else {}
    dt32, st32, kms32, kl32, out32, warm32 = time_mode(prob, a, FP32_MIXED, args.loops, args.steps, args.warmup,
                                                       eval_, tune=tune)
    log(f"  {dt32 / args.steps * 1e3:.1f} ms/RHS, element-nodes/RHS {st32['nodes'] / args.steps:.3e}")
    res = {}
    if not args.no_fp64:
        log("FP32 per-element leg (the same arithmetic without shared nodes)")
        dte, ste, kmse, kle, oute, _ = time_mode(prob, a, FP32_ELEMENT, args.loops, args.steps, args.warmup, eval_)
        log(f"  {dte / args.steps * 1e3:.1f} ms/RHS")
        res["fp32_element_ms_per_step"] = dte / args.steps * 1e3
        res["speedup_shared_vs_element"] = dte / dt32
    if not args.no_fp64:
        log("FP64 leg")
        dt64, st64, kms64, kl64, out64, _ = time_mode(prob, a, FP64, args.loops, args.steps, args.warmup, eval_)
        log(f"  {dt64 / args.steps * 1e3:.1f} ms/RHS")
        err_g = float(np.max(np.abs(out32.gamma - out64.gamma)) / np.max(np.abs(out64.gamma)))
        err_q3 = max(float(np.max(np.abs(out32.view(c, ch, "q3") - out64.view(c, ch, "q3"))) /
                           np.max(np.abs(out64.view(c, ch, "q3")))) for c in range(2) for ch in range(3))
        err_s = float(np.max(np.abs(out32.Sigma - out64.Sigma)) / np.max(np.abs(out64.Sigma)))
        res.update({"fp64_rhs_per_s": args.steps / dt64, "fp64_ms_per_step": dt64 / args.steps * 1e3,
                    "speedup_vs_fp64": dt64 / dt32,
                    "element_nodes_ratio_fp64_vs_shared": st64["nodes"] / max(1, st32["nodes"]),
                    "fp32_vs_fp64_rel_err": {"vertex": err_g, "q3_worst": err_q3, "sigma": err_s},
                    "fp64_kernel": {"kernel": "k_vertex_pass<PASS, double>", "ms_per_launch": kms64 / max(1, kl64),
                                    "gathered_GBps": st64["gathered_bytes"] / (kms64 * 1e-3) / 1e9 if kms64 else None}})
        log("per-pass statistics")
        tshape = tuple(tune["shape"][k] for k in ("waves", "batch_nodes", "max_nodes")) if tune else None
        res["per_pass"] = {"fp32_mixed_shared": per_pass_stats(prob, a, FP32_MIXED, eval_, tshape),
                           "fp64": per_pass_stats(prob, a, FP64, eval_)}
        ref_out = out64
    else:
        ref_out = out32
    # dominant kernel: the shared-node pass k_shared_pass (the timing covers it alone);
    # algorithmic bytes = the interpolation corners it gathers (2 components x 4 B per corner)
    bytes_per_launch = st32["gathered_bytes"] / max(1, kl32)
    avg_ms = kms32 / max(1, kl32)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic, traffic_note = committed_traffic("shared_pass_traffic.json")
    log("gather ceiling")
    ceil = gather_ceilings(prob, a)
    if "fp64_kernel" in res and res["fp64_kernel"]["gathered_GBps"]:
        res["fp64_kernel"]["frac_of_l2_gather_ceiling_3_waves"] = \
            res["fp64_kernel"]["gathered_GBps"] / ceil["fp64_l2_3_waves_GBps"]
    cpu = None
    if not args.no_cpu:
        log("CPU baseline leg (FP64 oracle, bounded sample)")
        v, cores, sample, extra = cpu_baseline(prob, a, ref_out, args.loops, eval_, args.cpu_budget)
        cpu = {"value": v, "unit": "RHS evals/s", "cores": cores, "kind": "port", "sample": sample, **extra}
        log(f"  cpu {v:.4g} RHS/s on {cores} threads")
    line = {
        "metric": METRIC,
        "value": args.steps / dt32,
        "unit": "RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (fp32 interpolation/products, fp64 quadrature)",
        "data": "synthetic (seeded N(0,1) vertex with 1/(1+|W|+|v|+|v'|) decay)",
        "config": {"workload": f"single RHS: square L={args.L} ({ns} sites) Heisenberg SU(2) {args.loops}-loop, "
                               f"N={args.N} (+1 points/axis), Lambda=1, eval={eval_}",
                   "lattice": "square", "L": args.L, "n_sites": ns, "N": args.N, "loops": args.loops},
        "element_nodes_per_rhs": st32["nodes"] / args.steps,
        "first_rhs_ms": warm32[0] if warm32 else None,
        "first_rhs_note": "the context's first RHS (untimed warmup): its shared-node passes dispatch their units "
                          "in the prior order (the autotuner's runs have already ordered the Katanin pass when it "
                          "ran); every later RHS orders them by the run times of the previous one",
        "autotune": tune,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "kernel": "k_shared_pass", "bytes_per_launch": bytes_per_launch, "ms_per_launch": avg_ms,
                     "launches_per_step": kl32 / max(1, args.steps), "source_hash": source_hash(),
                     "gather_ceiling": ceil,
                     "frac_of_l2_gather_ceiling_max_occupancy": achieved / ceil["fp32_l2_max_occupancy_GBps"],
                     "note": "gathered bytes are served from L2 and the Infinity Cache; the pass is bound by "
                             "gather latency and LDS traffic, not HBM"},
        "cpu_baseline": cpu,
    }
    line.update(res)
    print(json.dumps(line), flush=True)


FLOW_CONFIGS = {
    # launch!(f, "square", 3, "heisenberg", "su2", [1.0]; loops=1, parquet=false, initial=50.0, final=0.3,
    #         num_σ=num_Ω=num_ν=24, num_χ=10)
    "square": dict(name="square", size=3, N=24, loops=1, Li=50.0, Lf=0.3,
                   workload="1-loop flow square L=3 Heisenberg SU(2), N=24 (+1 points/axis), Lambda 50 -> 0.3"),
    # launch!(f, "pyrochlore", 4, "heisenberg", "su2", [1.0]; loops=3, Σ_corr=true, parquet=false,
    #         initial=50.0, final=0.1, num_σ=num_Ω=num_ν=32)
    "pyrochlore": dict(name="pyrochlore", size=4, N=32, loops=3, Li=50.0, Lf=0.1,
                       workload="3-loop flow with self-energy corrections, pyrochlore L=4 Heisenberg SU(2), "
                                "N=32 (+1 points/axis), Lambda 50 -> 0.1"),
}


def flow_workload(args):
    """Full flow (BS3 step control, χ measured and meshes rescanned after every accepted step) of
    the square L=3 1-loop or the pyrochlore L=4 3-loop configuration, in FP64 and in FP32-mixed
    on the GPU; the FP64 oracle's cost is estimated from a sample of its right-hand sides and χ
    at the flow's start and end."""
    cfg = dict(FLOW_CONFIGS[args.flow_config])
    if args.flow_N:
        cfg["N"] = args.flow_N
        cfg["workload"] = cfg["workload"].replace(f"N={FLOW_CONFIGS[args.flow_config]['N']}", f"N={args.flow_N}")
    N = cfg["N"]
    prob = F.Problem.build(cfg["name"], cfg["size"], [[1.0]], initial=cfg["Li"], num_sigma=N, num_Omega=N, num_nu=N,
                           num_chi=10)
    a0 = prob.bare_action()
    chi_tol = (1e-8, 1e-4)
    runs = {}
    end_state = None
    modes = [m.strip() for m in args.flow_modes.split(",") if m.strip()]
    for tag in modes:
        mode = {"fp64": FP64, "fp32": FP32_MIXED}[tag]
        ctx = prob.context(n_slots=9, precision=mode)
        log(f"flow {args.flow_config} {tag}: Lambda {cfg['Li']} -> {cfg['Lf']}, {cfg['loops']}-loop, "
            f"{prob.r.n_sites} sites, N={N}")
        t0 = time.perf_counter()
        res = F.launch_flow(ctx, prob, a0, cfg["Li"], cfg["Lf"], loops=cfg["loops"], chi_tol=chi_tol,
                            verbose=args.flow_verbose, max_seconds=args.flow_max_seconds)
        dt = time.perf_counter() - t0
        if end_state is None:
            end_state = (ctx.download(F.A), ctx.m)
        ctx.close()
        runs[tag] = (res, dt)
        log(f"  {tag}: {dt:.2f} s, {res.rhs_evals} RHS, {res.accepted} accepted / {res.rejected} rejected, "
            f"chi {res.chi_s:.2f} s, status {res.status}, Lambda_end {res.Lambdas[-1] if res.Lambdas else None}")
    r64, t64 = runs.get("fp64", (None, None))
    r32, t32 = runs.get("fp32", (None, None))
    common, chi_err, mesh_diff = [], [], []
    if r64 is not None and r32 is not None:
        common = [(i, j) for i, L in enumerate(r64.Lambdas) for j, L2 in enumerate(r32.Lambdas)
                  if abs(L - L2) <= 1e-12 * L]
        chi_err = [float(np.max(np.abs(r32.chi[j] - r64.chi[i])) / np.max(np.abs(r64.chi[i]))) for i, j in common]
        mesh_diff = [max(float(np.max(np.abs(getattr(r32.meshes[j], f) - getattr(r64.meshes[i], f)) /
                                   np.max(getattr(r64.meshes[i], f)))) for f in ("sigma", "Os", "ns", "Ot", "nt", "chi"))
                     for i, j in common if i < len(r64.meshes) and j < len(r32.meshes)]
    last = common[-1] if common else None
    rref = r64 if r64 is not None else r32
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O  # checker-only import, cpu_baseline leg
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        # a right-hand side and a χ at the flow's start (bare action, initial meshes) and at its end
        # (the first run's final action on its final meshes)
        samples = []
        from pffrg.mesh import get_eval
        a_end, m_end = end_state
        for L, act, m in ((cfg["Li"], a0, prob.mesh), (rref.Lambdas[-1], a_end, m_end)):
            o = O.Oracle(prob.r, m)
            o.set_threads(threads)
            slots = [A.Action.empty(prob.r.n_sites, m) for _ in range(5)]
            log(f"  oracle sample at Lambda {L:.4g}")
            t0 = time.perf_counter()
            o.compute_dsigma(L, act, slots[0])
            ev = get_eval(L)
            if cfg["loops"] == 1:
                o.compute_dgamma_1l(L, act, slots[0], ev)
            else:
                o.compute_dgamma_ml(L, cfg["loops"], act, slots[0], slots[1], slots[2], slots[3], slots[4], ev)
                o.compute_dsigma_corr(L, act, slots[0], slots[4])
            t_rhs = time.perf_counter() - t0
            t0 = time.perf_counter()
            o.compute_chi(L, act, chi_tol)
            t_chi = time.perf_counter() - t0
            samples.append((L, t_rhs, t_chi))
            log(f"    {t_rhs:.1f} s RHS, {t_chi:.1f} s chi")
        t_rhs = float(np.mean([x[1] for x in samples]))
        t_chi = float(np.mean([x[2] for x in samples]))
        est = rref.rhs_evals * t_rhs + rref.accepted * t_chi
        cpu = {"value": 1.0 / est, "unit": "flows/s", "cores": threads, "kind": "port",
               "sample": f"FP64 oracle RHS and chi at the flow's first and last Lambda {[round(x[0], 3) for x in samples]} "
                         f"(mean {t_rhs:.3f} s/RHS, {t_chi:.3f} s/chi) x the GPU flow's {rref.rhs_evals} RHS and "
                         f"{rref.accepted} chi; estimated {est:.1f} s per flow", **host_cpu_info()}
    tv = t32 if t32 is not None else t64
    line = {
        "metric": "full-flow wall time (FP32-mixed vs plain FP64 GPU) at chi rel err <= 1e-4",
        "value": 1.0 / tv, "unit": "flows/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": tv * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("fp32-mixed (fp32 interpolation/products, fp64 quadrature and step control)" if t32 is not None
                  else "f64"),
        "data": "bare couplings (deterministic)",
        "config": {"workload": cfg["workload"] + ", BS3, chi and mesh rescanning after every accepted step",
                   "modes": modes, "max_seconds": args.flow_max_seconds},
        "wall_s": {k: v[1] for k, v in runs.items()},
        "speedup_vs_fp64": (t64 / t32) if (t64 and t32) else None,
        "rhs_evals": {k: v[0].rhs_evals for k, v in runs.items()},
        "accepted": {k: v[0].accepted for k, v in runs.items()},
        "rejected": {k: v[0].rejected for k, v in runs.items()},
        "status": {k: v[0].status for k, v in runs.items()},
        "Lambda_end": {k: (v[0].Lambdas[-1] if v[0].Lambdas else None) for k, v in runs.items()},
        "chi_s": {k: v[0].chi_s for k, v in runs.items()},
        "chi_rel_err_fp32_vs_fp64": {"max_over_common_steps": max(chi_err) if chi_err else None,
                                     "at_last_common_Lambda": chi_err[-1] if chi_err else None,
                                     "last_common_Lambda": r64.Lambdas[last[0]] if last else None,
                                     "common_steps": len(common)},
        "mesh_rel_diff_fp32_vs_fp64": max(mesh_diff) if mesh_diff else None,
        "chi_final": {k: v[0].chi[-1][:, 0].tolist() for k, v in runs.items() if v[0].chi},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def mlp_workload(args):
    """Training steps (backward + Adam) of the warm start's network at the size it trains with:
    2 inputs, three hidden layers of 512 (SiLU), 256 outputs, a batch of 64 samples; BF16 MFMA
    products (k_gemm_bf16) with FP32 accumulation."""
    from pffrg import learn
    net = learn.MLP(2, 512, 4, 256, seed=3)
    rng = np.random.default_rng(SEED)
    x = rng.standard_normal((64, 2)).astype(np.float32)
    t = rng.standard_normal((64, 256)).astype(np.float32)
    steps = max(args.steps, 1) * 100
    for _ in range(max(args.warmup, 1) * 20):
        net.backward(x, t)
        net.adam(1e-3)
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = net.backward(x, t)
        net.adam(1e-3)
    dt = time.perf_counter() - t0
    log(f"  {dt / steps * 1e3:.3f} ms per training step, loss {loss:.3e}")
    print(json.dumps({
        "metric": "warm-start network training steps/s (backward + Adam, batch 64)", "value": steps / dt,
        "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": max(args.warmup, 1) * 20,
        "ms_per_step": dt / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 products, fp32 accumulation and parameters", "data": "synthetic (seeded N(0,1) inputs and targets)",
        "config": {"workload": "MLP 2 -> 512 x3 (SiLU) -> 256, batch 64, backward + Adam per step"},
        "final_loss": float(loss)}), flush=True)


def warmstart_workload(args):
    """Learned parquet warm start: kagome L=5 J1-J2 Heisenberg, parquet solutions (absolute
    residual 1e-8, at most 500 iterations) at seeded (J2/J1 in [-0.5, 0.5], Λ in [0.5, 5])
    train a 4-layer width-512 MLP on per-channel SVD coefficients; on held-out couplings the
    iteration count from the decoded warm start is compared with the bare start."""
    from pffrg import learn
    N = args.ws_N
    pts = learn.coupling_samples(args.ws_train + args.ws_test, seed=5)
    train, test = pts[:args.ws_train], pts[args.ws_train:]
    t0 = time.perf_counter()

    def solve(j2, lam, init=None):
        prob = F.Problem.build("kagome", 5, [[1.0], [j2]], initial=lam, num_sigma=N, num_Omega=N, num_nu=N,
                               num_chi=10)
        bare = prob.bare_action()
        with prob.context(n_slots=9) as ctx:
            ts = time.perf_counter()
            res = F.launch_parquet(ctx, lam, bare if init is None else init(prob, bare), max_iter=500,
                                   parquet_tol=(1e-8, 0.0), max_seconds=100.0, verbose=args.ws_verbose)
            a = ctx.download(F.A)
        log(f"  J2={j2:+.3f} L={lam:.3f} {'bare' if init is None else 'warm'}: {res.iterations} it "
            f"({res.sde_iterations} SDE), residual {res.abs_err:.1e}, {res.status}, {time.perf_counter() - ts:.1f} s")
        return res, a, prob

    log(f"warmstart: kagome L=5, N={N}, {len(train)} training + {len(test)} held-out parquet solutions")
    acts, it_train = [], []
    for j2, lam in train:
        res, a, _ = solve(j2, lam)
        acts.append(a)
        it_train.append(res.iterations)
    t_data = time.perf_counter() - t0
    ws = learn.WarmStart(max_coef=256, width=512, n_layers=4, seed=1)
    t1 = time.perf_counter()
    rep = ws.fit(train, acts, steps=args.ws_steps, lr=1e-3, verbose=True)
    t_fit = time.perf_counter() - t1
    log(f"  fit: coefficients {rep.n_coef}, loss {rep.losses[0]:.2e} -> {rep.losses[-1]:.2e}, "
        f"train rel err per block (s, t, u, Sigma) {['%.1e' % e for e in rep.train_block_err]}, {t_fit:.1f} s")
    rows = []
    held_acts = []
    for j2, lam in test:
        rb, ab, _ = solve(j2, lam)
        held_acts.append(ab)
        rw, aw, _ = solve(j2, lam, lambda prob, bare: ws.initial_action((j2, lam), bare))
        d = float(np.max(np.abs(aw.gamma - ab.gamma)) / np.max(np.abs(ab.gamma)))
        rows.append({"J2": float(j2), "Lambda": float(lam), "status": [rb.status, rw.status],
                     "bare_iterations": rb.iterations,
                     "warm_iterations": rw.iterations, "bare_residual": rb.abs_err, "warm_residual": rw.abs_err,
                     "solution_rel_diff": d})
    # the iteration count is compared where both starts reach the residual; a solve that stalls
    # above it (quadrature noise of the right-hand side) runs to max_iter from either start
    conv = [r for r in rows if r["status"] == ["converged", "converged"]]
    ib = np.array([r["bare_iterations"] for r in conv], dtype=float)
    iw = np.array([r["warm_iterations"] for r in conv], dtype=float)
    line = {
        "metric": "parquet iterations to residual 1e-8: bare start / learned warm start",
        "value": float(ib.sum() / iw.sum()) if conv else None, "unit": "x fewer iterations", "n_gpus": 1,
        "steps": len(rows),
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 MFMA (fp32 accumulation) network, fp32 decoder, fp64 parquet",
        "data": f"parquet solutions generated here ({len(train)} training, {len(test)} held-out, seed 5)",
        "config": {"workload": f"kagome L=5 J1-J2 Heisenberg, N={N} (+1 points/axis), J2/J1 in [-0.5,0.5], "
                               f"Lambda in [0.5,5], MLP 4 x 512 SiLU, SVD coefficients per channel {rep.n_coef}"},
        "mean_iterations": {"bare": float(ib.mean()) if conv else None, "warm": float(iw.mean()) if conv else None},
        "converged_both": len(conv), "stalled": len(rows) - len(conv),
        "held_out": rows, "train_iterations_mean": float(np.mean(it_train)),
        "times_s": {"data": t_data, "fit": t_fit, "total": time.perf_counter() - t0},
        "fit": {"loss_first": rep.losses[0], "loss_last": rep.losses[-1], "train_rel_err": rep.train_rel_err,
                "train_block_err": dict(zip(("s", "t", "u", "Sigma"), rep.train_block_err)),
                "held_out_block_err": dict(zip(("s", "t", "u", "Sigma"), ws.held_out_errors(test, held_acts)))},
    }
    print(json.dumps(line), flush=True)


def batch_workload(args):
    """16 square L=5 J1-J2 problems (J2/J1 = k/15) and square L=3..7 at J2/J1 = 0.5, 1-loop,
    N=32 by default (--batch-N), couplings normalised per problem, one RHS each at Λ=1 with a seeded
    synthetic vertex: all 21 right-hand sides as one batched launch, against the same 21
    evaluated one problem at a time."""
    from pffrg.native import Batch
    specs = [(5, k / 15.0) for k in range(16)] + [(L, 0.5) for L in (3, 4, 5, 6, 7)]
    t0 = time.time()
    N = args.batch_N
    probs = [F.Problem.build("square", L, [[1.0], [j2]], initial=1.0, num_sigma=N, num_Omega=N, num_nu=N,
                             num_chi=10) for L, j2 in specs]
    acts = [A.synthetic_action(p.r.n_sites, p.mesh, p.bare, seed=SEED + k) for k, p in enumerate(probs)]
    log(f"batch workload: {len(probs)} problems, sites {[p.r.n_sites for p in probs]} (setup {time.time() - t0:.1f}s)")
    mode = FP32_MIXED
    ctxs = [p.context(n_slots=3, precision=mode) for p in probs]
    for c, a in zip(ctxs, acts):
        c.upload(0, a)
    n = len(ctxs)
    # one problem at a time
    for c in ctxs[:2]:
        F.rhs(c, 1.0, 1, 0, 1, 10)
    for c in ctxs:
        c.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        for c in ctxs:
            F.rhs(c, 1.0, 1, 0, 1, 10)
    for c in ctxs:
        c.synchronize()
    dt_seq = time.perf_counter() - t1
    log(f"  sequential: {dt_seq / args.steps * 1e3:.1f} ms per sweep of {n} RHS")
    with Batch(ctxs) as b:
        b.rhs_1l([1.0] * n, 0, 2, [10] * n)
        t2 = time.perf_counter()
        for _ in range(args.steps):
            b.rhs_1l([1.0] * n, 0, 2, [10] * n)
        dt_b = time.perf_counter() - t2
    same = all(np.array_equal(c.download(1).gamma, c.download(2).gamma) for c in ctxs)
    log(f"  batched: {dt_b / args.steps * 1e3:.1f} ms per sweep; identical to sequential: {same}")
    line = {
        "metric": "aggregate RHS evals/s, batched J1-J2 sweep + uneven truncations",
        "value": n * args.steps / dt_b, "unit": "RHS evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": dt_b / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32-mixed (fp32 interpolation/products, fp64 quadrature)",
        "data": "synthetic (seeded N(0,1) vertex with 1/(1+|W|+|v|+|v'|) decay)",
        "config": {"workload": "21 x 1-loop RHS: square L=5 J2/J1=k/15 (k=0..15) + square L=3..7 J2/J1=0.5, "
                               f"N={N} (+1 points/axis), Lambda=1, eval=10, one batched launch",
                   "problems": n, "sites": [p.r.n_sites for p in probs]},
        "sequential_rhs_per_s": n * args.steps / dt_seq, "batch_speedup": dt_seq / dt_b,
        "batched_equals_sequential": bool(same),
    }
    for c in ctxs:
        c.close()
    print(json.dumps(line), flush=True)


def batchflow_workload(args):
    """The batched configuration as full flows: 16 square L=5 J1-J2 problems (J2/J1 = k/15) and
    square L=3..7 at J2/J1 = 0.5, 1-loop, N=32 by default (--batch-N), Λ 50 -> 0.3 (couplings
    normalised per problem),
    BS3 with χ and mesh rescanning after every accepted step, FP32-mixed: all 21 flows advanced
    together (one batched launch per BS3 stage) against the same flows run one at a time."""
    specs = [(5, k / 15.0) for k in range(16)] + [(L, 0.5) for L in (3, 4, 5, 6, 7)]
    N = args.batch_N
    probs = [F.Problem.build("square", L, [[1.0], [j2]], initial=50.0, num_sigma=N, num_Omega=N, num_nu=N,
                             num_chi=10) for L, j2 in specs]
    n = len(probs)
    log(f"batched flows: {n} problems, sites {[p.r.n_sites for p in probs]}")
    t0 = time.perf_counter()
    seq = []
    for k, p in enumerate(probs if not args.batch_skip_single else []):
        ts = time.perf_counter()
        with p.context(n_slots=9, precision=FP32_MIXED) as ctx:
            r = F.launch_flow(ctx, p, p.bare_action(), 50.0, 0.3, max_seconds=args.flow_max_seconds)
        seq.append(r)
        log(f"  flow {k} (L={specs[k][0]}, J2={specs[k][1]:.3f}): {time.perf_counter() - ts:.1f} s, {r.rhs_evals} RHS, "
            f"{r.status} at Lambda {r.Lambdas[-1]:.4g}")
    dt_seq = time.perf_counter() - t0
    log(f"  one at a time: {dt_seq:.1f} s")
    ctxs = [p.context(n_slots=9, precision=FP32_MIXED) for p in probs]
    t1 = time.perf_counter()
    res = F.launch_flow_batch(ctxs, probs, [p.bare_action() for p in probs], 50.0, 0.3, progress=10,
                              max_seconds=args.flow_max_seconds)
    dt_b = time.perf_counter() - t1
    for c in ctxs:
        c.close()
    same = None if not seq else all(r.Lambdas == q.Lambdas and all(np.array_equal(x, y) for x, y in zip(r.chi, q.chi))
                                    for r, q in zip(res, seq))
    log(f"  batched: {dt_b:.1f} s; identical flows: {same}")
    rhs = sum(r.rhs_evals for r in res)
    line = {
        "metric": "full-flow wall time, 21 flows (J1-J2 sweep + uneven truncations): batched vs one at a time",
        "value": n / dt_b, "unit": "flows/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": dt_b * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-mixed (fp32 interpolation/products, fp64 quadrature and step control)",
        "data": "bare couplings (deterministic)",
        "config": {"workload": "21 x 1-loop flow: square L=5 J2/J1=k/15 (k=0..15) + square L=3..7 J2/J1=0.5, "
                               f"N={N} (+1 points/axis), Lambda 50 -> 0.3, chi and mesh rescanning every accepted step",
                   "problems": n, "sites": [p.r.n_sites for p in probs], "max_seconds": args.flow_max_seconds},
        "batched_s": dt_b, "one_at_a_time_s": dt_seq if seq else None, "batch_speedup": dt_seq / dt_b if seq else None,
        "rhs_evals": rhs, "aggregate_rhs_per_s": rhs / dt_b, "batched_equals_one_at_a_time": same,
        "status": [r.status for r in res], "Lambda_end": [r.Lambdas[-1] for r in res],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
