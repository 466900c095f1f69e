"""Benchmark: homogenization solves per second on one MI355X.

Workload (one "step"): a 256^3 periodic design (seeded randc density, seed 0, rotate3 symmetry,
volume 0.2), E = 10, nu = 0.3, SIMP x = rho^3, Emin = 1e-9.  A step uploads nothing: the density is
resident in HBM.  It runs Homogenization::update (Galerkin coarse operators + coarsest FP64
factorisation), the six macro-strain load cases by MG-PCG from a zero initial guess to
relres = ||f - K u|| / ||f|| <= 1e-5 (FP64 norms), and C^H.

Prints one JSON line.  The dominant kernel's (the CG operator pass, fine_apply) average launch time
is measured with HIP events on the library's stream during the timed steps, with event pairs around
that kernel only; one more, untimed step with events around every finest-level kind logs their
launch times.  The CPU baseline is the FP64 oracle (test infrastructure) timed on the host on a
bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "bench_256cube_pcg_traffic.json")
PROF_KINDS = ("fine_gs", "fine_gs_zero", "fine_apply", "fine_apply_first", "fine_residual")
# rocprofv3 names of the kernel behind each timed kind (demangled, or mangled for the packed-FP16
# instantiations), for the traffic lookup
PROF_KERNEL = {"fine_gs": ("k_fine_gs2<6, true, -1", "k_fine_gs2ILi6ELb1ELin1E", "k_fine_gs<6, true, -1",
                           "k_fine_gsILi6ELb1ELin1E"),
               "fine_apply": ("k_fine_cg_stagedILb1ELi2E", "k_fine_cg_staged<true, 2>",
                              "k_fine_applyILi6ELi0ELb1EDv2_DF16_Lb1ELi2E", "k_fine_apply_efILi0ELb1ELb1ELi1ELi2E",
                              "k_fine_apply_efILi0ELb1ELb1ELi2ELi2E", "k_fine_apply<6, 0, true, half2, true, 2"),
               "fine_residual": ("k_fine_apply<6, 1, false", "k_fine_applyILi6ELi1ELb0E")}
DOMINANT = "fine_apply"  # the kernel with the largest total time per step
NET_WEIGHTS = os.path.join(ROOT, "homo3d_amd", "data", "warmstart_unet64.npz")
CH_FIXTURE = os.path.join(ROOT, "tests", "golden", "ch_256_rotate3_seed0.json")
METRIC = "solves/s to relres 1e-5 @256^3, learned-warm-start MG vs plain MG; stencil GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gs_colours_per_launch(n):
    """Colours per finest-level GS launch: two when whole grid rows of a parity block fit a 256-thread
    workgroup (the library fuses colour pairs then), else one."""
    h = n // 2
    return 2 if (n % 2 == 0 and h <= 256 and 256 % h == 0 and h ** 3 % 256 == 0) else 1


def kernel_bytes(name, n, R=6, s_z=2, s_f=4, s_w=4):
    """Algorithmic HBM bytes per launch of a finest-level kernel.  p, q, r, f, x are FP32 (s_f); the
    preconditioner iterate z (u of the V-cycle's finest level) is packed FP16 (s_z = 2) when n % 8 == 0."""
    N = n ** 3
    if name == "fine_apply":        # CG pass: read z, q_old, p_old, x, w; write q = K z + beta q_old, p, x
        return N * (3 * R * (s_z + 6 * s_f) + s_w)
    if name == "fine_apply_first":  # first CG pass of an inner solve: read z, w; write q = K z, p = z
        return N * (3 * R * (s_z + 2 * s_f) + s_w)
    if name == "fine_residual":   # r = f - K z: read z, f, write r, read w
        return N * (3 * R * (s_z + s_f + s_f) + s_w)
    if name == "fine_gs":         # c colours = c/8 of an in-place sweep: z read+write, f read, w read
        return N * (3 * R * (2 * s_z + s_f) + s_w) * gs_colours_per_launch(n) / 8.0
    return None


def gs_pair_must_read(n, R=6, s_z=2, s_f=4, s_w=4):
    """HBM bytes a colour-pair GS launch must move at least: every vertex neighbours the pair's colours, so
    the whole z field and every cell weight is read; f of the pair's vertices is read and their z words
    written (c/8 of the vertices)."""
    N = n ** 3
    c = gs_colours_per_launch(n) / 8.0
    return N * (3 * R * s_z + s_w) + N * c * (3 * R * s_f + 3 * R * s_z)


def measured_traffic(name, n):
    """HBM bytes per launch of `name` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (256^3 only)."""
    if n != 256 or not os.path.exists(TRAFFIC_JSON) or name not in PROF_KERNEL:
        return None
    data = json.load(open(TRAFFIC_JSON))["kernels"]
    vals = [v["traffic_bytes"] for k, v in data.items() if any(p in k for p in PROF_KERNEL[name])]
    return sum(vals) / len(vals) if vals else None


def cpu_baseline(n_sample, n_target, tol, iters):
    """The FP64 oracle (test infrastructure) on the host threads, as a bounded sample of the same solve.

    At the target size (the default): the full Galerkin set-up of the design's hierarchy is timed, then one
    operator application K u and one symmetric V-cycle, the work of one CG step of the oracle's MG-PCG; the
    six load cases are extrapolated as (steps + 1) of those each (the start forms one residual and one
    V-cycle), with `iters` the CG step counts of the measured GPU solve.  With a smaller --cpu-n: one full
    solve there, scaled by the cell count."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    rho = oracle.randc_density(n_sample, 0, 0.2, "rotate3").astype(np.float64)
    if n_sample == n_target:
        t0 = time.perf_counter()
        mg = oracle.MG(n_sample)
        mg.set_density(rho ** 3)
        t_setup = time.perf_counter() - t0
        f = oracle.macro_rhs(n_sample, mg.KE, mg.w, 0)
        u = np.zeros_like(f)
        t0 = time.perf_counter()
        oracle.apply_K(n_sample, mg.KE, mg.w, f)
        t_k = time.perf_counter() - t0
        t0 = time.perf_counter()
        mg.vcycle(u, f, symmetric=True)
        t_v = time.perf_counter() - t0
        del mg, f, u
        total = t_setup + sum((k + 1) * (t_k + t_v) for k in iters)
        sample = ("%d^3 measured: FP64 oracle on %d threads, Galerkin set-up %.1f s, one K u %.2f s, one symmetric "
                  "V-cycle %.2f s; six load cases at (CG steps + 1) x (K u + V-cycle) with the GPU run's CG steps "
                  "%s: %.0f s per solve" % (n_sample, threads, t_setup, t_k, t_v, list(iters), total))
        return {"value": 1.0 / total, "unit": "solves/s", "cores": threads, "kind": "port", "sample": sample}, None
    t0 = time.perf_counter()
    mg = oracle.MG(n_sample)
    mg.set_density(rho ** 3)
    C, chi, code, its = mg.homogenize(tol=tol, maxit=400, mode="pcg")
    dt = time.perf_counter() - t0
    del chi, mg
    scale = (n_sample / n_target) ** 3
    sample = ("FP64 oracle, full solve at %d^3 (Galerkin set-up + 6 load cases, MG-PCG to %.0e, iters %s) "
              "in %.1f s on %d threads; scaled by (%d/%d)^3 to %d^3" %
              (n_sample, tol, its, dt, threads, n_sample, n_target, n_target))
    return {"value": (1.0 / dt) * scale, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": sample}, C


def ch_errors(C, Cref):
    """Per-entry relative error of C^H against the oracle over the entries that are not zero by
    symmetry (|Cref_ij| > 1e-6 max|Cref|), and the largest absolute error relative to max|Cref|."""
    C, Cref = np.asarray(C, float), np.asarray(Cref, float)
    big = np.abs(Cref) > 1e-6 * np.abs(Cref).max()
    rel = np.abs(C - Cref)[big] / np.abs(Cref)[big]
    return {"max_rel_err_per_entry": float("%.3e" % rel.max()),
            "max_abs_err_over_max_entry": float("%.3e" % (np.abs(C - Cref).max() / np.abs(Cref).max())),
            "entries_compared": int(big.sum())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--cpu-n", type=int, default=0,
                    help="oracle grid of the CPU baseline (default --n: set-up and one CG step's work timed there; "
                         "a smaller grid: one full solve, scaled to --n)")
    ap.add_argument("--net", default=NET_WEIGHTS, help="trained warm-start network ('' to skip the learned leg)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=64,
                    help="designs of the batched uneven-design leg (sizes from {32,48,64,80,96}^3; 0 skips it)")
    ap.add_argument("--omega", type=float, default=None, help="smoother relaxation factor (default: the library's)")
    ap.add_argument("--mfma", type=int, default=None,
                    help="finest-level operator on FP16 fields in element form on the matrix cores (1) or row form (0); "
                         "default: the library default")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import homo3d_amd

    n = args.n
    log("bench: n=%d tol=%.1e steps=%d warmup=%d" % (n, args.tol, args.steps, args.warmup))
    rho = homo3d_amd.randc_density(n, 0, 0.2, "rotate3")
    x = homo3d_amd.DeviceBuffer.from_array((rho.astype(np.float64) ** 3).astype(np.float32))
    sync = homo3d_amd.DeviceBuffer.synchronize
    sync()
    kw = {} if args.mfma is None else {"fine_mfma": bool(args.mfma)}
    if args.omega is not None:
        kw["sor_omega"] = args.omega
    hom = homo3d_amd.Homogenization(homo3d_amd.HomoConfig(reso=n, solver="pcg", max_iter=400, **kw))

    def step():
        hom.update(x.value, n, on_device=True)
        st = hom.solve(args.tol, warm=False)
        C = hom.elasticMatrix(solve=False)
        return st, C

    def check(st, what):
        # every counted solve must have met the tolerance on all six load cases
        if st["status"] != 0 or not max(st["relres"]) <= args.tol:
            log("%s: status %d relres %s: not converged, no number reported" % (what, st["status"], st["relres"]))
            sys.exit(1)

    worst = 0.0
    for i in range(args.warmup):
        st, C = step()
        check(st, "warmup %d" % i)
        log("warmup %d: iters %s relres %.2e %.2fs" % (i, st["iters"], max(st["relres"]), st["seconds"]))
    hom.profile_enable(True, only=DOMINANT)
    sync()
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        st, C = step()
        stats.append(st)
    sync()
    dt = time.perf_counter() - t0
    for i, st in enumerate(stats):
        check(st, "step %d" % i)
        worst = max(worst, max(st["relres"]))
        log("step %d: iters %s relres %.2e" % (i, st["iters"], max(st["relres"])))
    dom = DOMINANT
    cnt, ms = hom.profile_get(dom)
    hom.profile_enable(False)
    s_z = 2 if n % 8 == 0 else 4
    # untimed step with events around every kind, for the per-kernel log
    hom.profile_enable(True)
    step()
    prof = {k: hom.profile_get(k) for k in PROF_KINDS}
    hom.profile_enable(False)
    for k, (c, t) in prof.items():
        if c:
            kb = kernel_bytes(k, n, s_z=s_z)
            log("kernel %-16s launches %4d total %8.2f ms avg %8.4f ms%s" %
                (k, c, t, t / c, "  %.1f GB/s algorithmic" % (kb / (t / c * 1e-3) / 1e9) if kb else ""))
    if max(prof.values(), key=lambda ct: ct[1]) != prof[dom]:
        log("note: %s is not the kernel kind with the largest total in this run" % dom)
    avg_s = ms / cnt * 1e-3
    achieved = kernel_bytes(dom, n, s_z=s_z) / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(dom, n), "kernel": dom,
                "avg_launch_ms": round(ms / cnt, 5),
                "algorithmic_bytes_per_launch": kernel_bytes(dom, n, s_z=s_z)}
    # the reverse colour-pair GS launch (VALU-bound): its must-read bytes against the counter traffic
    gs_c, gs_t = prof["fine_gs"]
    gs_pair = None
    if gs_c and gs_colours_per_launch(n) == 2:
        gb = gs_pair_must_read(n, s_z=s_z)
        gs_ms = gs_t / gs_c
        gtr = measured_traffic("fine_gs", n)
        gs_pair = {"kernel": "fine_gs", "avg_launch_ms": round(gs_ms, 5), "must_read_bytes_per_launch": gb,
                   "achieved": round(gb / (gs_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(gb / (gs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5), "traffic": gtr,
                   "traffic_over_must_read": round(gtr / gb, 3) if gtr else None}
    out = {
        "metric": METRIC,
        "value": round(args.steps / dt, 5),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",  # FP32 arithmetic, FP64 refinement residuals and dots; z and coarse stencils stored as FP16
        "data": "synthetic",
        "config": {"workload": "%d^3 periodic homogenization: 6 macro-strain load cases, plain MG-PCG "
                               "(V(1,1) 8-colour block GS, FP64 coarsest) to relres %.0e, seeded randc density "
                               "(seed 0, rotate3, vol 0.2), E=10 nu=0.3 p=3" % (n, args.tol),
                   "n": n, "tol": args.tol, "solver": "mg-pcg", "warm_start": "none",
                   "iters": stats[-1]["iters"], "relres": [float("%.3e" % r) for r in stats[-1]["relres"]],
                   "worst_relres_all_steps": float("%.3e" % worst)},
        "roofline": roofline,
        "gs_pair": gs_pair,
    }
    if n == 256 and os.path.exists(CH_FIXTURE):
        # the FP64 oracle's C^H of this design (tests/golden/make_ch256.py)
        ref = json.load(open(CH_FIXTURE))
        out["config"]["ch_vs_fp64_oracle"] = ch_errors(C, np.array(ref["C"]).reshape(6, 6))
        out["config"]["ch_vs_fp64_oracle"]["oracle_solve_s"] = round(ref["setup_s"] + ref["solve_s"], 1)
        out["config"]["ch_vs_fp64_oracle"]["oracle_threads"] = ref["threads"]
    if args.net and os.path.exists(args.net):
        out["learned"] = learned_leg(homo3d_amd, hom, x, n, args, sync, check, log)
        out["plain"] = {"value": out["value"], "ms_per_step": out["ms_per_step"]}
        out["learned_vs_plain"] = round(out["learned"]["value"] / out["value"], 4)
    if args.batch > 0:
        out["batch"] = batch_leg(homo3d_amd, args.batch, args.tol, log)
    if not args.no_cpu:
        cpu_n = args.cpu_n or n
        log("cpu baseline: oracle at %d^3" % cpu_n)
        out["cpu_baseline"], Cref = cpu_baseline(cpu_n, n, args.tol, stats[-1]["iters"])
    print(json.dumps(out), flush=True)


def batch_leg(homo3d_amd, count, tol, log):
    """Batched uneven designs: `count` seeded randc designs, sizes drawn uniformly from {32,48,64,80,96}^3 and
    volume fractions from U(0.1, 0.5), six load cases each, solved to relres `tol` through homo3d_batch_*
    (one handle and stream per design, several designs in flight).  One untimed call creates the
    hierarchies and autotunes each design's operator form; the timed call updates all densities and
    solves them.  The same designs one at a time (one in flight) give the sequential rate."""
    rng = np.random.default_rng(2024)
    sizes = [int(v) for v in rng.choice([32, 48, 64, 80, 96], count)]
    vols = rng.uniform(0.1, 0.5, count)
    xs = [(homo3d_amd.randc_density(nn, 100 + i, float(vols[i]), "none").astype(np.float64) ** 3).astype(np.float32)
          for i, nn in enumerate(sizes)]
    cfg = homo3d_amd.HomoConfig(solver="pcg", max_iter=400)
    rec = {"designs": count, "sizes": {str(k): sizes.count(k) for k in (32, 48, 64, 80, 96)}, "tol": tol}
    for inflight in (12, 1):  # 12 in flight measured best (4..32 swept: 63, 82, 95, 95, 91, 92 designs/s)
        b = homo3d_amd.Batch(sizes, cfg, max_inflight=inflight)
        b.solve(xs, tol, autotune=True)
        t0 = time.perf_counter()
        C, stats = b.solve(xs, tol)
        dt = time.perf_counter() - t0
        b.close()
        worst = max(max(st["relres"]) for st in stats)
        if any(st["status"] != 0 for st in stats) or worst > tol:
            raise RuntimeError("batch: a design missed relres %.0e (worst %.2e)" % (tol, worst))
        log("batch: %d designs, %d in flight: %.3f s (%.1f designs/s), worst relres %.2e" %
            (count, inflight, dt, count / dt, worst))
        if inflight > 1:
            rec.update({"value": round(count / dt, 3), "unit": "designs/s", "seconds": round(dt, 4),
                        "max_inflight": inflight, "worst_relres": float("%.3e" % worst),
                        "cg_steps_total": int(sum(sum(st["iters"]) for st in stats))})
        else:
            rec["sequential_designs_per_s"] = round(count / dt, 3)
    rec["speedup_vs_sequential"] = round(rec["value"] / rec["sequential_designs_per_s"], 3)
    return rec


def learned_leg(homo3d_amd, hom, x, n, args, sync, check, log):
    """The same steps with the learned warm start: update, network prediction as the FP64 start of the
    refined solve (a load case whose predicted start has relres >= 1 restarts from zero), solve, C^H."""
    net = homo3d_amd.WarmStartNet.load(args.net)

    def lstep():
        hom.update(x.value, n, on_device=True)
        net.warm_start(hom)
        st = hom.solve(args.tol, warm=True)
        C = hom.elasticMatrix(solve=False)
        return st, C

    for i in range(args.warmup):
        st, _ = lstep()
        check(st, "learned warmup %d" % i)
    sync()
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        st, C = lstep()
        stats.append(st)
    sync()
    dt = time.perf_counter() - t0
    for i, st in enumerate(stats):
        check(st, "learned step %d" % i)
        log("learned step %d: iters %s relres %.2e" % (i, st["iters"], max(st["relres"])))
    # untimed: the prediction's own FP64 relres per load case
    hom.update(x.value, n, on_device=True)
    net.warm_start(hom)
    start = hom.residual_norms()
    rec = {"value": round(args.steps / dt, 5), "ms_per_step": round(dt / args.steps * 1e3, 3),
           "iters": stats[-1]["iters"], "relres": [float("%.3e" % r) for r in stats[-1]["relres"]],
           "start_relres": [float("%.3e" % r) for r in start],
           "network": {"grid": net.grid, "levels": net.levels, "channels": net.base_channels,
                       "params": net.num_params}}
    if n == 256 and os.path.exists(CH_FIXTURE):
        ref = json.load(open(CH_FIXTURE))
        rec["ch_vs_fp64_oracle"] = ch_errors(C, np.array(ref["C"]).reshape(6, 6))
    return rec


if __name__ == "__main__":
    main()
