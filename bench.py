"""Benchmark of the restarted-PDHG engine on one MI355X.

Workload (one "step"): the 64-case price family of the refinery-style LP
(config C3: 12 periods x 40 streams x 36 products x 30 qualities, 555,852
nonzeros, rows/columns scaled by 10^U(-3,3); costs c_k = c0 + theta_k d on
four crude purchase prices, theta_k ~ U[-0.3, 0.3]^4, seed 4) solved by ONE
batched run of width 64 (FP64 iterates, hipGraphs of 64 iterations,
autotuned row mapping) to relative KKT <= 1e-6, then every answer certified
(implied bounds + rigorous safe dual bound, acceptance
|obj - LB| <= 1e-4 (1 + |obj|)).  A step is: (warm start only) MLP inference
theta -> POD coefficients and BF16-MFMA decode of the starting points, upload,
batched solve, download, certification.  The value is certified solves per
second of wall time over the timed steps (host <-> device transfers
included).

Start of a step: the learned warm start (with its per-case cold-start
fallback: a warm case not converged after 256k iterations restarts in its
lane from the reference's cold start) or the reference's cold start.  The
warmup steps run warm, then one run of the same family runs cold; the timed
steps use whichever certified more solves per second in this run (autotuning
seconds excluded), and the line reports both runs and the choice
(--start forces one).

Before the timer: the warm-start model is fitted once on 64 other parameters
of the family (solved at the same tolerance from a cold start; POD basis, MLP
trained on the GPU); its cost is reported as warm_start_fit_seconds.  Then
the per-launch roofline of the dominant iteration kernel, the standalone SpMV
rates on the refinery matrix and on a refinery LP beyond the 256 MB Infinity
Cache, plain single-instance GPU PDHG on a sample of the cases, and the
reference's 16-thread CPU PDHG on a bounded sample of the same workload.

Wall budget: the whole run stays inside --budget seconds (default 330, well
below a 600 s limit on the run): warmup and timed steps stop early when
the next step would not fit, and the line reports the steps actually timed.
After them, when the budget still allows: config C4's 256 uneven LPs in one
packed run.

    python bench.py [--steps 20] [--warmup 5] [--gpus 1] [--budget 330]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "optimized-gpu-accelerator_amd"))

import pdhg_amd as P  # noqa: E402
import warmstart as W  # noqa: E402

PEAK_HBM_GBS = 8000.0
SEED_LP, SEED_THETA, SEED_TRAIN, SEED_PACKED = 3, 4, 5, 7
T0 = time.perf_counter()  # process clock of the wall budget


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fit_warm_start(lp, n_train, tol, opts, steps):
    """PDHG solutions of n_train other family parameters at tolerance tol,
    POD basis and GPU-trained MLP.  The first 64 parameters are solved from a
    cold start; a model fitted to them warm-starts the rest, which stream
    through one batched run; the final model is fitted to all of them."""
    th, costs = P.gen_family(lp, SEED_TRAIN, n_train, 0.3)
    o = P.default_options(tolerance=tol, precision=opts.precision, max_iterations=opts.max_iterations,
                          progress_seconds=15.0, row_map=opts.row_map)
    first = min(64, n_train)
    res, st, _ = P.solve_batch(lp, costs[:first], o)
    log(f"warm-start snapshots {first}/{n_train} (cold start): {st['iterations']} iterations")

    def fit(k, rs):
        # Only converged solutions are snapshots: a case stopped at the
        # iteration limit is not a solution of its parameter.
        ok = [q for q in range(k) if rs[q].status == P.OPTIMAL]
        if len(ok) < 2:
            return None
        return W.WarmStart(lp, rank=32, hidden=128).fit(th[ok], np.stack([rs[q].x for q in ok]),
                                                       np.stack([rs[q].y for q in ok]), steps=steps, lr=1e-3,
                                                       omegas=np.array([rs[q].primal_weight for q in ok]))

    if n_train > first:
        ws0 = fit(first, res)
        x0 = y0 = w0 = None
        if ws0 is not None:
            x0, y0 = ws0.predict(th[first:])
            w0 = ws0.primal_weights(th[first:])
        rest, st, _ = P.solve_batch(lp, costs[first:], o, x0=x0, y0=y0, primal_weight=w0)
        log(f"warm-start snapshots {n_train}/{n_train} (warm start from the first {first}): "
            f"{st['iterations']} iterations")
        res = list(res) + list(rest)
    ws = fit(n_train, res)
    if ws is None:  # cold start is the fallback
        log("warm start: fewer than two training solves converged; the step runs from a cold start")
        return None
    log(f"warm-start MLP: loss {ws.losses[0]:.3g} -> {ws.losses[-1]:.3g}, POD energy kept "
        f"{ws.energy[:ws.used].sum() / max(ws.energy.sum(), 1e-300):.6f}, "
        f"fitted to the {sum(r.status == P.OPTIMAL for r in res)}/{n_train} converged training solves")
    return ws


def family_step(lp, thetas, costs, opts, ws=None, threads=16):
    x0 = y0 = w0 = None
    if ws is not None:
        x0, y0 = ws.predict(thetas)
        w0 = ws.primal_weights(thetas)
    res, st, _ = P.solve_batch(lp, costs, opts, x0=x0, y0=y0, primal_weight=w0)
    xs = np.stack([r.x for r in res])
    ys = np.stack([r.y for r in res])
    obj = np.array([r.objective for r in res])
    cert = P.certify_batch(lp, costs, xs, ys, obj, threads=threads)
    ok = np.array([r.status == P.OPTIMAL for r in res]) & (cert[:, 3] > 0)
    return res, st, cert, ok


def reference_check(res, cert, tol, k):
    """The step's answers on the cases that the reference's own CPU PDHG (its
    sources, FP64) solved ahead of time (tests/golden/c3_reference_cases.json,
    the default workload only): objective error and certified dual bound
    difference relative to the reference's, iteration counts apart."""
    path = os.path.join(ROOT, "tests", "golden", "c3_reference_cases.json")
    if tol != 1e-6 or k != 64 or not os.path.exists(path):
        return None
    with open(path) as f:
        cases = json.load(f)["cases"]
    obj, lb, its = [], [], []
    for q, rec in cases.items():
        r = res[int(q)]
        obj.append(abs(r.objective - rec["objective"]) / (1 + abs(rec["objective"])))
        lb.append(abs(cert[int(q), 0] - rec["safe_bound"]) / (1 + abs(rec["safe_bound"])))
        its.append(abs(r.iterations - rec["iterations"]))
    return {"cases": len(cases), "max_rel_objective_error": max(obj), "max_rel_bound_difference": max(lb),
            "max_iteration_difference": int(max(its)),
            "source": "tests/golden/c3_reference_cases.json (reference solvePdhg, its own sources, FP64)"}


def spmv_rates(lp, reps=200):
    """Standalone CSR SpMV on the scaled matrix: the op-level backend's k_spmv
    (no atomics, rows bucketed by length; the reference's Backend::spmv and
    spmvT, src/pdhg/pdhg.cpp:93-103), A x and A^T y at widths 1 and 64, FP64
    and FP32 values, HIP events over `reps` launches on the backend's stream.
    Algorithmic bytes: values and indices once, row pointers, input and output
    vectors once."""
    import scipy.sparse as sp
    pr = P.prepare(lp)
    A = sp.csr_matrix((pr["val"], pr["idx"], pr["ptr"]), shape=(lp.m, lp.n))
    AT = A.T.tocsr()
    AT.sort_indices()
    rfac, cfac = pr["beta"] / pr["R"], pr["gamma"] / pr["C"]
    rng = np.random.default_rng(0)
    out = {}
    for prec, name, vs in ((P.FP64, "fp64", 8), (P.FP32, "fp32", 4)):
        for k in (1, 64):
            be = P.Backend(A, AT, rfac, cfac, k=k, precision=prec)
            be.set(P.V_X, rng.standard_normal((lp.n, k)))
            be.set(P.V_Y, rng.standard_normal((lp.m, k)))
            r = {}
            for transpose, key, rows, cols in ((False, "Ax", lp.m, lp.n), (True, "ATy", lp.n, lp.m)):
                ms = be.time_spmv(transpose, reps)
                nbytes = lp.nnz * (vs + 4) + (rows + 1) * 4 + (rows + cols) * k * vs
                r[key] = {"us": ms * 1e3, "GBs": nbytes / (ms * 1e-3) / 1e9, "bytes": nbytes,
                          "layout": be.spmv_layout(transpose)}
            be.close()
            out[f"{name}_B{k}"] = r
    return out


def cpu_baseline(lp, costs, gpu_iters, seconds, threads):
    """The reference CPU PDHG (its own sources, 16 threads) on the same batched
    family, run for a bounded number of iterations; solves/s extrapolated to the
    iterations its batch would run (every case iterates until the slowest one
    stops, src/pdhg/pdhg.cpp:606-607: the largest per-case count, capped at the
    iteration limit).  The extrapolation factor is stated in the sample."""
    K = costs.shape[0]
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    if not O.ref_available():
        return None
    O.ref_set_threads(threads)
    probe = 64
    t = time.perf_counter()
    O.ref_solve_batch(lp, costs, tol=1e-30, max_iter=probe)
    per = (time.perf_counter() - t) / probe
    iters = max(probe, int(seconds / max(per, 1e-9)) // 64 * 64)
    t = time.perf_counter()
    O.ref_solve_batch(lp, costs, tol=1e-30, max_iter=iters)
    el = time.perf_counter() - t
    per = el / iters
    factor = gpu_iters / iters
    return {
        "value": K / (per * gpu_iters),
        "unit": "solves/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"the reference's batched CPU PDHG (its own sources, FP64, {threads} threads) on the same "
                   f"{K}-case family for {iters} iterations in {el:.1f} s = {per * 1e3:.2f} ms per iteration; "
                   f"solves/s = {K} cases / ({per * 1e3:.2f} ms x {gpu_iters} iterations, the iteration count its "
                   f"batch would run), an extrapolation by {factor:.0f}x in iterations; host nproc={os.cpu_count()}"),
        "seconds_per_iteration": per,
        "extrapolation_factor": factor,
    }


def spmv_beyond_cache(periods, reps=50):
    """Standalone CSR SpMV (the op-level backend's autotuned k_spmv, width 1)
    on a refinery LP of `periods` periods whose matrix exceeds the 256 MB
    Infinity Cache (1200 periods: 55.6M nonzeros, 445 MB of FP32 values and
    indices), A x and A^T y in FP32 and FP64, HIP events over `reps` launches.
    The matrix is the generator's unscaled one (the kernel's rate does not
    depend on the values), which saves the host set-up of the scaling."""
    import scipy.sparse as sp

    lp = P.gen_refinery(SEED_LP, periods=periods)
    A = sp.csc_matrix((lp.value, lp.index, lp.start), shape=(lp.m, lp.n)).tocsr()
    A.sort_indices()
    AT = A.T.tocsr()
    AT.sort_indices()
    rng = np.random.default_rng(0)
    out = {"workload": f"refinery LP, {periods} periods: m={lp.m} n={lp.n} nnz={lp.nnz}, width 1"}
    for prec, name, vs in ((P.FP32, "fp32", 4), (P.FP64, "fp64", 8)):
        be = P.Backend(A, AT, np.ones(lp.m), np.ones(lp.n), k=1, precision=prec)
        be.set(P.V_X, rng.standard_normal(lp.n))
        be.set(P.V_Y, rng.standard_normal(lp.m))
        for transpose, key, rows, cols in ((False, "Ax", lp.m, lp.n), (True, "ATy", lp.n, lp.m)):
            ms = be.time_spmv(transpose, reps)
            nbytes = lp.nnz * (vs + 4) + (rows + 1) * 4 + (rows + cols) * vs
            gbs = nbytes / (ms * 1e-3) / 1e9
            out[f"{name}_{key}"] = {"us": ms * 1e3, "GBs": gbs, "bytes": nbytes, "frac": gbs / PEAK_HBM_GBS,
                                    "layout": be.spmv_layout(transpose)}
        be.close()
    log("spmv beyond the Infinity Cache: " + ", ".join(
        f"{k} {v['GBs']:.0f} GB/s ({v['frac']:.2f})" for k, v in out.items() if k != "workload"))
    return out


def large_lp_rates(periods, reps=50):
    """Kernels on a refinery LP too large for the 256 MB Infinity Cache
    (`periods` periods of the staircase: 1200 gives 55.6M nonzeros, 445 MB of
    FP32 values and indices per matrix): the op-level CSR SpMV at width 1
    (A x and A^T y, FP32 and FP64 values), and the fused iteration kernels of
    one instance (FP32 and FP64 iterates), HIP events over `reps` launches."""
    lp = P.gen_refinery(SEED_LP, periods=periods)
    log(f"large LP: {periods} periods, m={lp.m} n={lp.n} nnz={lp.nnz}")
    import scipy.sparse as sp

    pr = P.prepare(lp)
    A = sp.csr_matrix((pr["val"], pr["idx"], pr["ptr"]), shape=(lp.m, lp.n))
    AT = A.T.tocsr()
    AT.sort_indices()
    rfac, cfac = pr["beta"] / pr["R"], pr["gamma"] / pr["C"]
    rng = np.random.default_rng(0)
    out = {"workload": f"refinery LP, {periods} periods: m={lp.m} n={lp.n} nnz={lp.nnz}", "spmv": {}, "iteration": {}}
    for prec, name, vs in ((P.FP32, "fp32", 4), (P.FP64, "fp64", 8)):
        be = P.Backend(A, AT, rfac, cfac, k=1, precision=prec)
        be.set(P.V_X, rng.standard_normal(lp.n))
        be.set(P.V_Y, rng.standard_normal(lp.m))
        r = {}
        for transpose, key, rows, cols in ((False, "Ax", lp.m, lp.n), (True, "ATy", lp.n, lp.m)):
            ms = be.time_spmv(transpose, reps)
            nbytes = lp.nnz * (vs + 4) + (rows + 1) * 4 + (rows + cols) * vs
            r[key] = {"us": ms * 1e3, "GBs": nbytes / (ms * 1e-3) / 1e9, "bytes": nbytes,
                      "frac": nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, "layout": be.spmv_layout(transpose)}
        be.close()
        out["spmv"][name] = r
        log(f"large-LP spmv {name}: A x {r['Ax']['us']:.1f} us {r['Ax']['GBs']:.0f} GB/s, "
            f"A^T y {r['ATy']['us']:.1f} us {r['ATy']['GBs']:.0f} GB/s")
        _, st, _ = P.solve(lp, P.default_options(tolerance=1e-30, max_iterations=64 * 4, precision=prec,
                                                 kernel_timing_reps=reps))
        it = {}
        for k in ("col", "row"):
            b, ms = st[f"{k}_kernel_bytes"], st[f"{k}_kernel_ms"]
            it[f"k_{k}"] = {"us": ms * 1e3, "bytes": b, "GBs": b / (ms * 1e-3) / 1e9,
                            "frac": b / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
        out["iteration"][name] = it
        log(f"large-LP iteration {name}: k_col {it['k_col']['us']:.1f} us {it['k_col']['GBs']:.0f} GB/s, "
            f"k_row {it['k_row']['us']:.1f} us {it['k_row']['GBs']:.0f} GB/s")
    return out


def elapsed():
    return time.perf_counter() - T0


def kernel_roofline(lp, costs, prec, row_map, xcd_remap, tol, graph=1):
    """Per-launch roofline of the iteration kernels at width 64: HIP events on
    the run's own stream over 200 launches each, after a short solve of the
    workload's first 64 cases with the tuned mapping."""
    _, kst, _ = P.solve_batch(lp, costs[:64], P.default_options(tolerance=tol, precision=prec, max_iterations=64 * 8,
                                                                kernel_timing_reps=200, row_map=row_map,
                                                                xcd_remap=xcd_remap, use_graph=graph))
    kern = {"k_col": (kst["col_kernel_bytes"], kst["col_kernel_ms"]),
            "k_row": (kst["row_kernel_bytes"], kst["row_kernel_ms"])}
    dom = max(kern, key=lambda k: kern[k][1])
    dbytes, dms = kern[dom]
    achieved = dbytes / (dms * 1e-3) / 1e9
    kprec = "fp64" if prec == P.FP64 else "fp32"
    rmap = kst["col_row_map" if dom == "k_col" else "row_row_map"]
    # LDS-staged tiles of 16, 8 or 4 rows: k_col_tiled, k_col_tiled8, k_col_tiled4
    name = dom + {3: "_tiled", 4: "_tiled8", 5: "_tiled4"}.get(rmap, "")
    traffic = None
    tf = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get(f"{name}_{kprec}")
    ibytes, ims = kern["k_col"][0] + kern["k_row"][0], kst["iter_kernel_ms"]
    log(f"kernels: k_col {kern['k_col'][1] * 1e3:.1f} us ({kern['k_col'][0] / kern['k_col'][1] / 1e6:.0f} GB/s), "
        f"k_row {kern['k_row'][1] * 1e3:.1f} us ({kern['k_row'][0] / kern['k_row'][1] / 1e6:.0f} GB/s), "
        f"iteration {ims * 1e3:.1f} us ({ibytes / ims / 1e6:.0f} GB/s)")
    return {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "algorithmic_bytes": dbytes,
            "launch_us": dms * 1e3, "xcd_remap": kst["col_xcd" if dom == "k_col" else "row_xcd"],
            "other_kernel": {k: {"launch_us": v[1] * 1e3, "algorithmic_bytes": v[0]} for k, v in kern.items()
                             if k != dom},
            "iteration": {"kernels": "k_col then k_row, in turn", "us": ims * 1e3, "algorithmic_bytes": ibytes,
                          "achieved": ibytes / (ims * 1e-3) / 1e9, "frac": ibytes / (ims * 1e-3) / 1e9 / PEAK_HBM_GBS}}


def plain_gpu(lp, costs, opts, sample):
    """Plain single-instance GPU PDHG (cold start, one case per run) on the
    first `sample` cases: certified solves per second."""
    t = time.perf_counter()
    ok, iters = 0, []
    for b in range(sample):
        r, _, _ = P.solve_batch(lp, costs[b:b + 1], opts)
        c = P.certify_batch(lp, costs[b:b + 1], r[0].x[None], r[0].y[None], np.array([r[0].objective]))
        ok += int(r[0].status == P.OPTIMAL and c[0, 3] > 0)
        iters.append(int(r[0].iterations))
        log(f"plain case {b}: {r[0].iterations} iterations, status {r[0].status}, {time.perf_counter() - t:.1f} s")
    el = time.perf_counter() - t
    return {"solves_per_s": ok / el if sample else None, "cases": sample, "certified": ok, "seconds": el,
            "iterations": iters}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--budget", type=float, default=330.0,
                    help="wall seconds for the whole run; warmup and timed steps stop when the next would not fit")
    ap.add_argument("--precision", choices=["fp32", "fp64", "fp32auto"], default="fp64")
    ap.add_argument("--row-map", type=int, default=0, choices=[0, 1, 2, 3, 4, 5],
                    help="row mapping of the iteration kernels: 0 autotuned, 1 buckets, 2 one mapping, "
                         "3, 4, 5 LDS tiles of 16, 8, 4 rows")
    ap.add_argument("--xcd-remap", type=int, default=-1, choices=[-1, 0, 1],
                    help="workgroup placement: -1 autotuned, 0 round-robin over XCDs, 1 contiguous ranges per XCD")
    ap.add_argument("--cases", type=int, default=64, help="family cases solved per step (config C3: 64)")
    ap.add_argument("--cold", action="store_true", help="no learned warm start (batched cold start only)")
    ap.add_argument("--train", type=int, default=64, help="family parameters solved to fit the warm start")
    ap.add_argument("--train-tol", type=float, default=1e-6, help="tolerance of the training solves")
    ap.add_argument("--train-steps", type=int, default=3000, help="Adam steps of the MLP")
    ap.add_argument("--warm-fallback", type=int, default=256_000,
                    help="iterations after which an unconverged warm-started case restarts from the cold start")
    ap.add_argument("--no-cold-batch", action="store_true", help="skip the cold-start batched comparison run")
    ap.add_argument("--no-packed", action="store_true", help="skip the packed run of 256 uneven LPs")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--plain-sample", type=int, default=1, help="instances solved one by one for the plain baseline")
    ap.add_argument("--hbm-periods", type=int, default=1200,
                    help="periods of the refinery LP for the SpMV rates beyond the Infinity Cache (0: skip)")
    ap.add_argument("--start", choices=["measured", "warm", "cold"], default="measured",
                    help="start of the timed steps: the faster of the warmup (learned warm start) and the cold "
                         "batch as measured in this run, or forced")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every iteration's kernels directly instead of replaying hipGraphs of 64 iterations")
    ap.add_argument("--max-iterations", type=int, default=1_000_000,
                    help="cap per instance (the reference default, src/pdhg/pdhg.h:34); unconverged = unsolved")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one device copy of MIB MiB (units of the HBM traffic counters)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="print the standalone SpMV rates on the refinery matrix (one JSON line) and exit")
    ap.add_argument("--large-lp", type=int, metavar="PERIODS",
                    help="kernel rates on a refinery LP of PERIODS periods (1200: beyond the Infinity Cache); "
                         "one JSON line, then exit")
    ap.add_argument("--plain-all", metavar="JSON",
                    help="solve every case of the family with plain single-instance GPU PDHG, write JSON, exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batched family shards by "
              "instance with no per-iteration communication, so run one process per GPU.", file=sys.stderr)
        sys.exit(2)
    info = P.gpu_info()
    if not info["available"]:
        raise SystemExit(f"no usable gfx950 device: {info}")
    if args.copy_calibration:
        # A plain device-to-device copy of a known size: its FETCH_SIZE /
        # WRITE_SIZE counter reading calibrates the units of the HBM-traffic
        # counters collected for the iteration kernels in the same process.
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        size = args.copy_calibration << 20
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(size)) == 0
        assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(size)) == 0
        assert hip.hipMemset(src, 1, ctypes.c_size_t(size)) == 0
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(size), 3) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(src)
        hip.hipFree(dst)
        log(f"calibration copy: {args.copy_calibration} MiB device to device")
    if args.spmv_only:
        print(json.dumps({"spmv": spmv_rates(P.gen_refinery(SEED_LP))}))
        return
    if args.large_lp:
        print(json.dumps(large_lp_rates(args.large_lp)))
        return
    prec = {"fp64": P.FP64, "fp32": P.FP32, "fp32auto": P.FP32_AUTO}[args.precision]
    lp = P.gen_refinery(SEED_LP)
    K = args.cases
    thetas, costs = P.gen_family(lp, SEED_THETA, K, 0.3)
    graph = 0 if args.no_graph else 1
    opts = P.default_options(tolerance=args.tol, precision=prec, max_iterations=args.max_iterations,
                             progress_seconds=15.0, row_map=args.row_map, xcd_remap=args.xcd_remap,
                             warm_fallback=args.warm_fallback, use_graph=graph)
    plain_opts = P.default_options(tolerance=args.tol, precision=prec, max_iterations=args.max_iterations,
                                   row_map=args.row_map, xcd_remap=args.xcd_remap, use_graph=graph)
    log(f"workload: refinery m={lp.m} n={lp.n} nnz={lp.nnz}, {K} instances, precision {args.precision}")
    if args.plain_all:
        pl = plain_gpu(lp, costs, plain_opts, K)
        pl.update(workload=f"refinery price family seed {SEED_LP}/{SEED_THETA}, {K} cases, tol {args.tol}, "
                           f"precision {args.precision}, one case per run (cold start)", gpu=info)
        with open(args.plain_all, "w") as f:
            json.dump(pl, f, indent=1)
        print(json.dumps({"plain_gpu_solves_per_s": pl["solves_per_s"], "cases": K, "file": args.plain_all}))
        return

    ws, fit_s = None, 0.0
    if not args.cold:
        t = time.perf_counter()
        ws = fit_warm_start(lp, args.train, args.train_tol, opts, args.train_steps)
        fit_s = time.perf_counter() - t
        log(f"warm-start fit: {fit_s:.1f} s ({args.train} parameters at tolerance {args.train_tol})")

    roof = kernel_roofline(lp, costs, prec, args.row_map, args.xcd_remap, args.tol, graph)
    spmv = spmv_rates(lp)
    log("spmv (op-level k_spmv, scaled refinery matrix): " + ", ".join(
        f"{k} A x {v['Ax']['GBs']:.0f} / A^T y {v['ATy']['GBs']:.0f} GB/s" for k, v in spmv.items()))
    plain = plain_gpu(lp, costs, plain_opts, args.plain_sample)
    log(f"plain single-instance GPU PDHG: {plain['solves_per_s']} solves/s over {plain['cases']} cases, "
        f"iterations {plain['iterations']}")
    plain_full = None
    pf = os.path.join(ROOT, "profiles", "plain_gpu_c3_fp64.json")
    if os.path.exists(pf) and prec == P.FP64 and K == 64:
        d = json.load(open(pf))
        plain_full = {k: d[k] for k in ("solves_per_s", "cases", "certified", "seconds")}
        plain_full["source"] = "profiles/plain_gpu_c3_fp64.json, recorded by bench.py --plain-all on this workload"

    hbm = None
    if args.hbm_periods > 0:
        hbm = spmv_beyond_cache(args.hbm_periods)

    # Warmup and timed steps inside the wall budget.  The warmup steps start
    # from the learned warm start; after them one run of the same family from
    # the cold start.  The timed steps then use whichever of the two certified
    # more solves per second of its run here (autotuning seconds excluded), or
    # the start forced by --start.
    step_s = []
    warm_done = 0
    reserve = 10.0  # the CPU baseline's and the report's own seconds come after the steps
    cpu_est = 0.0 if args.no_cpu else 2.2 * args.cpu_seconds
    warm_runs = []
    for w in range(args.warmup):
        if warm_done >= 1 and elapsed() + (1 + args.steps) * max(step_s) + cpu_est + reserve > args.budget:
            break  # one warmup at least; more only when every requested timed step still fits
        t = time.perf_counter()
        res, st, cert, ok = family_step(lp, thetas, costs, opts, ws)
        step_s.append(time.perf_counter() - t)
        warm_done += 1
        warm_runs.append({"seconds": step_s[-1], "tune_seconds": st["tune_seconds"], "certified": int(ok.sum()),
                          "iterations": st["iterations"], "cold_fallbacks": st["warm_fallbacks"],
                          "median_instance_iterations": float(np.median([r.iterations for r in res]))})
        log(f"warmup {w}: {step_s[-1]:.2f} s ({st['tune_seconds']:.2f} s autotuning), {st['iterations']} iterations, "
            f"{ok.sum()}/{K} certified, {st['warm_fallbacks']} cold fallbacks")

    # The reference's CPU PDHG on the same family, extrapolated to the
    # iterations of the slowest case of a step (from the warmup when there is one).
    cpu = None

    def run_cpu(case_iters):
        c = cpu_baseline(lp, costs, case_iters, args.cpu_seconds, args.cpu_threads)
        if c:
            log(f"cpu baseline: {c['value']:.4f} solves/s ({c['sample']})")
        return c

    if warm_done and not args.no_cpu:
        cpu = run_cpu(int(min(max(r.iterations for r in res), args.max_iterations)))
        cpu_est = 0.0

    # The same batched solve from a cold start (one run), before the timed
    # steps so that the budget cannot drop it.
    cold = None
    if ws is not None and not args.no_cold_batch:
        t = time.perf_counter()
        res_c, st_c, _, ok_c = family_step(lp, thetas, costs, opts, None)
        el = time.perf_counter() - t
        cold = {"solves_per_s": int(ok_c.sum()) / el, "seconds": el, "tune_seconds": st_c["tune_seconds"],
                "iterations": st_c["iterations"], "certified": int(ok_c.sum()),
                "median_instance_iterations": float(np.median([r.iterations for r in res_c]))}
        log(f"batched cold start: {el:.2f} s, {st_c['iterations']} iterations, {ok_c.sum()}/{K} certified")

    def rate(r):
        return r["certified"] / max(r["seconds"] - r["tune_seconds"], 1e-9)

    step_ws, choice = ws, None
    if ws is not None and cold is not None and warm_runs:
        warm_rate = max(rate(r) for r in warm_runs)
        cold_rate = rate(cold)
        use_warm = args.start == "warm" or (args.start == "measured" and warm_rate >= cold_rate)
        step_ws = ws if use_warm else None
        choice = {"rule": args.start, "warm_certified_per_s": warm_rate, "cold_certified_per_s": cold_rate,
                  "timed_start": "learned warm start" if use_warm else "cold"}
        log(f"start of the timed steps: {choice['timed_start']} (warm {warm_rate:.3f}, cold {cold_rate:.3f} "
            f"certified solves/s, rule {args.start})")
        if not use_warm:
            step_s = [cold["seconds"]]  # the budget's step estimate follows the chosen start

    certified, converged, iters, kkt, gaps, inst_iters, fallbacks = 0, 0, [], [], [], [], 0
    capped_kkt = []
    ref_check = None
    up_bytes = down_bytes = 0
    timed = []
    for s in range(args.steps):
        est = max(step_s) if step_s else 0.0
        if s >= 1 and elapsed() + est + cpu_est + reserve > args.budget:
            break
        ts = time.perf_counter()
        res, st, cert, ok = family_step(lp, thetas, costs, opts, step_ws)
        timed.append(time.perf_counter() - ts)
        step_s.append(timed[-1])
        up_bytes += st["bytes_uploaded"]
        down_bytes += st["bytes_downloaded"]
        certified += int(ok.sum())
        converged += sum(1 for r in res if r.status == P.OPTIMAL)
        capped_kkt += [r.relative_kkt for r in res if r.status != P.OPTIMAL]
        iters.append(st["iterations"])
        fallbacks += st["warm_fallbacks"]
        inst_iters += [r.iterations for r in res]
        kkt.append(max([r.relative_kkt for r in res if r.status == P.OPTIMAL] or [0.0]))
        ref_check = reference_check(res, cert, args.tol, K)
        gaps.append(float(cert[:, 1].max()))
        log(f"step {s}: {timed[-1]:.2f} s, {st['iterations']} iterations, {ok.sum()}/{K} certified, "
            f"{st['warm_fallbacks']} cold fallbacks, max rel-KKT of the converged {kkt[-1]:.2e}, "
            f"max certified gap {gaps[-1]:.2e}")
    steps = len(timed)
    elapsed_timed = sum(timed)
    value = certified / elapsed_timed
    gpu_iters = int(max(iters))  # iterations of the run
    case_iters = int(min(max(inst_iters), args.max_iterations))  # slowest case, capped

    if cpu is None and not args.no_cpu:
        cpu = run_cpu(case_iters)

    packed = None
    if not args.no_packed and elapsed() + 150.0 + reserve <= args.budget:
        # Config C4's uneven batch: 256 unrelated LPs (the config-0 generator,
        # nnz log-uniform in [2e3, 2e5], a non-binding box x <= 2) in one
        # packed run of persistent kernels, every answer through the certifier.
        lps = P.gen_uneven_batch(SEED_PACKED, 256)
        t = time.perf_counter()
        res_p, st_p = P.solve_packed(lps, P.default_options(tolerance=args.tol, precision=prec,
                                                            max_iterations=args.max_iterations))
        conv_p = sum(1 for r in res_p if r.status == P.OPTIMAL)
        cert_p = sum(1 for q, r in zip(lps, res_p)
                     if r.status == P.OPTIMAL and P.certify(q, r.x, r.y, r.objective)["accepted"])
        el = time.perf_counter() - t
        packed = {"lps": len(lps), "nnz": int(sum(q.nnz for q in lps)), "seconds": el, "converged": conv_p,
                  "certified": cert_p, "certified_per_s": cert_p / el,
                  "max_instance_iterations": int(max(r.iterations for r in res_p))}
        log(f"packed uneven batch: {conv_p}/{len(lps)} at rel-KKT <= {args.tol:g}, {cert_p} certified, in {el:.2f} s")

    mode = "batched+warm" if step_ws is not None else "batched"
    if ws is not None and step_ws is None:
        mode = "batched, cold start measured faster than the learned warm start"
    tol_label = ("%g" % args.tol).replace("e-0", "e-")  # 1e-06 -> 1e-6
    plain_rate = plain_full["solves_per_s"] if plain_full else plain["solves_per_s"]
    out = {
        "metric": f"LP solves/s at rel-KKT<={tol_label} ({mode}, 64-case refinery price family, certified)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm_done,
        "ms_per_step": elapsed_timed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {P.FP64: "f64", P.FP32: "f32", P.FP32_AUTO: "f32+f64"}[prec],
        "data": "synthetic",
        "config": {"workload": f"config C3: {K}-case price family of the refinery LP (m={lp.m} n={lp.n} "
                               f"nnz={lp.nnz}) in one batched run of width 64, tol {args.tol}",
                   "cases": K, "batch_width": 64, "start": "learned warm start" if step_ws is not None else "cold",
                   "warm_fallback_iterations": args.warm_fallback if step_ws is not None else None,
                   "warm_start_training": (f"{args.train} family parameters at rel-KKT <= {args.train_tol:g}"
                                           if ws is not None else None),
                   "iterations": gpu_iters, "median_instance_iterations": float(np.median(inst_iters)),
                   "cases_run": steps * K, "converged": converged, "certified": certified,
                   "iteration_limit": steps * K - converged, "cold_fallbacks": fallbacks},
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "budget_s": args.budget,
        "stopped_by_budget": steps < args.steps or warm_done < args.warmup,
        "roofline": roof,
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
        "spmv": spmv,
        "spmv_beyond_infinity_cache": hbm,
        "bytes_per_step": {"uploaded": up_bytes / steps, "downloaded": down_bytes / steps},
        "plain_gpu_solves_per_s": plain_rate,
        "plain_gpu_sample": plain,
        "plain_gpu_all_cases": plain_full,
        "vs_plain_gpu": (value / plain_rate) if plain_rate else None,
        "batched_warm": warm_runs[-1] if warm_runs and ws is not None else None,
        "batched_cold": cold,
        "start_choice": choice,
        "packed_uneven": packed,
        "warm_start_fit_seconds": fit_s if ws is not None else None,
        "max_rel_kkt": max(kkt),
        "rel_kkt_at_iteration_limit": capped_kkt,
        "max_certified_gap": max(gaps),
        "reference_check": ref_check,
        "wall_s": elapsed(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
