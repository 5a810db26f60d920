"""Benchmark of the MI355X two-stage symmetric eigensolver.

A step = one full FP32 eigen-decomposition (eigenvalues + eigenvectors:
SBR -> bulge chasing -> GPU divide and conquer -> back transformation) of the
headline workload: n = 16384, b = 32, nb = 1024, A = (X + X^T)/2 with X
i.i.d. uniform[-1,1] (Philox4x32-10, seed 0), synthetic and resident in HBM.
Each step first restores A from a pristine copy (the solver destroys its
input); that copy is inside the timed region.

value = EVD TFLOP/s = (10/3 n^3) / seconds per EVD (the one-stage textbook
count 4/3 n^3 + 2 n^3, D&C excluded), higher is better; ms_per_step is the
seconds per EVD.  The line also carries per-stage times, the roofline of the
dominant kernel (measured live with hipEvents on the solver's stream),
accuracy (backward error / orthogonality in FP64 on the GPU) and a host
LAPACK dsyevd baseline on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32_mfma_tflops": 157.3, "fp64_tflops": 78.6, "hbm_gbs": 8000.0}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "bcback_chase_n16384_hbm_traffic.md")


def profiled_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` at the headline size, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated)."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("| `") and kernel in line.split("`")[1]:
                return float(line.strip().strip("|").split("|")[-1]) * 1e9
    except OSError:
        pass
    return None


def golden_eigenvalues(n: int, seed: int):
    """Host LAPACK dsyevd (FP64) eigenvalues of the uniform workload at
    (n, seed), committed as a fixture (tests/golden/gen_golden_headline.py),
    with ||A||_2; None when no fixture exists for this order."""
    import numpy as np

    path = os.path.join(ROOT, "tests", "golden", f"headline_n{n}_s{seed}.npz")
    if not os.path.exists(path):
        return None
    g = np.load(path)
    return g["w"], float(g["norm2"])


def accuracy(E, ctx, A0, r, gold):
    """backward error, orthogonality (FP64 on the GPU) and, when the golden
    dsyevd eigenvalues exist, max |w - w_dsyevd| / ||A||_2."""
    out = {"backward_error": E.backward_error(ctx, A0, r.w, r.Z), "orthogonality": E.checkOrthogonality(ctx, r.Z)}
    if gold is not None:
        import numpy as np

        w0, nrm2 = gold
        out["eig_err"] = float(np.abs(r.w.cpu().numpy() - w0).max() / nrm2)
        out["eig_err_ref"] = "host LAPACK dsyevd FP64 on the same matrix, / ||A||_2"
    return out


def cpu_baseline(n: int = 4096, seed: int = 0):
    """Host LAPACK dsyevd (FP64, jobz=V) on an n x n sample of the same
    workload family, timed on this machine's cores."""
    import numpy as np
    import scipy.linalg as sl

    from oracle import evd_oracle as O

    A = O.gen_symmetric(n, seed, 0).astype(np.float64)
    # The threads LAPACK actually runs on: the BLAS pool's own size (capped by
    # OMP_NUM_THREADS / OPENBLAS_NUM_THREADS), not the machine's core count.
    threads = 0
    try:
        from threadpoolctl import threadpool_info

        threads = max((int(i.get("num_threads") or 0) for i in threadpool_info()
                       if i.get("user_api") == "blas"), default=0)
    except Exception:
        threads = 0
    if threads <= 0:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    sl.eigh(A, driver="evd")
    dt = time.perf_counter() - t0
    return {
        "value": (10.0 / 3.0) * n ** 3 / dt / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"host LAPACK dsyevd (scipy/OpenBLAS, FP64, jobz=V) at n={n}, uniform seed {seed}: {dt:.2f} s "
                  f"(same 10/3 n^3 normalisation); the reference's own CPU stage is MKL dstedc",
        "seconds": dt,
        "n": n,
    }


def serving_bench(E, torch, nreq: int = 60, seed: int = 1, budget: int = 200 << 30):
    """Serving form on a bounded sample: nreq requests, n drawn from
    {256: 50%, 1024: 30%, 4096: 15%, 16384: 5%}, 70% with eigenvectors,
    Gaussian-entry symmetric matrices resident in HBM.  Saturated throughput
    of one-at-a-time plain solves (one slot, no batching) and of the serving
    engine (4 concurrent slots, same-size batching), then the same Poisson
    arrival schedule (70% of the one-at-a-time capacity) on both; the engine
    also in the mixed-precision (split FP16) form."""
    import numpy as np

    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.choice([256, 1024, 4096, 16384], size=nreq, p=[0.5, 0.3, 0.15, 0.05])]
    jobz = ["V" if x else "N" for x in (rng.random(nreq) < 0.7)]
    gctx = E.Context(0)
    pristine = [E.gen_symmetric(gctx, n, 1000 + i, 1) for i, n in enumerate(sizes)]
    torch.cuda.synchronize()

    def run(slots, arrivals, batching, precision=E.PREC_FP32):
        work = []
        for p in pristine:
            w = E.colmajor(p.shape[0], p.shape[0])
            w.copy_(p)
            work.append(w)
        torch.cuda.synchronize()
        srv = E.Server(0, slots=slots, hbm_budget=budget, batching=batching, precision=precision)
        t0 = time.perf_counter()
        for i in np.argsort(arrivals, kind="stable"):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            n = sizes[i]
            srv.submit(work[i], b=32, nb=1024 if n >= 4096 else 256, jobz=jobz[i])
        srv.wait_all()
        st = srv.stats()
        srv.close()
        del work
        return {k: st[k] for k in ("solves_per_s", "p50_ms", "p99_ms")} | {"arena_gib": st["arena_bytes"] / 2**30}

    zero = np.zeros(nreq)
    sat1 = run(1, zero, False)
    sat4 = run(4, zero, True)
    rate = 0.7 * sat1["solves_per_s"]
    arr = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    poi1 = run(1, arr, False)
    poi4 = run(4, arr, True)
    sat4h = run(4, zero, True, E.PREC_FP16)
    poi4h = run(4, arr, True, E.PREC_FP16)
    return {"requests": nreq, "mix": "n in {256: 50%, 1024: 30%, 4096: 15%, 16384: 5%}, 70% jobz=V, Gaussian "
                                     f"entries, seed {seed}, plain FP32, HBM budget {budget >> 30} GiB",
            "saturated": {"one_at_a_time": sat1, "engine_4slots_batching": sat4,
                          "engine_4slots_batching_fp16": sat4h,
                          "throughput_ratio": sat4["solves_per_s"] / sat1["solves_per_s"],
                          "throughput_ratio_fp16": sat4h["solves_per_s"] / sat1["solves_per_s"]},
            "poisson": {"rate_per_s": rate, "one_at_a_time": poi1, "engine_4slots_batching": poi4,
                        "engine_4slots_batching_fp16": poi4h}}


def captured_bench(E, torch, reps: int = 20):
    """Small solves launched kernel by kernel from the host (evd_syevd_f32,
    evd_syevd_batched_f32) against the same solves replayed from a captured
    hipGraph (evd_graph_*): mean wall time per call over `reps` sequential
    calls, each restoring its input and waiting for its result."""
    out = {}
    ctx = E.Context(0)
    gctx = E.Context(0)
    try:
        for n in (256, 1024):
            M = E.gen_symmetric(ctx, n, 7, 0)
            A = E.colmajor(n, n)
            for _ in range(3):
                A.copy_(M)
                E.syevd(ctx, A, b=32, nb=256, overwrite=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                A.copy_(M)
                E.syevd(ctx, A, b=32, nb=256, overwrite=True)
            torch.cuda.synchronize()
            eager = (time.perf_counter() - t0) / reps
            g = E.SolveGraph(gctx, M, b=32, nb=256)
            try:
                for _ in range(3):
                    g.solve(M)
                t0 = time.perf_counter()
                for _ in range(reps):
                    g.solve(M)
                replay = (time.perf_counter() - t0) / reps
            finally:
                g.close()
            out[f"n{n}_jobzV"] = {"eager_ms": eager * 1e3, "graph_ms": replay * 1e3, "speedup": eager / replay}
        count, n = 32, 256
        mats = [E.gen_symmetric(ctx, n, 100 + q, 0) for q in range(count)]
        for _ in range(2):
            E.syevd_batched(ctx, mats, b=32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps // 4):
            E.syevd_batched(ctx, mats, b=32)
        torch.cuda.synchronize()
        eager = (time.perf_counter() - t0) / (reps // 4)
        g = E.BatchedSolveGraph(gctx, mats, b=32)
        try:
            for _ in range(2):
                g.solve(mats)
            t0 = time.perf_counter()
            for _ in range(reps // 4):
                g.solve(mats)
            replay = (time.perf_counter() - t0) / (reps // 4)
        finally:
            g.close()
        out[f"batch{count}_n{n}_jobzV"] = {"eager_ms": eager * 1e3, "graph_ms": replay * 1e3,
                                           "speedup": eager / replay,
                                           "solves_per_s_graph": count / replay}
    finally:
        gctx.close()
        ctx.close()
    return out


def batched_form_bench(E, torch, count: int = 512, seed: int = 1):
    """Batched form: `count` symmetric FP32 matrices, n log-uniform in
    [64, 2048] (seed 1), uniform[-1,1] entries symmetrised, half with
    eigenvectors: one evd_syevd_mixed_f32 call (size buckets, one batched
    solve per bucket and jobz) against the same problems solved one at a time
    with evd_syevd_f32; both with inputs resident in HBM."""
    import numpy as np

    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in np.exp(rng.uniform(np.log(64), np.log(2048), size=count)).round()]
    jobz = ["V" if i % 2 == 0 else "N" for i in range(count)]
    ctx = E.Context(0)
    try:
        mats = [E.gen_symmetric(ctx, n, 5000 + i, 0) for i, n in enumerate(sizes)]
        E.syevd_mixed(ctx, mats[:16], jobz[:16])  # warm-up: arena, plans
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E.syevd_mixed(ctx, mats, jobz)
        torch.cuda.synchronize()
        t_mixed = time.perf_counter() - t0
        t0 = time.perf_counter()
        for M, j in zip(mats, jobz):
            E.syevd(ctx, M, b=32, nb=256 if M.shape[0] < 4096 else 1024, jobz=j)
        torch.cuda.synchronize()
        t_single = time.perf_counter() - t0
    finally:
        ctx.close()
    return {"matrices": count, "sizes": f"log-uniform [64, 2048], seed {seed}, half jobz=V",
            "bucketed_s": t_mixed, "bucketed_solves_per_s": count / t_mixed,
            "one_at_a_time_s": t_single, "one_at_a_time_solves_per_s": count / t_single,
            "speedup": t_single / t_mixed}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--b", type=int, default=32)
    ap.add_argument("--nb", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision (FP16/BF16 split) runs")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop sample")
    ap.add_argument("--no-captured", action="store_true", help="skip the eager-vs-graph small-solve sample")
    ap.add_argument("--no-batched", action="store_true", help="skip the mixed-size batched sample")
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--cpu-n", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the eigensolver runs on a single MI355X and has no "
              "multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import evd4sc2025_amd as E

    n, b, nb = args.n, args.b, args.nb
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    ctx = E.Context(0, stream)
    ctx.reserve(E.syevd_workspace(n, b, nb, "V"))
    ctx.enable_timing(True)
    A0 = E.gen_symmetric(ctx, n, 0, 0)
    Aw = E.colmajor(n, n)
    torch.cuda.synchronize()

    def step():
        Aw.copy_(A0)
        return E.syevd(ctx, Aw, b=b, nb=nb, jobz="V", overwrite=True)

    for i in range(args.warmup):
        t = time.perf_counter()
        step()
        torch.cuda.synchronize()
        print(f"[bench] warmup {i}: {time.perf_counter() - t:.3f} s", file=sys.stderr, flush=True)
    stage_tot = {k: 0.0 for k in E.TIMERS}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step()
        for k, v in r.stage_ms.items():
            stage_tot[k] += v
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    stage_ms = {k: v / args.steps for k, v in stage_tot.items()}
    flops = (10.0 / 3.0) * n ** 3
    value = flops / dt / 1e12

    # dominant kernel: the BC back transformation, WY blocks of 32 sweeps on
    # FP32 MFMA (bcw_kernel, timed alone with hipEvents on the solver's
    # stream); 2 n^3 algorithmic flops (4 n per reflector entry) - the MFMA
    # issue 1.625x that, the staircase blocks' nonzero 16 x 16 tiles in full
    bc_ms = stage_ms["bc_back_kernel"]
    bc_flops = 2.0 * n ** 3
    achieved = bc_flops / (bc_ms * 1e-3) / 1e12
    traffic = profiled_traffic("bcw_kernel") if (n, b) == (16384, 32) else None
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK["fp32_mfma_tflops"], "unit": "TFLOP/s",
                "frac": achieved / PEAK["fp32_mfma_tflops"], "traffic": traffic, "kernel": "bcw_kernel",
                "note": "2 n^3 algorithmic flops per launch over the hipEvent time of the launch (the WY form issues "
                        "1.625x that on v_mfma_f32_16x16x4f32); FP32 MFMA peak 157.3 TF; traffic = copy-calibrated "
                        "FETCH_SIZE + WRITE_SIZE bytes per launch from the committed rocprofv3 summary"}
    # the bulge chase against HBM: (1.5 b + 1) n^2 FP64 window bytes moved per launch
    bc_bytes = (1.5 * b + 1) * n * n * 8
    chase = {"bound": "hbm", "achieved": bc_bytes / (stage_ms["bc"] * 1e-3) / 1e9 if stage_ms["bc"] else None,
             "peak": PEAK["hbm_gbs"], "unit": "GB/s", "kernel": "chase_kernel",
             "us_per_dependent_step": stage_ms["bc"] * 1e3 / (2.0 * n) if stage_ms["bc"] else None,
             "traffic": profiled_traffic("chase_kernel") if (n, b) == (16384, 32) else None}
    if chase["achieved"] is not None:
        chase["frac"] = chase["achieved"] / PEAK["hbm_gbs"]
    stage_tflops = {
        "sbr": (4.0 / 3.0) * n ** 3 / (stage_ms["sbr"] * 1e-3) / 1e12 if stage_ms["sbr"] else None,
        "bc_back": achieved,
        "sbr_back": 2.0 * n ** 3 / (stage_ms["sbr_back"] * 1e-3) / 1e12 if stage_ms["sbr_back"] else None,
    }

    acc = None
    gold = golden_eigenvalues(n, 0)
    if not args.no_accuracy:
        Aw.copy_(A0)
        r = E.syevd(ctx, Aw, b=b, nb=nb, jobz="V", overwrite=True)
        acc = accuracy(E, ctx, A0, r, gold)

    print(f"[bench] timed {args.steps} steps: {dt:.3f} s/EVD, stages {stage_ms}, accuracy {acc}", file=sys.stderr,
          flush=True)

    # mixed-precision forms of the same workload (split 16-bit MFMA products in
    # SBR and the back transformation; FP64 reflectors and tridiagonal stage)
    mixed = {}
    if not args.no_mixed:
        for name, prec in (("fp16", E.PREC_FP16), ("bf16", E.PREC_BF16)):
            ctx.set_precision(prec)
            step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mst = {k: 0.0 for k in E.TIMERS}
            for _ in range(args.steps):
                r = step()
                for k, v in r.stage_ms.items():
                    mst[k] += v
            torch.cuda.synchronize()
            mdt = (time.perf_counter() - t0) / args.steps
            entry = {"ms_per_step": mdt * 1e3, "tflops": flops / mdt / 1e12, "speedup_vs_fp32": dt / mdt,
                     "stage_ms": {k: v / args.steps for k, v in mst.items()}}
            if not args.no_accuracy:
                Aw.copy_(A0)
                r = E.syevd(ctx, Aw, b=b, nb=nb, jobz="V", overwrite=True)
                entry["accuracy"] = accuracy(E, ctx, A0, r, gold)
            mixed[name] = entry
            print(f"[bench] mixed {name}: {mdt:.3f} s/EVD ({dt / mdt:.2f}x), {entry.get('accuracy')}", file=sys.stderr,
                  flush=True)
        ctx.set_precision(E.PREC_FP32)
    serving = None
    if not args.no_serve:
        serving = serving_bench(E, torch, args.serve_requests)
        print(f"[bench] serving: {serving}", file=sys.stderr, flush=True)
    captured = None
    if not args.no_captured:
        captured = captured_bench(E, torch)
        print(f"[bench] captured: {captured}", file=sys.stderr, flush=True)
    batched = None
    if not args.no_batched:
        batched = batched_form_bench(E, torch)
        print(f"[bench] batched form: {batched}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n)
    line = {
        "metric": "EVD s & TFLOP/s at n=16384 FP32 b=32; mixed-size solves/s + p99 latency",
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"full symmetric EVD (jobz=V) n={n} b={b} nb={nb}, uniform[-1,1] symmetrised, seed 0",
                   "n": n, "b": b, "nb": nb, "precision": "plain FP32 (FP64 reflectors, band chase and D&C)"},
        "seconds_per_evd": dt,
        "stage_ms": stage_ms,
        "stage_tflops": stage_tflops,
        "roofline": roofline,
        "chase_roofline": chase,
        "cpu_baseline": cpu,
        "accuracy": acc,
        "mixed_precision": mixed or None,
        "serving": serving,
        "captured": captured,
        "batched_form": batched,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
