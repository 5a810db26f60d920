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


def _mkl():
    """The image's MKL runtime (the library the reference links for its host
    stage), or None."""
    import ctypes

    for path in ("/opt/conda/lib/libmkl_rt.so.1", "libmkl_rt.so.1", "libmkl_rt.so"):
        try:
            lib = ctypes.CDLL(path)
        except OSError:
            continue
        if hasattr(lib, "LAPACKE_dstedc") and hasattr(lib, "LAPACKE_dsyevd"):
            return lib
    return None


def host_threads() -> int:
    """Threads the host LAPACK runs on: OMP_NUM_THREADS when set (the CPU
    share the node gives this process), else the machine's core count."""
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)


def cpu_baseline(A_host, D, Ed, full: bool):
    """Host LAPACK on the headline workload, on this machine's cores: MKL
    LAPACKE_dstedc on the (D, E) the GPU chase produced for the headline
    matrix (the reference's own host stage, src/EVD/myEVD.cu:529) and, with
    full=True, LAPACK dsyevd (jobz=V) on the headline matrix itself; without
    it, dsyevd on the n = 8192 matrix of the same workload family (uniform,
    seed 0) as a bounded sample.  dsyevd comes from the SciPy build (OpenBLAS),
    with MKL as the fallback; dstedc from MKL, with the SciPy build as the
    fallback (MKL is optional: without it only the SciPy legs run); each
    failure is recorded.  value = (10/3 n^3) / seconds of
    the dsyevd run, the metric's normalisation."""
    import ctypes

    import numpy as np

    mkl = _mkl()
    threads = host_threads()
    if mkl is not None:
        mkl.MKL_Set_Num_Threads(ctypes.c_int(threads))
    dp = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    n = D.shape[0]
    if full:
        A = np.array(A_host, dtype=np.float64, order="F")
        nn = n
    else:
        from oracle import evd_oracle as O

        nn = 8192
        A = np.asfortranarray(O.gen_symmetric(nn, 0, 0).astype(np.float64))
    # the dsyevd sample: LAPACK from the SciPy build (OpenBLAS) first; MKL's
    # dsyevd was 40x slower on the host CPU of an MI355X node and then failed there
    # (a divide-and-conquer leaf, info > 0), so it is only the fallback
    syevd_lib, syevd_err = "SciPy LAPACK dsyevd (OpenBLAS)", {}
    try:
        from scipy.linalg import lapack

        t0 = time.perf_counter()
        _, _, info = lapack.dsyevd(A, compute_v=1, lower=1, overwrite_a=0)
        t_syevd = time.perf_counter() - t0
        if info != 0:
            syevd_err = {"host_dsyevd_scipy_error": f"SciPy dsyevd returned {info} at n={nn}"}
    except Exception as ex:  # noqa: BLE001 - reported in the line
        syevd_err = {"host_dsyevd_scipy_error": repr(ex)}
    if syevd_err:
        if mkl is None:
            return {"value": None, "error": f"{syevd_err}; no MKL runtime (libmkl_rt) on this host"}
        w = np.empty(nn)
        t0 = time.perf_counter()
        rc = mkl.LAPACKE_dsyevd(102, ctypes.c_char(b"V"), ctypes.c_char(b"L"), nn, dp(A), nn, dp(w))
        t_syevd = time.perf_counter() - t0
        if rc != 0:
            return {"value": None, "error": f"{syevd_err}; LAPACKE_dsyevd returned {rc} at n={nn}"}
        syevd_lib = "MKL LAPACKE_dsyevd"
    del A
    # the reference's own host stage on the headline (D, E): MKL dstedc; if
    # MKL reports a failure on this host, LAPACK dsbevd on the same matrix
    # stored as a band of width 1 (its tridiagonal stage is dstedc) from the
    # SciPy build, and both facts are recorded
    stedc = {"host_dstedc_n": n}
    d = np.array(D, dtype=np.float64)
    e = np.array(Ed, dtype=np.float64)
    Z = np.empty((n, n), dtype=np.float64, order="F")
    t0 = time.perf_counter()
    rc = mkl.LAPACKE_dstedc(102, ctypes.c_char(b"I"), n, dp(d), dp(e), dp(Z), n) if mkl is not None else None
    t_stedc = time.perf_counter() - t0
    del Z
    if rc == 0:
        stedc.update({"host_dstedc_s": t_stedc, "host_dstedc_lib": "MKL LAPACKE_dstedc"})
    else:
        stedc["host_dstedc_mkl_error"] = (f"LAPACKE_dstedc returned {rc} at n={n} after {t_stedc:.2f} s"
                                          if mkl is not None else "no MKL runtime (libmkl_rt) on this host")
        try:
            from scipy.linalg import lapack

            ab = np.zeros((2, n))
            ab[1] = d0 = np.array(D, dtype=np.float64)
            ab[0, 1:] = np.array(Ed, dtype=np.float64)[: n - 1]
            del d0
            t0 = time.perf_counter()
            _, _, info = lapack.dsbevd(ab, compute_v=1, lower=0)
            t_stedc = time.perf_counter() - t0
            if info == 0:
                stedc.update({"host_dstedc_s": t_stedc,
                              "host_dstedc_lib": "SciPy LAPACK dsbevd, band width 1 (dstedc inside; OpenBLAS)"})
            else:
                stedc["host_dstedc_error"] = f"dsbevd returned {info}"
        except Exception as ex:  # noqa: BLE001 - reported in the line
            stedc["host_dstedc_error"] = repr(ex)
    which = "the headline matrix" if full else f"the n={nn} matrix of the same workload (uniform[-1,1], seed 0)"
    st = (f"{stedc['host_dstedc_lib']} on the headline (D, E) - the reference's own host stage - "
          f"{stedc['host_dstedc_s']:.2f} s") if "host_dstedc_s" in stedc else "host dstedc: failed (see fields)"
    out = {
        "value": (10.0 / 3.0) * nn ** 3 / t_syevd / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{syevd_lib} (FP64, jobz=V) on {which}: {t_syevd:.2f} s (10/3 n^3 normalisation); {st}",
        "dsyevd_s": t_syevd,
        "dsyevd_n": nn,
        "dsyevd_lib": syevd_lib,
    }
    out.update(syevd_err)
    out.update(stedc)
    return out


def serving_bench(E, torch, nreq: int = 400, seed: int = 1, budget: int = 200 << 30, slots: int = 4):
    """Serving form on a bounded sample of the request stream: nreq requests
    in the exact proportions n = 256 (50%), 1024 (30%), 4096 (15%), 16384
    (5%), 70% with eigenvectors, Gaussian-entry symmetric matrices resident in
    HBM, against an HBM budget pool.  Runs: one-at-a-time plain FP32 solves
    (one slot, no batching) saturated; the engine (`slots` slots, same-size
    batching, largest-order-first admission) saturated in plain FP32 and in
    the split-FP16 form; then one Poisson schedule at 70% of the
    one-at-a-time capacity for one-at-a-time solves, and the same schedule at
    twice that rate for the engine: the target is 2x the solves/s of
    one-at-a-time plain solves at no worse p99."""
    import numpy as np

    rng = np.random.default_rng(seed)
    counts = {256: nreq // 2, 1024: nreq * 3 // 10, 4096: nreq * 3 // 20}
    counts[16384] = nreq - sum(counts.values())
    sizes = [n for n, c in counts.items() for _ in range(c)]
    rng.shuffle(sizes)
    nv = int(round(0.7 * nreq))
    jobz = np.array(["V"] * nv + ["N"] * (nreq - nv))
    rng.shuffle(jobz)
    gctx = E.Context(0)

    def run(nslots, arrivals, batching, precision=E.PREC_FP32, policy=0):
        # inputs regenerated from their seeds (the solver overwrites them)
        work = [E.gen_symmetric(gctx, n, 1000 + i, 1) for i, n in enumerate(sizes)]
        torch.cuda.synchronize()
        srv = E.Server(0, slots=nslots, hbm_budget=budget, batching=batching, precision=precision, policy=policy)
        t0 = time.perf_counter()
        for i in np.argsort(arrivals, kind="stable"):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            n = sizes[i]
            srv.submit(work[i], b=32, nb=1024 if n >= 4096 else 256, jobz=str(jobz[i]))
        srv.wait_all()
        st = srv.stats()
        srv.close()
        del work
        torch.cuda.empty_cache()
        return {k: st[k] for k in ("solves_per_s", "p50_ms", "p99_ms")} | {"arena_gib": st["arena_bytes"] / 2**30}

    zero = np.zeros(nreq)
    sat1 = run(1, zero, False, policy=1)
    print(f"[bench] serving one-at-a-time saturated: {sat1}", file=sys.stderr, flush=True)
    sat_plain = run(slots, zero, True)
    sat_fp16 = run(slots, zero, True, E.PREC_FP16)
    print(f"[bench] serving engine saturated: plain {sat_plain}, fp16 {sat_fp16}", file=sys.stderr, flush=True)
    rate = 0.7 * sat1["solves_per_s"]
    arr = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    poi1 = run(1, arr, False, policy=1)
    poi_fp16 = run(slots, arr / 2.0, True, E.PREC_FP16)
    poi_plain = run(slots, arr / 2.0, True)
    # the engine's throughput at the one-at-a-time p99: the same arrival
    # sequence compressed to 1.0x and 1.4x the one-at-a-time rate; the
    # highest of these rates (with the 2x run) whose p99 is no worse
    sweep = {2.0: poi_plain}
    for f in (1.0, 1.4):
        sweep[f] = run(slots, arr / f, True)
        print(f"[bench] serving engine at {f}x rate: {sweep[f]}", file=sys.stderr, flush=True)
    ok = [f for f in sorted(sweep) if sweep[f]["p99_ms"] <= poi1["p99_ms"]]
    matched = None
    if ok:
        f = ok[-1]
        matched = {"rate_factor": f, "rate_per_s": f * rate} | sweep[f] | {
            "throughput_ratio": sweep[f]["solves_per_s"] / poi1["solves_per_s"]}
    gctx.close()
    return {"requests": nreq, "mix": f"exact counts {counts}, 70% jobz=V, Gaussian entries, seed {seed}, "
                                      f"HBM budget {budget >> 30} GiB (one pool, lent per request)",
            "saturated": {"one_at_a_time": sat1, f"engine_{slots}slots_plain": sat_plain,
                          f"engine_{slots}slots_fp16": sat_fp16,
                          "throughput_ratio_plain": sat_plain["solves_per_s"] / sat1["solves_per_s"],
                          "throughput_ratio_fp16": sat_fp16["solves_per_s"] / sat1["solves_per_s"]},
            "poisson": {"one_at_a_time": {"rate_per_s": rate} | poi1,
                        f"engine_{slots}slots_fp16_at_2x_rate": {"rate_per_s": 2 * rate} | poi_fp16,
                        f"engine_{slots}slots_plain_at_2x_rate": {"rate_per_s": 2 * rate} | poi_plain,
                        "throughput_ratio_fp16": poi_fp16["solves_per_s"] / poi1["solves_per_s"],
                        "p99_ratio_fp16": poi_fp16["p99_ms"] / poi1["p99_ms"],
                        "throughput_ratio_plain": poi_plain["solves_per_s"] / poi1["solves_per_s"],
                        "p99_ratio_plain": poi_plain["p99_ms"] / poi1["p99_ms"],
                        f"engine_{slots}slots_plain_rate_sweep": {f"{f}x": v for f, v in sorted(sweep.items())},
                        "engine_plain_at_matched_p99": matched}}


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
    solve per bucket and jobz, bandwidth min(32, N / 4) per bucket order N),
    the same call captured as one hipGraph (evd_graph_create_mixed, mean of 3
    replays), and the same problems solved one at a time with evd_syevd_f32
    (b = 32); all with inputs resident in HBM."""
    import numpy as np

    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in np.exp(rng.uniform(np.log(64), np.log(2048), size=count)).round()]
    jobz = ["V" if i % 2 == 0 else "N" for i in range(count)]
    ctx = E.Context(0)
    try:
        mats = [E.gen_symmetric(ctx, n, 5000 + i, 0) for i, n in enumerate(sizes)]
        E.syevd_mixed(ctx, mats[:16], jobz[:16], b=0)  # warm-up: arena, plans
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eager = E.syevd_mixed(ctx, mats, jobz, b=0)
        torch.cuda.synchronize()
        t_mixed = time.perf_counter() - t0
        t0 = time.perf_counter()
        for M, j in zip(mats, jobz):
            E.syevd(ctx, M, b=32, nb=256 if M.shape[0] < 4096 else 1024, jobz=j)
        torch.cuda.synchronize()
        t_single = time.perf_counter() - t0
        # the same call captured once as a hipGraph (every bucket's padding,
        # batched solve and extraction) and replayed on the resident inputs
        g = E.MixedSolveGraph(ctx, mats, jobz=jobz, b=0)
        try:
            g.replay()
            reps = 3
            t0 = time.perf_counter()
            for _ in range(reps):
                g.launch()
            g.wait()
            t_graph = (time.perf_counter() - t0) / reps
            same = all(torch.equal(r.w, w) for r, w in zip(eager, g.w)) and all(
                r.Z is None or torch.equal(r.Z, Z) for r, Z in zip(eager, g.Z))
        finally:
            g.close()
    finally:
        ctx.close()
    return {"matrices": count, "sizes": f"log-uniform [64, 2048], seed {seed}, half jobz=V",
            "bucketed_s": t_mixed, "bucketed_solves_per_s": count / t_mixed,
            "captured_s": t_graph, "captured_solves_per_s": count / t_graph,
            "captured_equals_eager": bool(same),
            "one_at_a_time_s": t_single, "one_at_a_time_solves_per_s": count / t_single,
            "speedup": t_single / t_mixed}


def precision_form_bench(E, torch, n: int = 32768, b: int = 64, nb: int = 2048):
    """The precision form: n = 32768, FP32 storage, b = 64, nb = 2048, a
    spectrum of 16 clusters of 2048 eigenvalues (centres uniform [-1, 1],
    spread 1e-6) on a basis of Gaussian Householder reflectors, solved with
    split-FP16 and split-BF16 GEMMs (one untimed warm-up solve first).  The
    exact eigenvalues are the prescribed ones; backward error and
    orthogonality are taken in FP64 on the GPU."""
    import numpy as np

    rng = np.random.default_rng(4)
    centres = rng.uniform(-1.0, 1.0, 16)
    lam = np.concatenate([c + 1e-6 * rng.uniform(-1.0, 1.0, n // 16) for c in centres])
    ctx = E.Context(0)
    out = {"workload": f"n={n} b={b} nb={nb}, 16 clusters of {n // 16} eigenvalues (spread 1e-6), "
                       "Householder-product basis, FP32 storage"}
    try:
        dA = E.gen_spectrum(ctx, torch.from_numpy(lam).cuda(), 21)
        lam = np.sort(lam)
        ctx.enable_timing(True)
        ctx.set_precision(E.PREC_FP16)
        E.syevd(ctx, dA, b=b, nb=nb)  # warm-up: arena, merge plans
        for name, prec in (("fp16", E.PREC_FP16), ("bf16", E.PREC_BF16)):
            ctx.set_precision(prec)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = E.syevd(ctx, dA, b=b, nb=nb)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            stages = ctx.stage_ms()
            ctx.set_precision(E.PREC_FP32)
            w = r.w.cpu().numpy()
            out[name] = {"s_per_evd": dt, "stage_ms": stages,
                         "eig_err": float(np.abs(w - lam).max() / np.abs(lam).max()),
                         "backward_error": float(E.backward_error(ctx, dA, r.w, r.Z)),
                         "orthogonality": float(E.checkOrthogonality(ctx, r.Z))}
            del r
            print(f"[bench] precision form {name}: {dt:.3f} s/EVD", file=sys.stderr, flush=True)
        del dA
    finally:
        ctx.set_precision(E.PREC_FP32)
        ctx.close()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--b", type=int, default=32)
    ap.add_argument("--nb", type=int, default=1024)
    ap.add_argument("--chase-depth", type=int, default=0,
                    help="sweeps per workgroup of the bulge chase (0: the library default, one)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision (FP16/BF16 split) runs")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop sample")
    ap.add_argument("--no-captured", action="store_true", help="skip the eager-vs-graph small-solve sample")
    ap.add_argument("--no-batched", action="store_true", help="skip the mixed-size batched sample")
    ap.add_argument("--no-precision-form", action="store_true",
                    help="skip the n=32768 b=64 split-FP16 / split-BF16 solves")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--cpu-full", action="store_true",
                    help="time host MKL dsyevd on the headline matrix itself (minutes) instead of the n=8192 sample")
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
    if args.chase_depth > 0:
        ctx.set_knob("chase_depth", args.chase_depth)
    # the chase kernel that runs: the grouped kernel for 2-3 sweeps per
    # workgroup (b <= 32 and the window holding (2D + 2) b - D + 1 columns)
    depth = min(args.chase_depth or 1, 3, max(0, (255 - 2 * b) // (2 * b - 1))) if b <= 32 else 1
    chase_name = "chase_group_kernel" if depth >= 2 else "chase_kernel"
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

    # the two largest kernels, each against its own roofline, timed alone with
    # hipEvents on the solver's stream:
    #  * bcw_kernel (BC back transformation, WY blocks on FP32 MFMA): 2 n^3
    #    algorithmic flops per launch against the FP32 MFMA peak;
    #  * the bulge chase (HBM/latency bound): the algorithmic window bytes
    #    (1.5 b + 1) n^2 s, s = 4 bytes per element, per launch against HBM
    #    (the band is held in FP64, twice that element size)
    bc_ms = stage_ms["bc_back_kernel"]
    achieved = 2.0 * n ** 3 / (bc_ms * 1e-3) / 1e12 if bc_ms else None
    bcw = {"bound": "mfma", "achieved": achieved, "peak": PEAK["fp32_mfma_tflops"], "unit": "TFLOP/s",
           "frac": achieved / PEAK["fp32_mfma_tflops"] if achieved else None,
           "traffic": profiled_traffic("bcw_kernel") if (n, b) == (16384, 32) else None, "kernel": "bcw_kernel",
           "ms": bc_ms,
           "note": "2 n^3 algorithmic flops per launch over the hipEvent time of the launch (the WY form issues "
                   "1.625x that on v_mfma_f32_16x16x4f32); FP32 MFMA peak 157.3 TF; traffic = copy-calibrated "
                   "FETCH_SIZE + WRITE_SIZE bytes per launch from the committed rocprofv3 summary"}
    ch_ms = stage_ms.get("chase_kernel") or stage_ms["bc"]
    ch_bytes = (1.5 * b + 1) * n * n * 4
    ch_ach = ch_bytes / (ch_ms * 1e-3) / 1e9 if ch_ms else None
    chase = {"bound": "hbm", "achieved": ch_ach, "peak": PEAK["hbm_gbs"], "unit": "GB/s",
             "frac": ch_ach / PEAK["hbm_gbs"] if ch_ach else None,
             "traffic": profiled_traffic(chase_name) if (n, b) == (16384, 32) else None,
             "kernel": chase_name, "ms": ch_ms,
             "us_per_dependent_step": ch_ms * 1e3 / (2.0 * n) if ch_ms else None,
             "note": "(1.5 b + 1) n^2 x 4 bytes algorithmic window traffic per launch over the hipEvent time of "
                     "the launch; 8 TB/s HBM peak; the chase is bound by its 2n dependent steps"}
    roofline = bcw if (bc_ms or 0) >= (ch_ms or 0) else chase
    achieved = bcw["achieved"]
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
    # the headline (D, E) from the chase and the matrix itself, for the host
    # baseline; then the headline buffers are released before the other legs
    cpu_in = None
    if not args.no_cpu:
        Aw.copy_(A0)
        Ab, _, _ = E.sy2sb(ctx, Aw, b, nb)
        band = E.band_pack(ctx, Ab, b)
        Dd, Ed, _ = E.sb2st(ctx, band, n, b)
        cpu_in = (A0.t().contiguous().cpu().numpy().T if args.cpu_full else None, Dd.cpu().numpy(), Ed.cpu().numpy())
        del Ab, band, Dd, Ed
    ctx.close()
    del A0, Aw, r
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    captured = None
    if not args.no_captured:
        captured = captured_bench(E, torch)
        print(f"[bench] captured: {captured}", file=sys.stderr, flush=True)
    batched = None
    if not args.no_batched:
        batched = batched_form_bench(E, torch)
        print(f"[bench] batched form: {batched}", file=sys.stderr, flush=True)
    precision_form = None
    if not args.no_precision_form:
        precision_form = precision_form_bench(E, torch)
    serving = None
    if not args.no_serve:
        serving = serving_bench(E, torch, args.serve_requests)
        print(f"[bench] serving: {serving}", file=sys.stderr, flush=True)
    cpu = None if cpu_in is None else cpu_baseline(cpu_in[0], cpu_in[1], cpu_in[2], args.cpu_full)
    print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)
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
        "rooflines": {"bcw_kernel": bcw, chase_name: chase},
        "cpu_baseline": cpu,
        "accuracy": acc,
        "mixed_precision": mixed or None,
        "serving": serving,
        "captured": captured,
        "batched_form": batched,
        "precision_form": precision_form,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
