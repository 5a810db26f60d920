"""The batched form at its own workload: 512 symmetric FP32 matrices, n drawn
log-uniform in [64, 2048] (seed 1), uniform[-1,1] entries symmetrised, half
with eigenvectors (alternating), bandwidth min(32, N / 4) for each bucket
order N, all through one evd_syevd_mixed_f32 call.  Every problem's
eigenvalues are checked against host LAPACK (FP64) within 1e-5 ||A||_2, and
every eigenvector problem's backward error and orthogonality against 1e-6."""
import concurrent.futures as cf

import numpy as np
import pytest

from tests.helpers import to_host

pytestmark = pytest.mark.gpu


def _sizes(count=512, seed=1):
    rng = np.random.default_rng(seed)
    return [int(x) for x in np.exp(rng.uniform(np.log(64), np.log(2048), size=count)).round()]


def test_batched_form_512_matrices(ctx):
    import torch
    from threadpoolctl import threadpool_limits

    import evd4sc2025_amd as E

    sizes = _sizes()
    jobz = ["V" if i % 2 == 0 else "N" for i in range(len(sizes))]
    print(f"batched form: {len(sizes)} matrices, n in [{min(sizes)}, {max(sizes)}]", flush=True)
    mats = [E.gen_symmetric(ctx, n, 5000 + i, 0) for i, n in enumerate(sizes)]
    res = E.syevd_mixed(ctx, mats, jobz, b=0)
    torch.cuda.synchronize()
    hosts = [to_host(M).astype(np.float64) for M in mats]
    with threadpool_limits(1), cf.ThreadPoolExecutor(16) as ex:
        refs = list(ex.map(np.linalg.eigvalsh, hosts))
    print("batched form: host eigenvalues done", flush=True)
    worst_w = worst_be = worst_orth = 0.0
    for i, (M, r, w0) in enumerate(zip(mats, res, refs)):
        nrm2 = np.abs(w0).max()
        e = np.abs(to_host(r.w) - w0).max() / nrm2
        worst_w = max(worst_w, e)
        assert e <= 1e-5, f"problem {i} (n={sizes[i]}): eigenvalue error {e:.2e}"
        if r.Z is not None:
            be = E.backward_error(ctx, M, r.w, r.Z)
            orth = E.checkOrthogonality(ctx, r.Z)
            worst_be, worst_orth = max(worst_be, be), max(worst_orth, orth)
            assert be <= 1e-6 and orth <= 1e-6, f"problem {i} (n={sizes[i]}): {be:.2e} {orth:.2e}"
    print(f"batched form: worst eig err {worst_w:.2e}, backward {worst_be:.2e}, orth {worst_orth:.2e}", flush=True)


@pytest.mark.parametrize("n,b", [(64, 16), (96, 24), (100, 24), (70, 16)])
def test_small_orders_with_narrow_bands(ctx, n, b):
    """jobz = V at b = 16 and 24 (the small buckets' bandwidths): the
    small-order back transformation, single solves."""
    import evd4sc2025_amd as E

    dA = E.gen_symmetric(ctx, n, 60 + n + b, 0)
    A = to_host(dA).astype(np.float64)
    r = E.syevd(ctx, dA, b=b, nb=4 * b, jobz="V")
    w0 = np.linalg.eigvalsh(A)
    assert np.abs(to_host(r.w) - w0).max() <= 1e-5 * np.abs(w0).max()
    assert E.backward_error(ctx, dA, r.w, r.Z) <= 1e-6
    assert E.checkOrthogonality(ctx, r.Z) <= 1e-6
