"""Mixed-size batches (evd_syevd_mixed_f32): problems of different orders are
embedded in bucket orders with a diagonal pad above ||A||_inf and solved as
same-size batches.  The eigenvalues must match FP64 LAPACK within
1e-5 ||A||_2, the eigenvectors must meet the backward-error and
orthogonality bars, and every result must agree with the unpadded solve of
the same matrix."""
import numpy as np
import pytest

from oracle import evd_oracle as O
from tests.helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _check(A, w, Z):
    A64 = A.astype(np.float64)
    wref = np.linalg.eigvalsh(A64)
    nrm2 = max(np.abs(wref).max(), 1e-300)
    assert np.abs(w - wref).max() <= 1e-5 * nrm2
    if Z is not None:
        n = A.shape[0]
        Z64 = Z.astype(np.float64)
        R = A64 @ Z64 - Z64 * w[None, :]
        nA = np.linalg.norm(A64)
        if nA > 0:
            assert np.linalg.norm(R) / (n * nA) <= 1e-6
        assert np.linalg.norm(Z64.T @ Z64 - np.eye(n)) / n <= 1e-6


def test_mixed_sizes_match_lapack_and_single_solves(ctx):
    import evd4sc2025_amd as E

    sizes = [70, 100, 150, 260, 333, 64, 500, 97, 190, 100, 260, 700]
    jobz = ["V" if i % 2 == 0 else "N" for i in range(len(sizes))]
    mats = [O.gen_symmetric(n, 300 + i, i % 2) for i, n in enumerate(sizes)]
    print(f"mixed sizes {sizes}", flush=True)
    res = E.syevd_mixed(ctx, [to_dev(M) for M in mats], jobz)
    for i, (M, r) in enumerate(zip(mats, res)):
        w = to_host(r.w)
        Z = to_host(r.Z) if r.Z is not None else None
        _check(M, w, Z)
        single = E.syevd(ctx, to_dev(M), b=32, nb=256, jobz=jobz[i])
        ws = to_host(single.w)
        assert np.abs(w - ws).max() <= 1e-5 * max(np.abs(ws).max(), 1e-300), f"problem {i} (n={sizes[i]})"


def test_mixed_sizes_zero_and_tiny(ctx):
    import evd4sc2025_amd as E

    # a zero matrix (pad above a zero norm) and a small-norm one (pad scaled by ||A||)
    mats = [np.zeros((80, 80), np.float32), O.gen_symmetric(65, 9, 0) * np.float32(1e-3)]
    res = E.syevd_mixed(ctx, [to_dev(M) for M in mats], "V")
    w0, Z0 = to_host(res[0].w), to_host(res[0].Z)
    assert np.all(w0 == 0.0)
    assert np.linalg.norm(Z0.T.astype(np.float64) @ Z0 - np.eye(80)) / 80 <= 1e-6
    _check(mats[1].astype(np.float64), to_host(res[1].w), to_host(res[1].Z))


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_mixed_sizes_near_largest_bucket(ctx, prec):
    """Orders just above and below a bucket boundary of the batched form's
    range (bucket 2048: pads of 1 to 511 rows, ||A||_inf / ||A||_2 ~ 16), at
    the 1e-5 ||A||_2 bar, in the plain and the split-FP16 forms."""
    import evd4sc2025_amd as E

    sizes = [1537, 1800, 2047, 1600]
    jobz = ["V", "N", "V", "N"]
    mats = [O.gen_symmetric(n, 400 + i, 0) for i, n in enumerate(sizes)]
    print(f"mixed sizes {sizes} {prec}", flush=True)
    ctx.set_precision(E.PREC_FP16 if prec == "fp16" else E.PREC_FP32)
    try:
        res = E.syevd_mixed(ctx, [to_dev(M) for M in mats], jobz)
    finally:
        ctx.set_precision(E.PREC_FP32)
    for M, r in zip(mats, res):
        _check(M, to_host(r.w), to_host(r.Z) if r.Z is not None else None)
