"""Parity at the headline and precision-form shapes through properties that
do not need a host reference of the full size: closed-form spectra
(A = Q diag(lambda) Q^T with Q a product of Gaussian Householder reflectors),
backward error ||A Z - Z diag(w)||_F / (n ||A||_F) and orthogonality
||Z^T Z - I||_F / n computed in FP64 on the GPU."""
import numpy as np
import pytest

from tests.helpers import to_host

pytestmark = pytest.mark.gpu


def _known(ctx, n, lam, seed, b, nb, precision):
    import torch

    import evd4sc2025_amd as E

    dA = E.gen_spectrum(ctx, torch.from_numpy(lam).cuda(), seed)
    ctx.set_precision(precision)
    try:
        r = E.syevd(ctx, dA, b=b, nb=nb)
    finally:
        ctx.set_precision(E.PREC_FP32)
    torch.cuda.synchronize()
    w = to_host(r.w)
    err = np.abs(w - np.sort(lam)).max()
    be = E.backward_error(ctx, dA, r.w, r.Z)
    orth = E.checkOrthogonality(ctx, r.Z)
    print(f"n={n} b={b} nb={nb} prec={precision}: eig err {err:.3e}, backward {be:.3e}, orth {orth:.3e}",
          flush=True)
    return err, be, orth


def test_headline_size_known_spectrum(ctx):
    """n = 16384, b = 32, nb = 1024, plain FP32, lambda_i = +-10^(-6 i/(n-1))
    with random signs: ||A||_2 = 1, so the eigenvalue bar is 1e-5."""
    import evd4sc2025_amd as E

    n = 16384
    rng = np.random.default_rng(2)
    lam = 10.0 ** (-6.0 * np.arange(n) / (n - 1)) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    err, be, orth = _known(ctx, n, lam, 11, 32, 1024, E.PREC_FP32)
    assert err <= 1e-5
    assert be <= 1e-6
    assert orth <= 1e-6


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_precision_form_clustered_spectrum(ctx, prec):
    """The precision form's spectrum (16 clusters, centres uniform[-1,1],
    intra-cluster spread 1e-6) with b = 64, nb = 2048 in the split 16-bit
    forms, at n = 4096 (256 eigenvalues per cluster)."""
    import evd4sc2025_amd as E

    n = 4096
    rng = np.random.default_rng(3)
    centres = rng.uniform(-1.0, 1.0, 16)
    lam = np.concatenate([c + 1e-6 * rng.uniform(-1.0, 1.0, n // 16) for c in centres])
    p = E.PREC_FP16 if prec == "fp16" else E.PREC_BF16
    err, be, orth = _known(ctx, n, lam, 12, 64, 2048, p)
    assert err <= 1e-5 * np.abs(lam).max()
    assert be <= 1e-6
    assert orth <= 1e-6
