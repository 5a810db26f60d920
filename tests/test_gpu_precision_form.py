"""The precision form at its full shape: n = 32768, FP32 storage, b = 64,
nb = 2048, spectrum of 16 clusters of 2048 eigenvalues (centres uniform
[-1, 1], intra-cluster spread 1e-6), eigenbasis a product of Gaussian
Householder reflectors; SBR and back-transformation GEMMs in split FP16 and,
separately, split BF16.  The exact eigenvalues are the prescribed ones, so
the eigenvalue error needs no host reference; backward error and
orthogonality are computed in FP64 on the GPU."""
import numpy as np
import pytest

from tests.helpers import to_host

pytestmark = pytest.mark.gpu

N = 32768


@pytest.fixture(scope="module")
def problem():
    import torch

    import evd4sc2025_amd as E

    ctx = E.Context(0)
    rng = np.random.default_rng(4)
    centres = rng.uniform(-1.0, 1.0, 16)
    lam = np.concatenate([c + 1e-6 * rng.uniform(-1.0, 1.0, N // 16) for c in centres])
    print(f"precision form: generating n={N}", flush=True)
    dA = E.gen_spectrum(ctx, torch.from_numpy(lam).cuda(), 21)
    torch.cuda.synchronize()
    yield ctx, dA, np.sort(lam)
    del dA
    ctx.close()


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_precision_form_full_size(problem, prec):
    import torch

    import evd4sc2025_amd as E

    ctx, dA, lam = problem
    p = E.PREC_FP16 if prec == "fp16" else E.PREC_BF16
    print(f"precision form n={N} b=64 nb=2048 {prec}", flush=True)
    ctx.set_precision(p)
    try:
        r = E.syevd(ctx, dA, b=64, nb=2048)
    finally:
        ctx.set_precision(E.PREC_FP32)
    torch.cuda.synchronize()
    err = np.abs(to_host(r.w) - lam).max()
    be = E.backward_error(ctx, dA, r.w, r.Z)
    orth = E.checkOrthogonality(ctx, r.Z)
    print(f"{prec}: eig err {err:.3e}, backward {be:.3e}, orth {orth:.3e}", flush=True)
    del r
    assert err <= 1e-5 * np.abs(lam).max()
    assert be <= 1e-6
    assert orth <= 1e-6
