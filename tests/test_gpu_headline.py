"""Eigenvalue parity on the headline input: n = 16384, b = 32, nb = 1024,
A = (X + X^T)/2, X i.i.d. uniform[-1,1] from the Philox stream with seed 0.

The golden eigenvalues are host LAPACK dsyevd (FP64) of the same FP32 matrix
(tests/golden/gen_golden_headline.py).  The device generator's output is
checked against the fixture's SHA-256 first, so the GPU solves the very
matrix LAPACK solved; then every precision form must match within
1e-5 ||A||_2 and meet the backward-error and orthogonality bars of 1e-6."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

from tests.helpers import to_host

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).parent / "golden" / "headline_n16384_s0.npz"


@pytest.fixture(scope="module")
def headline(ctx):
    import torch

    import evd4sc2025_amd as E

    g = np.load(GOLD)
    A = E.gen_symmetric(ctx, 16384, 0, 0)
    torch.cuda.synchronize()
    # column-major storage = the row-major bytes of the transposed view
    raw = A.t().contiguous().cpu().numpy().tobytes()
    assert hashlib.sha256(raw).digest() == g["input_sha256_colmajor"].tobytes(), "device input differs from the fixture"
    return A, g["w"], float(g["norm2"])


@pytest.mark.parametrize("prec", ["fp32", "fp16", "bf16"])
def test_headline_eigenvalues_match_lapack(ctx, headline, prec):
    import torch

    import evd4sc2025_amd as E

    A, w0, nrm2 = headline
    p = {"fp32": E.PREC_FP32, "fp16": E.PREC_FP16, "bf16": E.PREC_BF16}[prec]
    print(f"headline n=16384 b=32 nb=1024 seed 0, {prec}", flush=True)
    ctx.set_precision(p)
    try:
        r = E.syevd(ctx, A, b=32, nb=1024, jobz="V")
    finally:
        ctx.set_precision(E.PREC_FP32)
    torch.cuda.synchronize()
    err = np.abs(to_host(r.w) - w0).max() / nrm2
    be = E.backward_error(ctx, A, r.w, r.Z)
    orth = E.checkOrthogonality(ctx, r.Z)
    print(f"{prec}: eig err / ||A||_2 {err:.3e}, backward {be:.3e}, orth {orth:.3e}", flush=True)
    assert err <= 1e-5
    assert be <= 1e-6
    assert orth <= 1e-6
