"""The FP64 forms of the reference entry points (the reference's api/ library
is instantiated for double only): panel QR, explicit thin Q, SYR2K, SBR, the
WY helpers and the back transformation of the chase reflectors, each against
the oracle at FP64 tolerances."""
import numpy as np
import pytest

from oracle import evd_oracle as O
from tests.helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,b", [(64, 32), (1000, 32), (3000, 64), (20, 32)])
def test_panel_qr_f64_matches_oracle(ctx, m, b):
    import torch

    import evd4sc2025_amd as E

    P = O.gen_symmetric(max(m, b), 900 + m).astype(np.float64)[:m, :b]
    dA = to_dev(P)
    k = min(m, b)
    dW = E.colmajor(m, b, dtype=torch.float64, fill=0.0)
    dR = E.colmajor(b, b, dtype=torch.float64, fill=0.0)
    E.panelQR(ctx, dA, dW, dR)
    V0, T0, R0, _ = O.panel_qr(P)
    Y = np.tril(to_host(dA), -1)[:, :k] + np.eye(m, k)
    np.testing.assert_allclose(Y, V0, atol=1e-10)
    np.testing.assert_allclose(np.triu(to_host(dR))[:k, :b], R0[:k, :b], atol=1e-10 * np.abs(P).max() * np.sqrt(m))
    np.testing.assert_allclose(to_host(dW)[:, :k], V0 @ T0, atol=1e-10)


def test_hou_tsqr_panel_f64(ctx):
    import torch

    import evd4sc2025_amd as E

    m, b = 700, 32
    P = O.gen_symmetric(m, 71)[:, :b].astype(np.float64)
    dA = to_dev(P)
    dR = E.colmajor(b, b, dtype=torch.float64, fill=0.0)
    E.hou_tsqr_panel(ctx, dA, dR)
    Q, R = to_host(dA), np.triu(to_host(dR))
    assert O.orthogonality(Q) < 1e-14
    assert np.abs(Q @ R - P).max() < 1e-13 * np.sqrt(m)


@pytest.mark.parametrize("n,k", [(100, 30), (513, 64), (1024, 1024)])
def test_syr2k_f64(ctx, n, k):
    import evd4sc2025_amd as E

    rng = np.random.default_rng(n + k)
    A, B = rng.standard_normal((n, k)), rng.standard_normal((n, k))
    C = rng.standard_normal((n, n))
    C = C + C.T
    dC = to_dev(C)
    E.tc_ozimmu_syr2k(ctx, -0.5, to_dev(A), to_dev(B), 1.5, dC)
    got = to_host(dC)
    want = O.syr2k(-0.5, A, B, 1.5, C)
    assert np.array_equal(got, got.T)
    assert np.abs(got - want).max() < 1e-13 * np.sqrt(k) * np.abs(want).max()


@pytest.mark.parametrize("n,b,nb", [(130, 32, 64), (600, 32, 128), (300, 64, 128)])
def test_sy2sb_f64(ctx, n, b, nb):
    import evd4sc2025_amd as E

    A = O.gen_symmetric(n, 950 + n).astype(np.float64)
    dA = to_dev(A)
    _, dW, dY = E.sy2sb(ctx, dA, b, nb)
    Bf = to_host(dA)
    B = O.band_of(Bf, b)
    nrm = np.abs(np.linalg.eigvalsh(A)).max()
    # the band holds the spectrum, and Q_SBR^T A Q_SBR = B with Q_SBR from W, Y
    np.testing.assert_allclose(np.linalg.eigvalsh(B), np.linalg.eigvalsh(A), atol=1e-12 * nrm)
    Q = O.q_sbr(to_host(dW), to_host(dY), b)
    assert O.orthogonality(Q) < 1e-14
    assert np.abs(Q.T @ A @ Q - B).max() < 1e-12 * nrm


def test_wy_helpers_f64(ctx):
    import torch

    import evd4sc2025_amd as E

    rng = np.random.default_rng(5)
    M, slide, nblk = 300, 32, 6
    W = 0.1 * rng.standard_normal((M, slide * nblk))
    Y = 0.1 * rng.standard_normal((M, slide * nblk))
    want = O.q_from_wy(O.wy_from_slide(W, Y, slide), Y)
    dW, dY = to_dev(W), to_dev(Y)
    E.computerWYFromSlide(ctx, slide, dW, dY)
    dQ = E.colmajor(M, M, dtype=torch.float64)
    E.computerQFromWY(ctx, dQ, dW, dY)
    assert np.abs(to_host(dQ) - want).max() < 1e-13


@pytest.mark.parametrize("n,order", [(70, 0), (70, 1), (300, 0), (300, 1), (1100, 0), (1100, 1), (33, 0), (3, 0)])
def test_bc_back_f64_matches_oracle(ctx, n, order):
    """my_BC_back_trans_v8_10_noBandU in its reference precision (order 0 is
    its product Q <- H_K .. H_1 Q), b = 32."""
    import evd4sc2025_amd as E

    b = 32
    print(f"bc_back_f64 n={n} order={order}", flush=True)
    A = O.gen_symmetric(n, 960 + n).astype(np.float64)
    Ab, _, _ = O.sy2sb(A, b, 2 * b)
    _, _, U = O.sb2st(O.band_of(Ab, b), b)
    X = np.random.default_rng(n).standard_normal((n, 150))
    dX, dU = to_dev(X), to_dev(U)
    E.bc_back(ctx, dX, dU, n, b, order)
    want = O.bc_back(X, U, n, b, order)
    assert np.abs(to_host(dX) - want).max() < 1e-12 * np.abs(X).max()
