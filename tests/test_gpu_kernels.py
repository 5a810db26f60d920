"""GPU parity of the building blocks against the oracle / FP64 references,
through the C ABI (libevd4sc.so)."""
import numpy as np
import pytest

from oracle import evd_oracle as O
from tests.helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize(
    "m,n,k",
    [(300, 200, 100), (4000, 32, 3000), (32, 32, 5000), (64, 500, 70), (1, 7, 3),
     # long-K skinny products (256 x 32 tiles, K split): interior tiles on the
     # two-slab pipeline, an edge tile, a partial last slab, an unaligned lda
     (5000, 32, 4100), (2048, 32, 4096), (4999, 32, 4103),
     # the same for b = 64 (256 x 64 tiles), and a 48-column edge tile
     (5000, 64, 4100), (4999, 64, 4103), (300, 48, 5000),
     # small K splits (64 x 64 and 64 x 32 tiles)
     (96, 96, 1024), (700, 32, 900)])
def test_sgemm_matches_fp64(ctx, ta, tb, m, n, k):
    import evd4sc2025_amd as E

    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k)).astype(np.float32)
    B = rng.standard_normal((n, k) if tb == "T" else (k, n)).astype(np.float32)
    C0 = rng.standard_normal((m, n)).astype(np.float32)
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C0)
    E.sgemm(ctx, ta, tb, 0.75, dA, dB, -0.5, dC)
    opA = A.T if ta == "T" else A
    opB = B.T if tb == "T" else B
    want = 0.75 * (opA.astype(np.float64) @ opB.astype(np.float64)) - 0.5 * C0
    got = to_host(dC)
    err = np.abs(got - want).max() / max(1.0, np.abs(want).max())
    assert err < 5e-6 * np.sqrt(k), err


def _view(a, off, ld):
    """Device column-major view of a at row offset `off` inside columns of
    leading dimension `ld` (base alignment and ld chosen by the caller)."""
    import torch

    rows, cols = a.shape
    full = np.zeros((ld, cols), dtype=a.dtype)
    full[off:off + rows] = a
    return torch.from_numpy(np.ascontiguousarray(full.T)).to("cuda").t()[off:off + rows]


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("off,pad", [(4, 7), (1, 3), (0, 0)])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_gemm_strided_views(ctx, ta, tb, off, pad, dt):
    """Operands at 16-byte aligned and unaligned bases with padded leading
    dimensions and edges that cut a 16-byte vector (m, k odd): the vector and
    element load paths of the tile loader give the FP64 product."""
    import evd4sc2025_amd as E

    m, n, k = 301, 203, 101
    f = np.float32 if dt == "f32" else np.float64
    rng = np.random.default_rng(off + pad)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k)).astype(f)
    B = rng.standard_normal((n, k) if tb == "T" else (k, n)).astype(f)
    C0 = rng.standard_normal((m, n)).astype(f)
    ld = lambda r: (r + off + pad + 3) // 4 * 4 if pad else r + off
    dA, dB, dC = _view(A, off, ld(A.shape[0])), _view(B, off, ld(B.shape[0])), to_dev(C0)
    (E.sgemm if dt == "f32" else E.dgemm)(ctx, ta, tb, 0.75, dA, dB, -0.5, dC)
    want = 0.75 * ((A.T if ta == "T" else A).astype(np.float64) @ (B.T if tb == "T" else B).astype(np.float64)) \
        - 0.5 * C0
    err = np.abs(to_host(dC) - want).max() / np.abs(want).max()
    assert err < (5e-6 * np.sqrt(k) if dt == "f32" else 1e-13), err


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_dgemm_matches_fp64(ctx, ta, tb):
    import evd4sc2025_amd as E

    rng = np.random.default_rng(5)
    m, n, k = 130, 70, 90
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    C0 = rng.standard_normal((m, n))
    dC = to_dev(C0)
    E.dgemm(ctx, ta, tb, 1.5, to_dev(A), to_dev(B), 0.25, dC)
    want = 1.5 * ((A.T if ta == "T" else A) @ (B.T if tb == "T" else B)) + 0.25 * C0
    np.testing.assert_allclose(to_host(dC), want, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("n,k", [(200, 64), (1000, 257), (129, 1)])
def test_syr2k_lower_mirror(ctx, n, k):
    import evd4sc2025_amd as E

    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, k)).astype(np.float32)
    B = rng.standard_normal((n, k)).astype(np.float32)
    C = rng.standard_normal((n, n)).astype(np.float32)
    C = (C + C.T) / 2
    dC = to_dev(C)
    E.tc_ozimmu_syr2k(ctx, -1.0, to_dev(A), to_dev(B), 1.0, dC)
    got = to_host(dC)
    want = O.syr2k(-1.0, A.astype(np.float64), B.astype(np.float64), 1.0, C.astype(np.float64))
    assert np.array_equal(got, got.T)
    assert np.abs(got - want).max() < 1e-5 * np.sqrt(k) * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("n,seed", [(64, 0), (130, 1), (333, 7)])
def test_generator_bit_exact(ctx, n, seed):
    import evd4sc2025_amd as E

    A = to_host(E.gen_symmetric(ctx, n, seed, 0))
    assert np.array_equal(A, O.gen_symmetric(n, seed, 0))


@pytest.mark.parametrize("n,seed", [(64, 3), (257, 11)])
def test_gaussian_generator_matches_oracle(ctx, n, seed):
    """kind 1 (Box-Muller on the same Philox words as kind 0): the device's
    FP64 log / cos and the host's libm may differ in the last bits, so the
    FP32 entries agree to one FP32 ulp, and nearly all of them exactly."""
    import evd4sc2025_amd as E

    A = to_host(E.gen_symmetric(ctx, n, seed, 1))
    W = O.gen_symmetric(n, seed, 1)
    assert (np.abs(A - W) <= np.spacing(np.abs(W))).all()
    assert np.mean(A == W) > 0.999


@pytest.mark.parametrize("m,b", [(64, 32), (200, 32), (1000, 32), (5000, 32), (160, 64), (3000, 64), (20, 32)])
def test_panel_qr_matches_oracle(ctx, m, b):
    import evd4sc2025_amd as E

    torch = _torch()
    P = O.gen_symmetric(max(m, b), 40 + m)[:m, :b]
    V, T, R, tau = O.panel_qr(P.astype(np.float64))
    k = V.shape[1]
    dA = to_dev(P)
    dW = E.colmajor(m, b, fill=0.0)
    dR = E.colmajor(b, b, fill=0.0)
    E.panelQR(ctx, dA, dW, dR)
    torch.cuda.synchronize()
    Y = to_host(dA)
    scale = max(1.0, np.abs(P).max())
    np.testing.assert_allclose(Y[:, :k], V, atol=2e-5 * np.sqrt(m))
    np.testing.assert_allclose(to_host(dW)[:, :k], V @ T, atol=2e-5 * np.sqrt(m))
    kr = min(b, m)
    np.testing.assert_allclose(np.triu(to_host(dR))[:kr], R[:kr], atol=2e-5 * scale * np.sqrt(m))
    # the factorisation reproduces the panel: (I - W Y^T)[R;0] = P
    W = to_host(dW).astype(np.float64)
    Rf = np.zeros((m, b))
    Rf[:kr] = np.triu(to_host(dR))[:kr]
    rec = Rf - W @ (Y.astype(np.float64).T @ Rf)
    assert np.abs(rec - P).max() < 1e-5 * scale * np.sqrt(m)


def _panel_case(kind, m, b, seed):
    rng = np.random.default_rng(seed)
    P = rng.uniform(-1, 1, (m, b))
    if kind == "dup":  # rank deficient: two repeated columns
        P[:, 5] = P[:, 2]
        P[:, 20] = -P[:, 7]
    elif kind == "zero":
        P[:] = 0.0
    elif kind == "illcond":  # full rank, condition number about 1e9
        U, _ = np.linalg.qr(rng.standard_normal((m, b)))
        Vt, _ = np.linalg.qr(rng.standard_normal((b, b)))
        P = (U * np.logspace(0, -9, b)) @ Vt
    elif kind == "scaled":  # columns of very different norms, well conditioned
        P = P * np.logspace(3, -3, b)
    elif kind == "moderate":  # condition number 1e4: two Cholesky passes
        U, _ = np.linalg.qr(rng.standard_normal((m, b)))
        Vt, _ = np.linalg.qr(rng.standard_normal((b, b)))
        P = (U * np.logspace(0, -4, b)) @ Vt
    return P.astype(np.float32)


@pytest.mark.parametrize("kind", ["random", "dup", "zero", "illcond", "scaled", "moderate"])
@pytest.mark.parametrize("m", [64, 700, 5000])
def test_panel_cholqr_and_fallback_match_oracle(ctx, kind, m):
    """Panels of 32 columns go through Cholesky QR + Householder
    reconstruction (one Cholesky pass for well-conditioned FP32 panels, two
    for condition numbers above 1e3); rank-deficient and ill-conditioned ones
    are flagged on the device and factored by the column-exchange kernel.  Every kind against the
    dgeqrf-pinned oracle, and the Cholesky path against the exchange path
    (knob panel_exchange = 1)."""
    import evd4sc2025_amd as E

    torch = _torch()
    b = 32
    P = _panel_case(kind, m, b, 70 + m)
    V, T, R, tau = O.panel_qr(P.astype(np.float64))
    outs = []
    for exch in (0, 1):
        ctx.set_knob("panel_exchange", exch)
        try:
            dA, dW, dR = to_dev(P), E.colmajor(m, b, fill=0.0), E.colmajor(b, b, fill=0.0)
            E.panelQR(ctx, dA, dW, dR)
            torch.cuda.synchronize()
        finally:
            ctx.set_knob("panel_exchange", 0)
        outs.append((to_host(dA), to_host(dW), np.triu(to_host(dR))))
    scale = max(1.0, np.abs(P).max())
    for Y, W, Rg in outs:
        assert np.all(np.isfinite(Y)) and np.all(np.isfinite(W)) and np.all(np.isfinite(Rg))
        # the factorisation reproduces the panel: (I - W Y^T)[R; 0] = P
        Rf = np.zeros((m, b))
        Rf[:b] = Rg
        rec = Rf - W.astype(np.float64) @ (Y.astype(np.float64).T @ Rf)
        assert np.abs(rec - P).max() < 1e-5 * scale * np.sqrt(m), kind
        # Q = I - W Y^T is orthogonal on the panel's span
        Qt = np.eye(m)[:, :b] - W.astype(np.float64) @ Y.astype(np.float64)[:b].T
        assert np.abs(Qt.T @ Qt - np.eye(b)).max() < 1e-4, kind
        if kind in ("random", "scaled", "moderate"):
            np.testing.assert_allclose(Y, V, atol=2e-5 * np.sqrt(m))
            np.testing.assert_allclose(W, V @ T, atol=2e-5 * np.sqrt(m))
            np.testing.assert_allclose(Rg, R[:b], atol=2e-5 * scale * np.sqrt(m))
    if kind in ("random", "scaled", "moderate"):
        (Y0, W0, R0), (Y1, W1, R1) = outs
        assert np.abs(Y0 - Y1).max() < 1e-5 and np.abs(R0 - R1).max() < 1e-5 * scale


@pytest.mark.parametrize("n,b,nb", [(64, 32, 32), (130, 32, 64), (512, 32, 128), (700, 64, 128), (1000, 32, 1024)])
def test_sy2sb_band_spectrum(ctx, n, b, nb):
    import evd4sc2025_amd as E

    A = O.gen_symmetric(n, 200 + n)
    dA = to_dev(A)
    _, dW, dY = E.sy2sb(ctx, dA, b, nb)
    Ab = to_host(dA).astype(np.float64)
    B = O.band_of(Ab, b)
    wA = np.linalg.eigvalsh(A.astype(np.float64))
    wB = np.linalg.eigvalsh(B)
    nrm = np.abs(wA).max()
    assert np.abs(wA - wB).max() < 2e-6 * nrm * max(1.0, np.sqrt(n / 512))
    # Q_SBR^T A Q_SBR = B with Q from the returned W, Y
    Q = O.q_sbr(to_host(dW).astype(np.float64), to_host(dY).astype(np.float64), b)
    assert O.orthogonality(Q) < 1e-6
    assert np.abs(Q.T @ A.astype(np.float64) @ Q - B).max() < 1e-5 * nrm


@pytest.mark.parametrize("n,b", [(40, 32), (130, 32), (600, 32), (300, 64), (1500, 32), (200, 20), (257, 48), (3, 32)])
def test_sb2st_matches_oracle(ctx, n, b):
    import evd4sc2025_amd as E

    A = O.gen_symmetric(n, 300 + n).astype(np.float64)
    Ab, _, _ = O.sy2sb(A, b, 2 * b)
    B = O.band_of(Ab, b)
    D0, E0, U0 = O.sb2st(B, b)
    band = E.band_pack(ctx, to_dev(B.astype(np.float32)), b)
    D, Ee, U = E.sb2st(ctx, band, n, b)
    D, Ee, U = to_host(D), to_host(Ee), to_host(U)
    # same algorithm in FP64 on the same (FP32-rounded) band
    B32 = B.astype(np.float32).astype(np.float64)
    D1, E1, U1 = O.sb2st(B32, b)
    np.testing.assert_allclose(D, D1, atol=1e-10)
    np.testing.assert_allclose(Ee, E1, atol=1e-10)
    np.testing.assert_allclose(U, U1, atol=1e-5)
    wT = np.sort(np.linalg.eigvalsh(np.diag(D) + np.diag(Ee, 1) + np.diag(Ee, -1)))
    np.testing.assert_allclose(wT, np.linalg.eigvalsh(B32), atol=1e-10)


def _tridiag_cases():
    rng = np.random.default_rng(17)
    cases = []
    for n in [1, 2, 3, 17, 100, 513, 2000]:
        cases.append((f"rand{n}", rng.uniform(-1, 1, n), rng.uniform(-1, 1, max(n - 1, 0))))
    k = 10
    cases.append(("wilkinson21", np.abs(np.arange(-k, k + 1)).astype(float), np.ones(2 * k)))
    cases.append(("diag", rng.uniform(-1, 1, 50), np.zeros(49)))
    cases.append(("toeplitz", 2 * np.ones(300), np.ones(299)))
    d = np.repeat(rng.uniform(-1, 1, 8), 64) + 1e-9 * rng.standard_normal(512)
    cases.append(("cluster", d, 1e-7 * rng.standard_normal(511)))
    cases.append(("glued", np.ones(400), np.where(np.arange(399) % 50 == 49, 1e-12, 1.0)))
    # decoupled 2x2 blocks: every other off-diagonal entry is zero
    cases.append(("blocks2x2", rng.uniform(-1, 1, 64), np.where(np.arange(63) % 2 == 0, rng.uniform(-1, 1, 63), 0.0)))
    return cases


@pytest.mark.parametrize("name,d,e", _tridiag_cases(), ids=[c[0] for c in _tridiag_cases()])
def test_stedc_matches_lapack(ctx, name, d, e):
    import evd4sc2025_amd as E

    n = d.size
    w, Z = E.stedc(ctx, to_dev(d), to_dev(e))
    w, Z = to_host(w), to_host(Z)
    w0 = O.tridiag_eig(d, e, vectors=False)
    scale = max(1.0, np.abs(w0).max())
    np.testing.assert_allclose(w, w0, atol=1e-13 * scale * max(1, n) ** 0.5)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    resid = np.linalg.norm(T @ Z - Z * w[None, :]) / (scale * max(n, 1))
    assert resid < 1e-14, resid
    assert O.orthogonality(Z) < 1e-14


@pytest.mark.parametrize("n", [4096, 6001])
def test_stedc_library_products_match_grouped_kernel(ctx, n):
    """Merges of order >= 4096 form their FP64 products with rocBLAS (device
    problem list read back, product scattered to its columns); the knob
    own_gemm = 1 keeps them on the grouped tile kernel.  Both against LAPACK,
    and the two eigenvector sets against each other."""
    import evd4sc2025_amd as E

    rng = np.random.default_rng(n)
    d, e = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n - 1)
    w1, Z1 = E.stedc(ctx, to_dev(d), to_dev(e))
    ctx.set_knob("own_gemm", 1)
    try:
        w2, Z2 = E.stedc(ctx, to_dev(d), to_dev(e))
    finally:
        ctx.set_knob("own_gemm", 0)
    w1, Z1, w2, Z2 = to_host(w1), to_host(Z1), to_host(w2), to_host(Z2)
    w0 = O.tridiag_eig(d, e, vectors=False)
    scale = max(1.0, np.abs(w0).max())
    for w in (w1, w2):
        np.testing.assert_allclose(w, w0, atol=1e-13 * scale * n ** 0.5)
    np.testing.assert_allclose(w1, w2, atol=1e-14 * scale)
    assert np.abs(Z1 - Z2).max() < 1e-8
    T_Z1 = d[:, None] * Z1
    T_Z1[:-1] += e[:, None] * Z1[1:]
    T_Z1[1:] += e[:, None] * Z1[:-1]
    assert np.linalg.norm(T_Z1 - Z1 * w1[None, :]) / (scale * n) < 1e-14


@pytest.mark.parametrize("valu", [0, 1])
@pytest.mark.parametrize("n,b,order", [(70, 32, 1), (70, 32, 0), (300, 32, 1), (300, 32, 0), (257, 64, 1),
                                       (257, 64, 0), (1100, 32, 1), (66, 64, 1), (130, 64, 1), (700, 64, 1),
                                       (1100, 64, 1)])
def test_bc_back_matches_oracle(ctx, n, b, order, valu):
    """order 1 runs as WY blocks on FP32 MFMA (valu = 0; b = 32 and b = 64
    kernels) or on the FP32 VALU kernel (valu = 1); order 0 always takes the
    VALU kernel.  150 columns: the b = 64 kernel's last 32-column workgroup
    is partial."""
    import evd4sc2025_amd as E

    ctx.set_knob("bc_valu", valu)

    A = O.gen_symmetric(n, 400 + n).astype(np.float64)
    Ab, _, _ = O.sy2sb(A, b, 2 * b)
    _, _, U = O.sb2st(O.band_of(Ab, b), b)
    U32 = U.astype(np.float32)
    X = np.random.default_rng(n).standard_normal((n, 150)).astype(np.float32)
    dX = to_dev(X)
    dU = to_dev(U32)
    E.bc_back(ctx, dX, dU, n, b, order)
    ctx.set_knob("bc_valu", 0)
    want = O.bc_back(X.astype(np.float64), U32.astype(np.float64), n, b, order)
    assert np.abs(to_host(dX) - want).max() < 2e-5 * np.abs(X).max()


@pytest.mark.parametrize("n,b", [(256, 32), (1024, 32), (700, 64)])
def test_sbr_back_and_wy(ctx, n, b):
    import evd4sc2025_amd as E

    A = O.gen_symmetric(n, 500 + n)
    dA = to_dev(A)
    _, dW, dY = E.sy2sb(ctx, dA, b, 2 * b)
    W, Y = to_host(dW).astype(np.float64), to_host(dY).astype(np.float64)
    Q = O.q_sbr(W, Y, b)
    X = np.random.default_rng(1).standard_normal((n, 64)).astype(np.float32)
    dX = to_dev(X)
    E.sbr_back(ctx, dX, dW, dY, b)
    assert np.abs(to_host(dX) - Q @ X).max() < 2e-5 * np.sqrt(n / 256) * np.abs(X).max()


def test_wy_from_slide_and_q_from_wy(ctx):
    import evd4sc2025_amd as E

    rng = np.random.default_rng(3)
    M, slide, nblk = 300, 32, 6
    W = (0.1 * rng.standard_normal((M, slide * nblk))).astype(np.float32)
    Y = (0.1 * rng.standard_normal((M, slide * nblk))).astype(np.float32)
    want = O.q_from_wy(O.wy_from_slide(W.astype(np.float64), Y.astype(np.float64), slide), Y.astype(np.float64))
    dW, dY = to_dev(W), to_dev(Y)
    E.computerWYFromSlide(ctx, slide, dW, dY)
    dQ = E.colmajor(M, M)
    E.computerQFromWY(ctx, dQ, dW, dY)
    assert np.abs(to_host(dQ) - want).max() < 1e-4


@pytest.mark.parametrize("m,b", [(300, 32), (2000, 64)])
def test_hou_tsqr_panel_explicit_q(ctx, m, b):
    import evd4sc2025_amd as E

    P = O.gen_symmetric(m, 70 + m)[:, :b]
    dA = to_dev(P)
    dR = E.colmajor(b, b, fill=0.0)
    E.hou_tsqr_panel(ctx, dA, dR)
    Q = to_host(dA).astype(np.float64)
    R = np.triu(to_host(dR).astype(np.float64))
    assert O.orthogonality(Q) < 1e-6
    assert np.abs(Q @ R - P).max() < 1e-5 * np.sqrt(m)
    _, _, R0, _ = O.panel_qr(P.astype(np.float64))
    np.testing.assert_allclose(R, R0, atol=1e-4)


@pytest.mark.parametrize("n,seed", [(64, 3), (257, 8)])
def test_gaussian_generator_matches_oracle(ctx, n, seed):
    """kind = 1 (Box-Muller): the device's FP64 log / cos may differ from
    NumPy's in the last bit, so the FP32 entries must agree to one rounding
    (relative 2^-23) and almost all of them exactly."""
    import evd4sc2025_amd as E

    A = to_host(E.gen_symmetric(ctx, n, seed, 1)).astype(np.float64)
    B = O.gen_symmetric(n, seed, 1).astype(np.float64)
    assert np.all(np.abs(A - B) <= 2.0 ** -23 * np.abs(B) + 1e-30)
    assert np.mean(A == B) > 0.99


@pytest.mark.parametrize("n,b,depth", [(600, 32, 3), (1500, 32, 3), (1500, 32, 2), (130, 32, 3), (40, 32, 3),
                                       (5, 32, 3), (3, 32, 3), (700, 16, 6), (700, 16, 4), (500, 20, 3),
                                       (333, 8, 6), (2000, 32, 1)])
def test_sb2st_sweep_depth_matches_oracle(ctx, n, b, depth):
    """The grouped chase (depth sweeps per workgroup, handed over in LDS) and
    the one-sweep-per-workgroup chase (depth 1) against the oracle."""
    import evd4sc2025_amd as E

    print(f"chase n={n} b={b} depth={depth}", flush=True)
    A = O.gen_symmetric(n, 500 + n + b, 0).astype(np.float64)
    Ab, _, _ = O.sy2sb(A, b, 2 * b)
    B32 = O.band_of(Ab, b).astype(np.float32).astype(np.float64)
    band = E.band_pack(ctx, to_dev(B32.astype(np.float32)), b)
    ctx.set_knob("chase_depth", depth)
    try:
        D, Ee, U = E.sb2st(ctx, band, n, b)
    finally:
        ctx.set_knob("chase_depth", 0)
    D, Ee, U = to_host(D), to_host(Ee), to_host(U)
    D1, E1, U1 = O.sb2st(B32, b)
    np.testing.assert_allclose(D, D1, atol=1e-10)
    np.testing.assert_allclose(Ee, E1, atol=1e-10)
    np.testing.assert_allclose(U, U1, atol=1e-5)
