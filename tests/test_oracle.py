"""Pins the CPU oracle (oracle/evd_oracle.py) before it is trusted as the
checker of the GPU library: published known-answer vectors, LAPACK outputs
(committed golden fixtures and live scipy), closed forms and brute force."""
import hashlib
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import evd_oracle as O

G = np.load(Path(__file__).parent / "golden" / "golden.npz")


def test_philox_known_answers():
    # Random123 philox4x32-10 known-answer vectors
    kat = [
        ([0, 0, 0, 0], (0, 0), [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, (0xFFFFFFFF, 0xFFFFFFFF), [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], (0xA4093822, 0x299F31D0),
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        got = O.philox4x32(np.array([ctr], dtype=np.uint32), key)[0]
        assert [int(x) for x in got] == want


def test_generator_matches_fixture_hash_and_properties():
    for n, seed in [(64, 0), (130, 1), (512, 2)]:
        A = O.gen_symmetric(n, seed)
        assert A.dtype == np.float32 and np.array_equal(A, A.T)
        assert hashlib.sha256(A.tobytes()).digest() == G[f"input_sha256_n{n}_s{seed}"].tobytes()
        assert A.min() >= -1.0 and A.max() <= 1.0


@pytest.mark.parametrize("m,b,seed", [(64, 32, 5), (200, 32, 6), (160, 64, 7)])
def test_panel_qr_matches_lapack_geqrf(m, b, seed):
    P = O.gen_symmetric(m, seed)[:, :b].astype(np.float64)
    V, T, R, tau = O.panel_qr(P)
    np.testing.assert_allclose(R, G[f"geqrf_r_m{m}_b{b}_s{seed}"], atol=1e-12 * np.abs(R).max())
    np.testing.assert_allclose(tau, G[f"geqrf_tau_m{m}_b{b}_s{seed}"], atol=1e-13)
    # V is LAPACK's V as well
    qr, _, _, _ = sl.lapack.dgeqrf(P)
    np.testing.assert_allclose(np.tril(V, -1), np.tril(qr[:, :b], -1), atol=1e-12)
    # dlarft: I - V T V^T equals the product of the reflectors (brute force)
    H = np.eye(m)
    for j in range(b):
        H = H @ (np.eye(m) - tau[j] * np.outer(V[:, j], V[:, j]))
    np.testing.assert_allclose(np.eye(m) - V @ T @ V.T, H, atol=1e-13)
    Rf = np.zeros((m, b))
    Rf[:b] = R
    np.testing.assert_allclose(H @ Rf, P, atol=1e-12)


def test_panel_qr_zero_column_gives_tau_zero():
    P = np.zeros((10, 3))
    P[:, 1] = np.arange(10)
    P[0, 2] = 1.0
    V, T, R, tau = O.panel_qr(P)
    assert tau[0] == 0.0 and np.all(V[:, 0] == np.eye(10)[:, 0])


def test_syr2k_is_textbook():
    rng = np.random.default_rng(0)
    A, B, C = rng.standard_normal((7, 3)), rng.standard_normal((7, 3)), rng.standard_normal((7, 7))
    C = C + C.T
    want = np.zeros((7, 7))
    for i in range(7):
        for j in range(7):
            want[i, j] = -1.5 * sum(A[i, l] * B[j, l] + B[i, l] * A[j, l] for l in range(3)) + 0.5 * C[i, j]
    np.testing.assert_allclose(O.syr2k(-1.5, A, B, 0.5, C), want, atol=1e-12)


@pytest.mark.parametrize("n,b,nb", [(64, 8, 16), (130, 32, 64), (200, 32, 128), (97, 16, 32)])
def test_sy2sb_band_and_similarity(n, b, nb):
    A = O.gen_symmetric(n, n).astype(np.float64)
    Ab, W, Y = O.sy2sb(A, b, nb)
    assert np.abs(np.tril(Ab, -b - 1)).max() == 0.0
    B = O.band_of(Ab, b)
    Q = O.q_sbr(W, Y, b)
    np.testing.assert_allclose(Q.T @ A @ Q, B, atol=1e-12)
    np.testing.assert_allclose(Q.T @ Q, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(np.linalg.eigvalsh(B), np.linalg.eigvalsh(A), atol=1e-12)
    # Q_SBR fixes e_1 (reflectors start below row b)
    np.testing.assert_allclose(Q[:, 0], np.eye(n)[:, 0], atol=1e-15)


@pytest.mark.parametrize("n,seed", [(40, 3), (64, 4)])
def test_sb2st_matches_lapack_sytrd(n, seed):
    A = O.gen_symmetric(n, seed).astype(np.float64)
    b = 8
    Ab, _, _ = O.sy2sb(A, b, 2 * b)
    D, E, U = O.sb2st(O.band_of(Ab, b), b)
    np.testing.assert_allclose(D, G[f"sytrd_d_n{n}_s{seed}"], atol=1e-11)
    np.testing.assert_allclose(np.abs(E), G[f"sytrd_abse_n{n}_s{seed}"], atol=1e-11)


@pytest.mark.parametrize("n,b", [(50, 8), (130, 32), (100, 64)])
def test_sb2st_similarity_and_reflector_layout(n, b):
    A = O.gen_symmetric(n, 100 + n).astype(np.float64)
    Ab, _, _ = O.sy2sb(A, b, 2 * b)
    B = O.band_of(Ab, b)
    D, E, U = O.sb2st(B, b)
    T = np.diag(D) + np.diag(E, 1) + np.diag(E, -1)
    Qb = O.q_bc(U, n, b)
    np.testing.assert_allclose(Qb.T @ B @ Qb, T, atol=1e-12)
    np.testing.assert_allclose(sl.eigvalsh_tridiagonal(D, E), np.linalg.eigvalsh(A), atol=1e-12)
    # every stored reflector has ||u||^2 = 2 (or is zero) and sits in its slot
    for s, p, ln in O.chase_steps(n, b):
        u = U[p:p + ln, s]
        assert abs(u @ u - 2.0) < 1e-12 or not np.any(u)
    assert not np.any(U[n:, :])


def test_bc_back_orders_are_inverse_transposes():
    n, b = 90, 16
    A = O.gen_symmetric(n, 9).astype(np.float64)
    Ab, _, _ = O.sy2sb(A, b, 2 * b)
    _, _, U = O.sb2st(O.band_of(Ab, b), b)
    Qbc = O.q_bc(U, n, b)
    X = np.random.default_rng(1).standard_normal((n, 5))
    np.testing.assert_allclose(O.bc_back(X, U, n, b, 1), Qbc @ X, atol=1e-12)
    np.testing.assert_allclose(O.bc_back(X, U, n, b, 0), Qbc.T @ X, atol=1e-12)


def test_wy_from_slide_and_q_from_wy_are_products():
    rng = np.random.default_rng(3)
    M, slide, nblk = 40, 4, 5
    W = rng.standard_normal((M, slide * nblk))
    Y = rng.standard_normal((M, slide * nblk))
    want = np.eye(M)
    for k in range(nblk):
        s = slice(k * slide, (k + 1) * slide)
        want = want @ (np.eye(M) - W[:, s] @ Y[:, s].T)
    Wm = O.wy_from_slide(W, Y, slide)
    np.testing.assert_allclose(O.q_from_wy(Wm, Y), want, atol=1e-10)


def test_tridiagonal_oracle_against_fixtures_and_closed_forms():
    k = 10
    d = np.abs(np.arange(-k, k + 1)).astype(np.float64)
    np.testing.assert_allclose(O.tridiag_eig(d, np.ones(2 * k), vectors=False), G["wilkinson21_w"], atol=1e-13)
    np.testing.assert_allclose(O.tridiag_eig(G["cluster128_d"], G["cluster128_e"], vectors=False),
                               G["cluster128_w"], atol=1e-13)
    # closed form: the (1, 2, 1) Toeplitz tridiagonal, lambda_j = 2 + 2 cos(j pi/(n+1))
    n = 33
    w = O.tridiag_eig(2 * np.ones(n), np.ones(n - 1), vectors=False)
    want = np.sort(2 + 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))
    np.testing.assert_allclose(w, want, atol=1e-13)
    # closed form: diagonal T has its diagonal as spectrum, Z a permutation
    rng = np.random.default_rng(3)
    d = rng.uniform(-1, 1, 40)
    w, Z = O.tridiag_eig(d, np.zeros(39))
    np.testing.assert_allclose(w, np.sort(d), atol=1e-15)
    np.testing.assert_allclose(np.abs(Z).sum(axis=0), 1.0, atol=1e-15)
    # closed form: decoupled 2x2 blocks [[a, e], [e, c]] have
    # lambda = (a + c)/2 -+ sqrt(((a - c)/2)^2 + e^2)
    a, c, e = rng.uniform(-1, 1, (3, 16))
    d2 = np.stack([a, c], axis=1).ravel()
    e2 = np.zeros(31)
    e2[0::2] = e
    m, r = (a + c) / 2, np.hypot((a - c) / 2, e)
    want = np.sort(np.concatenate([m - r, m + r]))
    np.testing.assert_allclose(O.tridiag_eig(d2, e2, vectors=False), want, atol=1e-14)


@pytest.mark.parametrize("n,b,nb", [(64, 32, 32), (130, 32, 64), (200, 16, 64)])
def test_oracle_pipeline_matches_dsyevd(n, b, nb):
    A = O.gen_symmetric(n, 7 + n).astype(np.float64)
    w, Z = O.syevd(A, b, nb)
    np.testing.assert_allclose(w, O.dsyevd(A, vectors=False), atol=1e-12)
    assert O.backward_error(A, w, Z) < 1e-15
    assert O.orthogonality(Z) < 1e-14


def test_golden_dsyevd_fixture_is_lapack():
    for n, seed in [(64, 0), (130, 1)]:
        A = O.gen_symmetric(n, seed).astype(np.float64)
        np.testing.assert_allclose(O.dsyevd(A, vectors=False), G[f"syevd_w_n{n}_s{seed}"], atol=1e-12)


def test_planted_mistake_is_caught():
    """A dropped term in the ZY update (the -1/2 Y W^T Z correction) must
    break the band similarity that the SBR tests check."""
    n, b, nb = 64, 8, 16
    A = O.gen_symmetric(n, 1).astype(np.float64)
    orig = O.panel_qr

    Ab, W, Y = O.sy2sb(A, b, nb)
    good = np.abs(np.linalg.eigvalsh(O.band_of(Ab, b)) - np.linalg.eigvalsh(A)).max()

    def broken_panel(P):
        V, T, R, tau = orig(P)
        return V, -T, R, tau  # wrong sign of the block reflector

    O.panel_qr = broken_panel
    try:
        Ab2, _, _ = O.sy2sb(A, b, nb)
    finally:
        O.panel_qr = orig
    bad = np.abs(np.linalg.eigvalsh(O.band_of(Ab2, b)) - np.linalg.eigvalsh(A)).max()
    assert good < 1e-12 and bad > 1e-3


def test_gaussian_generator_moments():
    """kind = 1: off-diagonal entries (x + y)/2 of two independent N(0, 1)
    draws are N(0, 1/2); the diagonal entries are N(0, 1).  Moments and a
    Kolmogorov-Smirnov test pin the Box-Muller mapping (a wrong scale, a
    missing sqrt(-2 log u) factor or a uniform in place of a normal fails)."""
    import scipy.stats as st

    n = 768
    A = O.gen_symmetric(n, 5, 1).astype(np.float64)
    off = A[np.tril_indices(n, -1)]
    N = off.size
    assert abs(off.mean()) < 5.0 * np.sqrt(0.5 / N)
    assert abs(off.var() - 0.5) < 5.0 * 0.5 * np.sqrt(2.0 / N)
    kurt = ((off - off.mean()) ** 4).mean() / off.var() ** 2
    assert abs(kurt - 3.0) < 0.1
    assert st.kstest(off, "norm", args=(0.0, np.sqrt(0.5))).pvalue > 1e-3
    assert st.kstest(np.diag(A), "norm", args=(0.0, 1.0)).pvalue > 1e-3
    assert np.array_equal(A, A.T)


def test_metric_definitions_known_deviation():
    """The metrics against deviations of known size, so that a dropped
    normalisation (the 1/n of either metric, the 1/||A||_F of the backward
    error) or a norm other than Frobenius fails."""
    n, dlt = 50, 1e-3
    # Q = (1 + d) I: Q^T Q - I = (2d + d^2) I, ||.||_F = (2d + d^2) sqrt(n)
    Q = (1.0 + dlt) * np.eye(n)
    np.testing.assert_allclose(O.orthogonality(Q), (2 * dlt + dlt ** 2) * np.sqrt(n) / n, rtol=1e-12)
    # a single off-diagonal unit: ||Q^T Q - I||_F^2 = 2 + 1 (two symmetric
    # entries of 1 and one diagonal entry of 1)
    Q = np.eye(n)
    Q[3, 7] = 1.0
    np.testing.assert_allclose(O.orthogonality(Q), np.sqrt(3.0) / n, rtol=1e-12)
    # A = diag(a), Z = I, w = a + d e_k: A Z - Z diag(w) = -d e_k e_k^T
    rng = np.random.default_rng(1)
    a = rng.uniform(-2, 2, n)
    w = a.copy()
    w[11] += dlt
    np.testing.assert_allclose(O.backward_error(np.diag(a), w, np.eye(n)), dlt / (n * np.linalg.norm(a)), rtol=1e-12)
