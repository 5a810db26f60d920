"""CPU oracle of the EVD hot path - TEST INFRASTRUCTURE ONLY.

This module restates, in FP64 NumPy, the algorithm of each stage of the
reference (hansheng1001/EVD4SC2025) so that the GPU library can be checked
against it on the same seeded inputs.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import it; the product never does.

Pins (what each function is checked against in tests/test_oracle.py):
  philox4x32 ............ Random123 known-answer vectors (published KATs)
  panel_qr .............. LAPACK dgeqrf (scipy) R, V, tau; dlarft identity
                          I - V T V^T = H_1 ... H_k by brute force
  syr2k ................. textbook dsyr2k definition (brute force)
  sy2sb ................. eig(B) = dsyevd(A); Q^T A Q = B; band structure
  sb2st ................. eig(T) = dsyevd(B); D, |E| = dsytrd(A) for tiny n
                          (implicit-Q theorem: both reductions fix e_1);
                          Q_BC^T B Q_BC = T by brute-force reflector products
  bc_back ............... brute-force sequential products (both orders)
  wy_from_slide /
  q_from_wy ............. product of (I - W_k Y_k^T) by brute force
  tridiag_eig ........... LAPACK dstemr/dsyevd (scipy); third-party routine,
                          no reference test exists for it
  syevd ................. dsyevd eigenvalues; residual and orthogonality
"""
from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------------------
# Philox4x32-10 (Salmon et al., SC'11) - the device generator's stream
# ---------------------------------------------------------------------------
_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(ctr: np.ndarray, key: tuple[int, int]) -> np.ndarray:
    """ctr: (N, 4) uint32, key: (k0, k1) -> (N, 4) uint32 after 10 rounds."""
    c = [ctr[:, i].astype(np.uint64) for i in range(4)]
    k0, k1 = key[0] & 0xFFFFFFFF, key[1] & 0xFFFFFFFF
    for _ in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return np.stack([x.astype(np.uint32) for x in c], axis=1)


def _u53(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return ((a >> 5).astype(np.float64) * 67108864.0 + (b >> 6).astype(np.float64)) / 9007199254740992.0


def rand_entries(idx: np.ndarray, stream: int, seed: int, kind: int) -> np.ndarray:
    idx = idx.astype(np.uint64)
    ctr = np.zeros((idx.size, 4), dtype=np.uint32)
    ctr[:, 0] = (idx & _MASK).astype(np.uint32)
    ctr[:, 1] = (idx >> np.uint64(32)).astype(np.uint32)
    ctr[:, 2] = stream
    w = philox4x32(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF))
    if kind == 0:
        return 2.0 * _u53(w[:, 0], w[:, 1]) - 1.0
    u1 = _u53(w[:, 0], w[:, 1]) + 0.5 / 9007199254740992.0
    u2 = _u53(w[:, 2], w[:, 3])
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def gen_symmetric(n: int, seed: int, kind: int = 0) -> np.ndarray:
    """(X + X^T)/2 rounded to FP32, X[r, c] from counter r + c*n (the device
    generator's mapping)."""
    r, c = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = rand_entries((r + c * n).ravel(), 0, seed, kind).reshape(n, n)
    return (0.5 * (x + x.T)).astype(np.float32)


# ---------------------------------------------------------------------------
# Panel QR - api/panelQR.cu:20-246, util/inc/kernelQR.h:267-311 (reflector
# convention R_jj = -sign(x_1)||x||, sign(0)=+), LAPACK dgeqrf/dlarft form
# ---------------------------------------------------------------------------

def panel_qr(A: np.ndarray):
    """Householder QR of the m x nc panel.  Returns (V, T, R, tau) with
    A = (I - V T V^T) [R; 0], V unit lower (m x k), T upper (k x k)."""
    A = np.array(A, dtype=np.float64, copy=True)
    m, nc = A.shape
    k = min(m, nc)
    tau = np.zeros(k)
    for j in range(k):
        alpha = A[j, j]
        x = A[j + 1:, j]
        sigma = float(x @ x)
        if sigma == 0.0:
            tau[j], beta, scale = 0.0, alpha, 0.0
        else:
            nrm = np.sqrt(alpha * alpha + sigma)
            beta = -nrm if alpha >= 0.0 else nrm
            tau[j] = (beta - alpha) / beta
            scale = 1.0 / (alpha - beta)
        A[j + 1:, j] = x * scale
        A[j, j] = beta
        v = np.concatenate([[1.0], A[j + 1:, j]])
        w = tau[j] * (v @ A[j:, j + 1:])
        A[j:, j + 1:] -= np.outer(v, w)
    V = np.tril(A[:, :k], -1) + np.eye(m, k)
    R = np.triu(A[:k, :])
    T = larft(V, tau)
    return V, T, R, tau


def larft(V: np.ndarray, tau: np.ndarray) -> np.ndarray:
    """dlarft forward/columnwise: I - V T V^T = H_1 ... H_k."""
    k = V.shape[1]
    T = np.zeros((k, k))
    for j in range(k):
        T[j, j] = tau[j]
        if j:
            T[:j, j] = -tau[j] * (T[:j, :j] @ (V[:, :j].T @ V[:, j]))
    return T


def syr2k(alpha, A, B, beta, C):
    """api/TCblas/tc_ozimmu_syr2k.cu:209-341 semantics (full symmetric result)."""
    return alpha * (A @ B.T + B @ A.T) + beta * C


# ---------------------------------------------------------------------------
# SBR - api/zy_zy_sy2sb.cu:24-486 (ZY lazy two-sided update)
# ---------------------------------------------------------------------------

def sy2sb(A: np.ndarray, b: int, nb: int):
    """Returns (A_band_full, W, Y): the lower band of A_band_full is
    B = Q^T A Q with Q = prod_p (I - W_p Y_p^T) (panel order)."""
    A = np.array(A, dtype=np.float64, copy=True)
    A = np.tril(A) + np.tril(A, -1).T
    n = A.shape[0]
    W = np.zeros((n, n))
    Y = np.zeros((n, n))
    P = nb // b
    c0 = 0
    while n - c0 - b > 1:
        Z = np.zeros((n, nb))
        p = 0
        while True:
            c = c0 + p * b
            r0 = c + b
            m = n - r0
            V, T, R, _ = panel_qr(A[r0:, c:c + b])
            k = V.shape[1]
            Wp = np.zeros((m, b))
            Vp = np.zeros((m, b))
            Wp[:, :k] = V @ T
            Vp[:, :k] = V
            W[r0:, c:c + b] = Wp
            Y[r0:, c:c + b] = Vp
            Rfull = np.zeros((m, b))
            Rfull[: R.shape[0], :] = R
            A[r0:, c:c + b] = Rfull
            Zp = A[r0:, r0:] @ Wp
            if p > 0:
                Yq = Y[r0:, c0:c]
                Zq = Z[r0:, : p * b]
                Zp -= Yq @ (Zq.T @ Wp) + Zq @ (Yq.T @ Wp)
            Zp -= 0.5 * Vp @ (Wp.T @ Zp)
            Z[r0:, p * b:(p + 1) * b] = Zp
            kk = (p + 1) * b
            Ya = Y[r0:, c0:c0 + kk]
            Za = Z[r0:, :kk]
            more = (p + 1 < P) and (n - (r0 + b) > 1)
            if more:
                A[r0:, r0:r0 + b] -= Ya @ Za[:b].T + Za @ Ya[:b].T
                p += 1
                continue
            T2 = A[r0:, r0:] - Ya @ Za.T - Za @ Ya.T
            A[r0:, r0:] = np.tril(T2) + np.tril(T2, -1).T
            break
        c0 += (p + 1) * b
    return A, W, Y


def band_of(A: np.ndarray, b: int) -> np.ndarray:
    """Dense symmetric band matrix from the lower band of A."""
    n = A.shape[0]
    L = np.tril(A)
    L[np.subtract.outer(np.arange(n), np.arange(n)) > b] = 0.0
    return L + np.tril(L, -1).T


def q_sbr(W: np.ndarray, Y: np.ndarray, b: int) -> np.ndarray:
    n = W.shape[0]
    Q = np.eye(n)
    for c in range(0, n, b):
        Wc, Yc = W[:, c:c + b], Y[:, c:c + b]
        if np.any(Yc):
            Q = Q @ (np.eye(n) - Wc @ Yc.T)
    return Q


# ---------------------------------------------------------------------------
# Bulge chasing - util/inc/bugle_chasingV10.h:21-392 (reflector
# u = (x/||x|| + sign(x_1) e_1)/sqrt(|.|), H = I - u u^T, stored in
# U[p:p+len, s]); sequential sweep-major order (the order the pipelined
# kernel is equivalent to)
# ---------------------------------------------------------------------------

def chase_steps(n: int, b: int):
    """(s, p, len) of every reflector in generation order."""
    out = []
    for s in range(max(n - 2, 0)):
        p = s + 1
        ln = min(b, n - p)
        while True:
            out.append((s, p, ln))
            ln2 = max(0, min(b, n - p - ln))
            if ln2 <= 1:
                break
            p += ln
            ln = ln2
    return out


def sb2st(B: np.ndarray, b: int):
    """B dense symmetric band (bandwidth b).  Returns (D, E, U) with U of
    shape (n + b, n) in the "noBandU" layout."""
    A = np.array(B, dtype=np.float64, copy=True)
    n = A.shape[0]
    U = np.zeros((n + b, n))
    for s in range(max(n - 2, 0)):
        p = s + 1
        ln = min(b, n - p)
        col = s
        while True:
            x = A[p:p + ln, col].copy()
            nrm = np.sqrt(x @ x)
            if nrm > 0.0:
                u = x / nrm
                u[0] += 1.0 if u[0] >= 0.0 else -1.0
                u /= np.sqrt(abs(u[0]))
                beta = -nrm if x[0] >= 0.0 else nrm
            else:
                u = np.zeros(ln)
                beta = x[0]
            U[p:p + ln, s] = u
            # rows/columns p..p+ln have their nonzeros within 2b of the diagonal
            lo, hi = max(0, p - 2 * b), min(n, p + ln + 2 * b)
            A[p:p + ln, lo:hi] -= np.outer(u, u @ A[p:p + ln, lo:hi])
            A[lo:hi, p:p + ln] -= np.outer(A[lo:hi, p:p + ln] @ u, u)
            A[p:p + ln, col] = 0.0
            A[col, p:p + ln] = 0.0
            A[p, col] = A[col, p] = beta
            ln2 = max(0, min(b, n - p - ln))
            if ln2 <= 1:
                break
            col = p
            p += ln
            ln = ln2
    return np.diag(A).copy(), np.diag(A, -1).copy(), U


def bc_back(Q: np.ndarray, U: np.ndarray, n: int, b: int, order: int) -> np.ndarray:
    """api/BC_backTransV8_10_noBandU.cu:60-313.  order 0: Q <- H_K..H_1 Q;
    order 1: Q <- H_1..H_K Q."""
    Q = np.array(Q, dtype=np.float64, copy=True)
    steps = chase_steps(n, b)
    if order == 1:
        steps = steps[::-1]
    for s, p, ln in steps:
        u = U[p:p + ln, s]
        Q[p:p + ln] -= np.outer(u, u @ Q[p:p + ln])
    return Q


def q_bc(U: np.ndarray, n: int, b: int) -> np.ndarray:
    """Q_BC = H_1 ... H_K (generation order)."""
    return bc_back(np.eye(n), U, n, b, 1)


# ---------------------------------------------------------------------------
# WY helpers - api/computerWYFromSlide.cu:5-25, api/computerQFromWY.cu:5-18
# ---------------------------------------------------------------------------

def wy_from_slide(W: np.ndarray, Y: np.ndarray, slide: int) -> np.ndarray:
    W = np.array(W, dtype=np.float64, copy=True)
    M, N = W.shape
    for i in range(2 * slide, N + 1, slide):
        work = Y[:, : i - slide].T @ W[:, i - slide:i]
        W[:, i - slide:i] -= W[:, : i - slide] @ work
    return W


def q_from_wy(W: np.ndarray, Y: np.ndarray) -> np.ndarray:
    return np.eye(W.shape[0]) - W @ Y.T


# ---------------------------------------------------------------------------
# Tridiagonal eigensolver - the reference calls MKL LAPACKE_dstedc
# (src/EVD/myEVD.cu:529); the oracle is LAPACK through scipy.
# ---------------------------------------------------------------------------

def tridiag_eig(D: np.ndarray, E: np.ndarray, vectors: bool = True):
    import scipy.linalg as sl

    if vectors:
        return sl.eigh_tridiagonal(D, E)
    return sl.eigh_tridiagonal(D, E, eigvals_only=True)


def dsyevd(A: np.ndarray, vectors: bool = True):
    import scipy.linalg as sl

    return sl.eigh(np.asarray(A, dtype=np.float64), driver="evd", eigvals_only=not vectors)


# ---------------------------------------------------------------------------
# Metrics - util/tool/checkMetric.cu:5-25, src/EVD/myEVD.cu:598-639
# ---------------------------------------------------------------------------

def orthogonality(Q: np.ndarray) -> float:
    Q = np.asarray(Q, dtype=np.float64)
    n = Q.shape[1]
    return float(np.linalg.norm(Q.T @ Q - np.eye(n)) / n)


def backward_error(A: np.ndarray, w: np.ndarray, Z: np.ndarray) -> float:
    A = np.asarray(A, dtype=np.float64)
    Z = np.asarray(Z, dtype=np.float64)
    n = A.shape[0]
    return float(np.linalg.norm(A @ Z - Z * w[None, :]) / (n * np.linalg.norm(A)))


# ---------------------------------------------------------------------------
# Whole pipeline (src/EVD/myEVD.cu:251-571) - SBR -> BC -> tridiagonal ->
# back transformation, all FP64
# ---------------------------------------------------------------------------

def syevd(A: np.ndarray, b: int, nb: int, vectors: bool = True):
    n = A.shape[0]
    Ab, W, Y = sy2sb(A, b, nb)
    B = band_of(Ab, b)
    D, E, U = sb2st(B, b)
    if not vectors:
        return tridiag_eig(D, E, vectors=False)
    w, Zt = tridiag_eig(D, E)
    Z = bc_back(Zt, U, n, b, 1)
    Z = q_sbr(W, Y, b) @ Z
    return w, Z
