"""CPU model of the bulge chase's inter-sweep protocol (chase.hip): sweeps
run concurrently, each step split into a read phase (S and N windows, B1
write-back) and a write phase (S, N at the end, the next reflector's beta,
progress word), interleaved at random subject to the progress rule.  The
rule used on the GPU - sweep s may start row p once sweep s-1 published a
next-row >= p + 2b - 1, with beta of the next step published early - must
give the oracle's (sequential) result; a one-step-earlier rule, or the same
rule without the early beta, must not."""
import random

import numpy as np
import pytest

from oracle import evd_oracle as O


def _refl(x):
    nrm = np.sqrt(x @ x)
    if nrm > 0:
        u = x / nrm
        u[0] += 1.0 if x[0] >= 0 else -1.0
        u /= np.sqrt(abs(u[0]))
        return u, (-nrm if x[0] >= 0 else nrm)
    return np.zeros_like(x), x[0]


def _chase(Bd, b, seed, lag, early_beta=True, window=6):
    """lag 1: the GPU rule (p + 2b - 1); lag 0: the reference rule (p + 2b)."""
    n = Bd.shape[0]
    bd = np.zeros((2 * b, n))
    for c in range(n):
        for d in range(min(2 * b, n - c)):
            bd[d, c] = Bd[c + d, c]
    get = lambda r, c: bd[r - c, c]

    def put(r, c, v):
        assert 0 <= r - c < 2 * b
        bd[r - c, c] = v

    U = np.zeros((n + b, n))
    INF = 1 << 40
    com = [0] * max(n - 2, 0)
    rng = random.Random(seed)
    nsw, st, ticket, done = n - 2, {}, 0, 0
    while done < nsw:
        cand = [("new", ticket)] if ticket < nsw and len(st) < window else []
        for s, x in st.items():
            if x["ph"] == "wait":
                if s == 0 or com[s - 1] >= x["p"] + 2 * b - lag:
                    cand.append(("start", s))
            else:
                cand.append(("finish", s))
        kind, s = rng.choice(cand)
        if kind == "new":
            st[s] = dict(ph="wait", p=s + 1, len=min(b, n - s - 1), first=True, cB1=s, colB1=1)
            ticket += 1
            continue
        x = st[s]
        p, ln = x["p"], x["len"]
        if kind == "start":
            if x["first"]:
                col = np.array([get(s + 1 + i, s) for i in range(ln)])
                x["B1"] = col.reshape(ln, 1)
                x["u"], x["beta"] = _refl(col.copy())
                U[p:p + ln, s] = x["u"]
            ln2 = max(0, min(b, n - p - ln))
            S = np.array([[get(p + max(i, k), p + min(i, k)) for k in range(ln)] for i in range(ln)])
            N = np.array([[get(p + ln + i, p + k) for k in range(ln)] for i in range(ln2)]).reshape(ln2, ln)
            u, B1 = x["u"], x["B1"]
            nb1 = B1 - np.outer(u, B1.T @ u)
            nb1[:, 0] = 0.0
            nb1[0, 0] = x["beta"]
            for i in range(ln):
                for c in range(x["colB1"]):
                    if c == 0 and i == 0 and not x["first"] and early_beta:
                        continue
                    put(p + i, x["cB1"] + c, nb1[i, c])
            x.update(S=S, N=N, len2=ln2, ph="run")
        else:
            u, S, N, ln2 = x["u"], x["S"], x["N"], x["len2"]
            y = S @ u
            w = y - 0.5 * (u @ y) * u
            S = S - np.outer(u, w) - np.outer(w, u)
            N = N - np.outer(N @ u, u)
            for i in range(ln):
                for k in range(i + 1):
                    put(p + i, p + k, S[i, k])
            if ln2 <= 1:
                for i in range(ln2):
                    for k in range(ln):
                        put(p + ln + i, p + k, N[i, k])
                com[s] = INF
                del st[s]
                done += 1
                continue
            u2, beta2 = _refl(N[:, 0].copy())
            U[p + ln:p + ln + ln2, s] = u2
            if early_beta:
                put(p + ln, p, beta2)
            com[s] = p + ln
            x.update(ph="wait", B1=N, cB1=p, colB1=ln, p=p + ln, len=ln2, first=False, u=u2, beta=beta2)
    return bd[0].copy(), bd[1, :n - 1].copy(), U


def _err(res, ref):
    return max(np.abs(a - b).max() for a, b in zip(res, ref))


@pytest.mark.parametrize("n,b", [(40, 4), (67, 5), (50, 3)])
def test_gpu_rule_matches_sequential_chase(n, b):
    A = O.gen_symmetric(n, 3, 0).astype(np.float64)
    B = np.tril(np.triu(A, -b), b)
    ref = O.sb2st(B, b)
    for seed in range(3):
        assert _err(_chase(B, b, seed, lag=1), ref) < 1e-11


def test_protocol_model_catches_unsafe_rules():
    n, b = 40, 4
    A = O.gen_symmetric(n, 3, 0).astype(np.float64)
    B = np.tril(np.triu(A, -b), b)
    ref = O.sb2st(B, b)
    # the reference's rule is safe without the early beta
    assert _err(_chase(B, b, 0, lag=0, early_beta=False), ref) < 1e-11
    # the shorter lag needs the early beta ...
    assert _err(_chase(B, b, 0, lag=1, early_beta=False), ref) > 1e-3
    # ... and one more step is unsafe even with it
    assert _err(_chase(B, b, 0, lag=b + 1), ref) > 1e-3
