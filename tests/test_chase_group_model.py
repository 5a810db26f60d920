"""CPU model of the grouped bulge chase (chase.hip, chase_group_kernel): a
workgroup owns D consecutive sweeps, one wave per sweep, stepping in
lockstep (sweep s0 + i does step t - 2i at time step t) on a window of band
columns held in LDS (column c in slot c mod W).  The I/O wave loads the front
sweep's new columns from HBM three steps ahead (into registers) once the
previous group has published enough progress, and writes them into the
window one step ahead (synthesising the known zeros below the
early-published beta of the last column); it writes the columns the group's
last sweep has finished back to HBM and publishes the progress one step
later, after the stores have completed.

The model checks, for random interleavings of the groups' time steps:
  * the result equals the sequential chase (oracle) - D, E and the reflectors;
  * within a time step the waves' LDS read/write sets are disjoint;
  * every column a wave touches is resident in its slot, and a slot is only
    overwritten after its previous column was written back in an earlier step
    (a column on which the last sweep's step rows start is written back down
    to its beta only: the fill below is being zeroed in that same step);
  * no group waits forever.
A rule that skips the synthesised zeros, or a window one step too small,
must fail one of these checks."""
import random

import numpy as np
import pytest

from oracle import evd_oracle as O

INF = 1 << 40


def _refl(x):
    nrm = np.sqrt(x @ x)
    if nrm > 0:
        u = x / nrm
        u[0] += 1.0 if x[0] >= 0 else -1.0
        u /= np.sqrt(abs(u[0]))
        return u, (-nrm if x[0] >= 0 else nrm)
    return np.zeros_like(x), x[0]


def sweep_steps(n, b, s):
    """(p, len, len2) of every step of sweep s."""
    p, ln, out = s + 1, min(b, n - s - 1), []
    while True:
        ln2 = max(0, min(b, n - p - ln))
        out.append((p, ln, ln2))
        if ln2 <= 1:
            return out
        p, ln = p + ln, ln2


class Conflict(AssertionError):
    pass


class Group:
    def __init__(self, m, g):
        self.m, self.g = m, g
        n, b, D = m.n, m.b, m.D
        self.s0 = g * D
        self.nsw = min(D, n - 2 - self.s0)
        self.steps = [sweep_steps(n, b, self.s0 + i) for i in range(self.nsw)]
        F = self.steps[0]
        self.cols = []  # columns the front needs first at step j
        for j, (p, ln, _) in enumerate(F):
            lo = self.s0 if j == 0 else p
            # the front's last step also brings in the column (at most one)
            # past its window that the later sweeps' last steps reach
            self.cols.append((lo, n if j == len(F) - 1 else min(n, p + ln)))
            # the kernel's fixed chunk buffer holds b + 2 columns
            assert self.cols[-1][1] - self.cols[-1][0] <= b + 2
        self.need = [p + 2 * b - 1 for (p, _, _) in F]
        self.t = -3
        self.known = -1
        self.regs = {}  # step j -> {column: data} loaded, not yet in LDS
        self.win = {}  # slot -> [column, data(2b), retired_at]
        self.st = [dict(first=True) for _ in range(self.nsw)]
        self.wb = self.s0  # next column to write back
        self.pending = None  # progress to publish at the next step
        self.done = False

    # ---- window access with conflict tracking
    def _slot(self, c):
        return c % self.m.W

    def _check_resident(self, c):
        e = self.win.get(self._slot(c))
        if e is None or e[0] != c:
            raise Conflict(f"group {self.g} t={self.t}: column {c} not resident")
        return e

    def rd(self, who, r, c):
        e = self._check_resident(c)
        self.acc.setdefault(who, [set(), set()])[0].add((r, c))
        return e[1][r - c]

    def wr(self, who, r, c, v):
        assert 0 <= r - c < 2 * self.m.b
        e = self._check_resident(c)
        self.acc.setdefault(who, [set(), set()])[1].add((r, c))
        e[1][r - c] = v

    # ---- one time step of the whole workgroup (all waves, then the barrier)
    def runnable(self):
        j = self.t + 3  # the I/O wave issues the loads of front step t + 3
        if self.g == 0 or j >= len(self.cols):
            return True
        return max(self.known, self.m.com[self.g - 1]) >= self.need[j]

    def step(self):
        m, n, b = self.m, self.m.n, self.m.b
        t = self.t
        self.acc = {}
        # -- store wave: publish last step's write-back, then write back the
        # columns the last sweep finished at t - 1
        if self.pending is not None:
            m.com[self.g] = self.pending
            self.pending = None
            if m.com[self.g] == INF:
                self.done = True
                return
        tail = self.nsw - 1
        k = t - 1 - 2 * tail
        if 0 <= k < len(self.steps[tail]):
            P, ln, ln2 = self.steps[tail][k]
            last = k == len(self.steps[tail]) - 1
            # column P is written down to its early beta only: the last sweep
            # zeroes the fill below the beta during this same step (the next
            # group's load chunks end on these columns and synthesise the zeros)
            hi = self.loaded_hi if last else P + 1
            for c in range(self.wb, hi):
                dmax = 2 * b if (last or c < P) else b + 1
                for d in range(dmax):
                    if c + d < n:
                        m.hbm[d, c] = self.rd("store", c + d, c)
                self.win[self._slot(c)][2] = t  # retired at this step
            self.wb = hi
            self.pending = INF if last else P + ln
        # -- I/O wave: LDS <- columns of front step t + 1 (loaded at t - 2),
        # then issue the loads of front step t + 3
        j = t + 1
        if 0 <= j < len(self.cols) and j in self.regs:
            for c, data in self.regs.pop(j).items():
                old = self.win.get(self._slot(c))
                if old is not None and (old[2] is None or old[2] >= t):
                    raise Conflict(f"group {self.g} t={t}: slot of column {old[0]} reused before write-back")
                self.win[self._slot(c)] = [c, data, None]
                self.acc.setdefault("load", [set(), set()])[1].update((c + d, c) for d in range(2 * b))
            self.loaded_hi = self.cols[j][1]
        j = t + 3
        if 0 <= j < len(self.cols):
            if self.g > 0:
                self.known = max(self.known, m.com[self.g - 1])
                assert self.known >= self.need[j]
            lo, hi = self.cols[j]
            p = self.steps[0][j][0]
            regs = {}
            for c in range(lo, hi):
                col = np.array([m.hbm[d, c] if c + d < n else 0.0 for d in range(2 * b)])
                if m.synth and c == p + b - 1:
                    col[b + 1:] = 0.0  # zeros the previous sweep writes below its early beta
                regs[c] = col
            self.regs[j] = regs
        # -- worker waves
        for i in range(self.nsw):
            js = t - 2 * i
            if 0 <= js < len(self.steps[i]):
                self._work(i, js)
        # -- speculative poll for the next step
        if self.g > 0:
            self.known = max(self.known, m.com[self.g - 1])
        # the waves' accesses within the step must be disjoint
        keys = list(self.acc)
        for a in keys:
            for c in keys:
                if a < c:
                    ra, wa = self.acc[a]
                    rc, wc = self.acc[c]
                    if wa & (rc | wc) or wc & ra:
                        raise Conflict(f"group {self.g} t={t}: {a} and {c} overlap")
        self.t += 1

    def _work(self, i, j):
        m, b = self.m, self.m.b
        who = f"w{i}"
        s = self.s0 + i
        p, ln, ln2 = self.steps[i][j]
        x = self.st[i]
        if x["first"]:
            col = np.array([self.rd(who, s + 1 + r, s) for r in range(ln)])
            x["u"], x["beta"] = _refl(col)
            m.U[p:p + ln, s] = x["u"]
            x["cB1"], x["colB1"] = s, 1
        u = x["u"]
        # B1 <- H B1 (column 0 becomes beta e1, written only at the first step)
        cB1, colB1 = x["cB1"], x["colB1"]
        for c in range(1, colB1):
            v = np.array([self.rd(who, p + r, cB1 + c) for r in range(ln)])
            v = v - u * (u @ v)
            for r in range(ln):
                self.wr(who, p + r, cB1 + c, v[r])
        for r in range(ln):
            if r == 0 and not x["first"]:
                continue
            self.wr(who, p + r, cB1, x["beta"] if r == 0 else 0.0)
        # S <- H S H, N <- N H
        S = np.array([[self.rd(who, p + max(r, k), p + min(r, k)) for k in range(ln)] for r in range(ln)])
        N = np.array([[self.rd(who, p + ln + r, p + k) for k in range(ln)] for r in range(ln2)]).reshape(ln2, ln)
        y = S @ u
        w = y - 0.5 * (u @ y) * u
        S = S - np.outer(u, w) - np.outer(w, u)
        N = N - np.outer(N @ u, u)
        for r in range(ln):
            for k in range(r + 1):
                self.wr(who, p + r, p + k, S[r, k])
        for r in range(ln2):
            for k in range(ln):
                self.wr(who, p + ln + r, p + k, N[r, k])
        if ln2 <= 1:
            return
        col = np.array([self.rd(who, p + ln + r, p) for r in range(ln2)])
        u2, beta2 = _refl(col)
        m.U[p + ln:p + ln + ln2, s] = u2
        self.wr(who, p + ln, p, beta2)
        x.update(first=False, u=u2, beta=beta2, cB1=p, colB1=ln)


class Model:
    def __init__(self, Bd, b, D, W, G, seed, synth=True):
        n = Bd.shape[0]
        self.n, self.b, self.D, self.W, self.synth = n, b, D, W, synth
        self.hbm = np.zeros((2 * b, n))
        for c in range(n):
            for d in range(min(2 * b, n - c)):
                self.hbm[d, c] = Bd[c + d, c]
        self.U = np.zeros((n + b, n))
        ngroups = (n - 2 + D - 1) // D
        self.com = [-1] * ngroups
        rng = random.Random(seed)
        ticket, active = 0, []
        while True:
            while len(active) < G and ticket < ngroups:
                active.append(Group(self, ticket))
                ticket += 1
            if not active:
                break
            run = [x for x in active if x.runnable()]
            if not run:
                raise Conflict("no group can make progress")
            grp = rng.choice(run)
            grp.step()
            if grp.done:
                active.remove(grp)
        self.D_, self.E_ = self.hbm[0].copy(), self.hbm[1, :n - 1].copy()


def _band(n, b, seed):
    A = O.gen_symmetric(n, seed, 0).astype(np.float64)
    return np.tril(np.triu(A, -b), b)


def _err(model, ref):
    D0, E0, U0 = ref
    return max(np.abs(model.D_ - D0).max(), np.abs(model.E_ - E0).max(), np.abs(model.U - U0).max())


@pytest.mark.parametrize("n,b,D", [(40, 4, 3), (67, 5, 3), (50, 3, 2), (61, 4, 4), (23, 4, 3), (9, 4, 3), (4, 4, 3),
                                   (70, 8, 3), (45, 8, 1), (57, 6, 5)])
def test_group_chase_matches_sequential_chase(n, b, D):
    B = _band(n, b, 5)
    ref = O.sb2st(B, b)
    W = (2 * D + 2) * b - D + 1  # the kernel's bound (its window is the next power of two)
    for seed in range(3):
        m = Model(B, b, D, W, G=4, seed=seed)
        assert _err(m, ref) < 1e-11


def test_group_chase_model_catches_unsafe_variants():
    n, b, D = 50, 8, 3
    B = _band(n, b, 5)
    ref = O.sb2st(B, b)
    # without the synthesised zeros the later sweeps of a group read stale fill
    # (for b >= D + 2 the fill an earlier group left in those columns is nonzero)
    assert _err(Model(B, b, D, 64, G=4, seed=0, synth=False), ref) > 1e-6
    # the window needs (2D + 2) b - D + 1 columns (the kernel's bound); a
    # window a few columns short reuses a slot before its column is written back
    B = _band(70, b, 5)
    Model(B, b, D, (2 * D + 2) * b - D + 1, G=4, seed=0)
    with pytest.raises(Conflict):
        Model(B, b, D, (2 * D + 2) * b - 2 * D, G=4, seed=0)
