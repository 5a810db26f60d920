"""Captured solves (evd_graph_*): a whole solve recorded once as a hipGraph and
replayed.  Every replay must return the eager solve's bits on the same input,
with the program's own eager work and synchronisations between replays, and
the graph must keep the context's arena in place while it lives."""
import numpy as np
import pytest

from oracle import evd_oracle as O
from tests.helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture()
def own_ctx():
    import evd4sc2025_amd as E

    c = E.Context(0)
    yield c
    c.close()


# n = 256: single-workgroup panels; n = 700 and 1024: panels of several
# workgroups (cooperative launches inside the graph), odd trailing sizes;
# n = 2304: the SBR back transformation's first block product (2272 x 2304 x
# 1024) runs on the library GEMM, which the capture records on its stream
@pytest.mark.parametrize("n,jobz", [(256, "V"), (256, "N"), (700, "V"), (1024, "V"), (1024, "N"), (2304, "V")])
def test_graph_replays_equal_eager(own_ctx, n, jobz):
    import evd4sc2025_amd as E

    torch = pytest.importorskip("torch")
    b, nb = 32, 256
    print(f"graph case n={n} jobz={jobz}", flush=True)
    mats = [O.gen_symmetric(n, 900 + s, 0) for s in range(4)]
    g = E.SolveGraph(own_ctx, to_dev(mats[0]), b=b, nb=nb, jobz=jobz)
    try:
        ref_ctx = E.Context(0)
        try:
            eager = [E.syevd(ref_ctx, to_dev(M), b=b, nb=nb, jobz=jobz) for M in mats]
            torch.cuda.synchronize()
            eager = [(to_host(r.w), to_host(r.Z) if r.Z is not None else None) for r in eager]
        finally:
            ref_ctx.close()
        # creation solved mats[0] eagerly on the graph's buffers
        torch.cuda.synchronize()
        assert np.array_equal(to_host(g.w), eager[0][0])
        # replays, with eager work on the same context between them
        for rep, q in enumerate([1, 2, 0, 3, 1]):
            r = g.solve(to_dev(mats[q]))
            w = to_host(r.w)
            assert np.array_equal(w, eager[q][0]), f"replay {rep} (matrix {q}): eigenvalues differ"
            if jobz == "V":
                assert np.array_equal(to_host(r.Z), eager[q][1]), f"replay {rep} (matrix {q}): vectors differ"
            if rep == 1:
                E.syevd(own_ctx, to_dev(mats[2]), b=b, nb=nb, jobz=jobz)
                torch.cuda.synchronize()
        # accuracy of the last replay against the FP64 oracle
        wref = np.linalg.eigvalsh(mats[1].astype(np.float64))
        assert np.abs(w - wref).max() <= 1e-5 * np.abs(wref).max()
    finally:
        g.close()


def test_graph_nonfinite_input_is_reported(own_ctx):
    import evd4sc2025_amd as E

    n = 256
    M = O.gen_symmetric(n, 5, 0)
    g = E.SolveGraph(own_ctx, to_dev(M), b=32, nb=256)
    try:
        bad = M.copy()
        bad[7, 3] = np.nan
        with pytest.raises(E.EvdError) as ei:
            g.solve(to_dev(bad))
        assert ei.value.code == 3
        # the next replay is clean and exact
        r = g.solve(to_dev(M))
        ref = E.syevd(own_ctx, to_dev(M), b=32, nb=256)
        assert np.array_equal(to_host(r.w), to_host(ref.w))
    finally:
        g.close()


def test_graph_pins_the_arena(own_ctx):
    import evd4sc2025_amd as E

    g = E.SolveGraph(own_ctx, to_dev(O.gen_symmetric(128, 1, 0)), b=32, nb=128)
    try:
        with pytest.raises(E.EvdError) as ei:
            E.syevd(own_ctx, to_dev(O.gen_symmetric(1024, 2, 0)), b=32, nb=256)
        assert ei.value.code == 5  # EVD_E_UNSUPPORTED: the arena may not move
        # a solve that fits the arena still runs
        r = E.syevd(own_ctx, to_dev(O.gen_symmetric(100, 3, 0)), b=32, nb=128)
        assert np.isfinite(to_host(r.w)).all()
    finally:
        g.close()
    r = E.syevd(own_ctx, to_dev(O.gen_symmetric(1024, 2, 0)), b=32, nb=256)
    assert np.isfinite(to_host(r.w)).all()


@pytest.mark.parametrize("count,n,jobz", [(8, 200, "V"), (5, 640, "V"), (16, 96, "N")])
def test_batched_graph_replays_equal_eager(own_ctx, count, n, jobz):
    import evd4sc2025_amd as E

    torch = pytest.importorskip("torch")
    print(f"batched graph case count={count} n={n} jobz={jobz}", flush=True)
    sets = [[to_dev(O.gen_symmetric(n, 1000 * s + q, 0)) for q in range(count)] for s in range(3)]
    g = E.BatchedSolveGraph(own_ctx, sets[0], b=32, nb=0, jobz=jobz)
    try:
        ref_ctx = E.Context(0)
        try:
            eager = []
            for mats in sets:
                w, Z = E.syevd_batched(ref_ctx, mats, b=32, nb=0, jobz=jobz)
                torch.cuda.synchronize()
                eager.append((to_host(w), to_host(Z.contiguous()) if Z is not None else None))
        finally:
            ref_ctx.close()
        for rep, s in enumerate([1, 2, 0, 1]):
            w, Z = g.solve(sets[s])
            assert np.array_equal(to_host(w), eager[s][0]), f"replay {rep} (set {s}): eigenvalues differ"
            if jobz == "V":
                assert np.array_equal(to_host(Z.contiguous()), eager[s][1]), f"replay {rep} (set {s}): vectors differ"
    finally:
        g.close()


def test_mixed_graph_replays_equal_eager(own_ctx):
    """evd_graph_create_mixed: the whole mixed-size call (five buckets, both
    eigenvector flags, bandwidth min(32, N / 4)) as one graph.  Every replay
    equals evd_syevd_mixed_f32 on the same inputs bit for bit, with eager
    solves on the same context between replays; a replay without new inputs
    solves A again (A is not modified)."""
    import evd4sc2025_amd as E

    torch = pytest.importorskip("torch")
    orders = [70, 300, 64, 129, 700, 96, 300, 1000, 70, 150]
    jobz = ["V" if q % 3 else "N" for q in range(len(orders))]
    print(f"mixed graph case orders={orders}", flush=True)
    sets = [[to_dev(O.gen_symmetric(n, 3000 * s + q, 0)) for q, n in enumerate(orders)] for s in range(3)]
    g = E.MixedSolveGraph(own_ctx, sets[0], jobz=jobz, b=0)
    try:
        ref_ctx = E.Context(0)
        try:
            eager = []
            for mats in sets:
                res = E.syevd_mixed(ref_ctx, mats, jobz, b=0)
                torch.cuda.synchronize()
                eager.append([(to_host(r.w), to_host(r.Z) if r.Z is not None else None) for r in res])
        finally:
            ref_ctx.close()

        def check(res, s, what):
            for q, (r, (w0, Z0)) in enumerate(zip(res, eager[s])):
                assert np.array_equal(to_host(r.w), w0), f"{what} (set {s}), problem {q}: eigenvalues differ"
                if Z0 is not None:
                    assert np.array_equal(to_host(r.Z), Z0), f"{what} (set {s}), problem {q}: vectors differ"

        # creation solved sets[0] eagerly on the graph's buffers
        torch.cuda.synchronize()
        check([E.EigResult(w, Z, {}) for w, Z in zip(g.w, g.Z)], 0, "creation")
        for rep, s in enumerate([1, 2, 0, 1]):
            check(g.solve(sets[s]), s, f"replay {rep}")
            if rep == 1:
                E.syevd(own_ctx, sets[2][4], b=32, nb=256, jobz="V")
                torch.cuda.synchronize()
        check(g.replay(), 1, "replay on resident inputs")
        # accuracy of one problem of the last replay against the FP64 oracle
        M = to_host(sets[1][7]).astype(np.float64)
        wref = np.linalg.eigvalsh(M)
        assert np.abs(to_host(g.w[7]) - wref).max() <= 1e-5 * np.abs(wref).max()
    finally:
        g.close()
