"""Serving loop (evd_server_*): continuous admission of uneven-size requests on
concurrent slots, HBM-budget admission, and results that do not depend on the
mix they ran in."""
import numpy as np
import pytest

from oracle import evd_oracle as O
from tests.helpers import to_dev, to_host

pytestmark = pytest.mark.gpu

# (n, jobz, seed): ragged sizes, both job kinds
MIX = [(256, "V", 1), (1024, "V", 2), (300, "N", 3), (2048, "V", 4), (513, "V", 5), (700, "N", 6), (256, "V", 7),
       (1536, "V", 8)]


def _solve_all(srv, reqs):
    out = []
    for n, jobz, seed in reqs:
        A = to_dev(O.gen_symmetric(n, seed, 0))
        out.append((n, srv.submit(A, b=32, nb=256, jobz=jobz)))
    srv.wait_all()
    return [(to_host(w)[:n], None if Z is None else to_host(Z)) for n, (_, w, Z) in out]


def test_server_results_identical_alone_and_in_mix():
    import evd4sc2025_amd as E

    alone = []
    srv1 = E.Server(0, slots=1)
    for req in MIX:
        alone += _solve_all(srv1, [req])
    srv1.close()
    srv4 = E.Server(0, slots=4)
    mixed = _solve_all(srv4, MIX)
    st = srv4.stats()
    srv4.close()
    assert st["completed"] == len(MIX)
    for (n, jobz, seed), (w1, Z1), (w4, Z4) in zip(MIX, alone, mixed):
        assert np.array_equal(w1, w4), (n, seed)
        if jobz == "V":
            assert np.array_equal(Z1, Z4), (n, seed)
        # and each result is a correct eigen-decomposition
        A = O.gen_symmetric(n, seed, 0).astype(np.float64)
        wr = np.linalg.eigvalsh(A)
        assert np.abs(np.sort(w4) - wr).max() < 1e-5 * np.abs(wr).max()


def test_server_budget_admission():
    import evd4sc2025_amd as E

    ws = E.syevd_workspace(1024, 32, 256, "V")
    # a request whose workspace alone exceeds the budget is refused up front
    small = E.Server(0, slots=2, hbm_budget=ws // 2)
    with pytest.raises(E.EvdError):
        small.submit(to_dev(O.gen_symmetric(1024, 9, 0)), b=32, nb=256)
    small.close()
    # a budget for one request at a time still completes every request, and
    # the slot arenas never exceed it
    srv = E.Server(0, slots=4, hbm_budget=int(ws * 1.5))
    res = _solve_all(srv, [(1024, "V", 10), (1024, "V", 11), (512, "V", 12), (1024, "N", 13)])
    st = srv.stats()
    srv.close()
    assert st["completed"] == 4
    assert st["arena_bytes"] <= int(ws * 1.5)
    for (w, Z), (n, seed) in zip(res, [(1024, 10), (1024, 11), (512, 12), (1024, 13)]):
        wr = np.linalg.eigvalsh(O.gen_symmetric(n, seed, 0).astype(np.float64))
        assert np.abs(np.sort(w) - wr).max() < 1e-5 * np.abs(wr).max()


def test_server_latency_statistics():
    import evd4sc2025_amd as E

    srv = E.Server(0, slots=3)
    ids = []
    for n, jobz, seed in MIX[:5]:
        rid, _, _ = srv.submit(to_dev(O.gen_symmetric(n, seed, 0)), b=32, nb=256, jobz=jobz)
        ids.append(rid)
    srv.wait_all()
    st = srv.stats()
    lat = [srv.latency_ms(i) for i in ids]
    srv.close()
    assert st["completed"] == 5 and st["solves_per_s"] > 0
    assert min(lat) > 0 and st["p50_ms"] <= st["p99_ms"] <= max(lat) + 1e-9


def test_server_same_size_requests_match_single_solves():
    """Requests of one size queued behind a running one are solved as a batch
    (evd_syevd_batched_f32); each result must equal the single solve."""
    import evd4sc2025_amd as E

    reqs = [(384, "V", 20 + i) for i in range(6)] + [(384, "N", 30), (384, "N", 31)]
    srv = E.Server(0, slots=1)
    res = _solve_all(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs)
    ctx = E.Context(0)
    for (n, jobz, seed), (w, Z) in zip(reqs, res):
        r = E.syevd(ctx, to_dev(O.gen_symmetric(n, seed, 0)), b=32, nb=256, jobz=jobz)
        assert np.array_equal(w, to_host(r.w)[:n]), seed
        if jobz == "V":
            assert np.array_equal(Z, to_host(r.Z)), seed
    ctx.close()


def test_server_fp16_precision_matches_single_fp16_solves():
    """A server switched to the split-FP16 form (evd_server_set_precision)
    returns, alone or batched, exactly the single FP16 solve of each request."""
    import evd4sc2025_amd as E

    reqs = [(256, "V", 40 + i) for i in range(4)] + [(600, "V", 50), (256, "N", 51)]
    srv = E.Server(0, slots=2, precision=E.PREC_FP16)
    res = _solve_all(srv, reqs)
    srv.close()
    ctx = E.Context(0)
    ctx.set_precision(E.PREC_FP16)
    for (n, jobz, seed), (w, Z) in zip(reqs, res):
        r = E.syevd(ctx, to_dev(O.gen_symmetric(n, seed, 0)), b=32, nb=256, jobz=jobz)
        assert np.array_equal(w, to_host(r.w)[:n]), seed
        if jobz == "V":
            assert np.array_equal(Z, to_host(r.Z)), seed
    ctx.close()


def test_server_admission_waits_for_memory_without_hanging():
    """A budget that holds three eigenvalue-only requests but not a fourth,
    larger request at the same time: the larger one is admitted once memory
    is returned (the budget is one pool lent out and returned per request),
    and every request completes within a bounded wait."""
    import threading

    import evd4sc2025_amd as E

    wsn = E.syevd_workspace(1024, 32, 256, "N")
    wsv = E.syevd_workspace(1024, 32, 256, "V")
    budget = 3 * wsn + (4 << 20) * 4 + (16 << 20)  # three N at once, plus the slots' plan pools
    assert wsv < budget < 3 * wsn + wsv + (4 << 20) * 4
    srv = E.Server(0, slots=4, hbm_budget=budget, batching=False, policy=1)
    reqs = [(1024, "N", 60), (1024, "N", 61), (1024, "N", 62), (1024, "V", 63)]
    done = {}

    def run():
        done["res"] = _solve_all(srv, reqs)

    th = threading.Thread(target=run)
    th.start()
    th.join(timeout=120)
    assert not th.is_alive(), "admission waited forever"
    st = srv.stats()
    srv.close()
    assert st["completed"] == 4 and st["arena_bytes"] <= budget
    for (w, Z), (n, jobz, seed) in zip(done["res"], reqs):
        wr = np.linalg.eigvalsh(O.gen_symmetric(n, seed, 0).astype(np.float64))
        assert np.abs(np.sort(w) - wr).max() < 1e-5 * np.abs(wr).max()
