"""Autotuner: tuning a problem class fills the cache, and solves that take
their knobs from the cache (nb = 0, knobs left at 0) stay at the accuracy
bar for every knob setting the tuner can pick."""
import numpy as np
import pytest

from oracle import evd_oracle as O
from tests.helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture()
def tctx():
    import evd4sc2025_amd as E

    E.tune_cache_clear()
    c = E.Context(0)
    yield c
    c.close()  # not left to garbage collection, which may run on another thread
    E.tune_cache_clear()


def _check(E, ctx, A, b, nb):
    n = A.shape[0]
    r = E.syevd(ctx, to_dev(A), b=b, nb=nb)
    w_ref = np.linalg.eigvalsh(A.astype(np.float64))
    assert np.abs(np.sort(to_host(r.w)) - w_ref).max() < 1e-5 * np.abs(w_ref).max()
    assert E.backward_error(ctx, to_dev(A), r.w, r.Z) < 1e-6
    assert E.checkOrthogonality(ctx, r.Z) < 1e-6


def test_autotune_fills_cache_and_solves(tctx):
    import evd4sc2025_amd as E

    knobs = E.autotune(tctx, 600, 32, reps=1)
    assert E.tune_cache_get(600, 32, E.PREC_FP32) == knobs
    assert knobs["nb"] in (256, 512, 1024) and knobs["bc_valu"] in (0, 1)
    assert knobs["panel_rows"] in (0, 256) and knobs["gemm_tile"] in (0, 1, 2) and knobs["chase_depth"] in (0, 2, 3)
    _check(E, tctx, O.gen_symmetric(600, 11, 0), 32, 0)


@pytest.mark.parametrize("knobs", [{"nb": 256, "chase_groups": 64, "bc_valu": 1, "sbr_back_block": 512,
                                    "bc_sweeps": 32, "bc_waves": 4},
                                   {"nb": 512, "chase_groups": 256, "bc_valu": 0, "sbr_back_block": 2048,
                                    "chase_depth": 3, "panel_rows": 256, "gemm_tile": 2},
                                   {"nb": 1024, "chase_depth": 2, "gemm_tile": 1}])
def test_cached_knobs_keep_accuracy(tctx, knobs):
    import evd4sc2025_amd as E

    E.tune_cache_put(700, 32, E.PREC_FP32, knobs)
    _check(E, tctx, O.gen_symmetric(700, 12, 0), 32, 0)
    tctx.use_tune_cache(False)
    _check(E, tctx, O.gen_symmetric(700, 12, 0), 32, 0)
    tctx.use_tune_cache(True)


@pytest.mark.parametrize("n", [3000, 9000])
def test_sbr_back_block_2048_at_odd_multiples_of_1024(tctx, n):
    """W and Y are padded to a multiple of 1024 columns; with 3 or 9 such
    blocks a 2048-wide SBR back-transformation block does not divide the
    padding, so the block shrinks to 1024 (both as a knob and from the cache
    entry of the order's bucket)."""
    import evd4sc2025_amd as E

    print(f"sbr_back_block 2048 at n={n}", flush=True)
    E.tune_cache_put(n, 32, E.PREC_FP32, {"nb": 1024, "chase_groups": 0, "bc_valu": 0, "sbr_back_block": 2048})
    A = O.gen_symmetric(n, 13, 0)
    _check(E, tctx, A, 32, 0)
    tctx.use_tune_cache(False)
    tctx.set_knob("sbr_back_block", 2048)
    _check(E, tctx, A, 32, 1024)
    tctx.set_knob("sbr_back_block", 0)
    tctx.use_tune_cache(True)
