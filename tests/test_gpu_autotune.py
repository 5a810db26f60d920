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
    _check(E, tctx, O.gen_symmetric(600, 11, 0), 32, 0)


@pytest.mark.parametrize("knobs", [{"nb": 256, "chase_groups": 64, "bc_valu": 1, "sbr_back_block": 512},
                                   {"nb": 512, "chase_groups": 256, "bc_valu": 0, "sbr_back_block": 2048}])
def test_cached_knobs_keep_accuracy(tctx, knobs):
    import evd4sc2025_amd as E

    E.tune_cache_put(700, 32, E.PREC_FP32, knobs)
    _check(E, tctx, O.gen_symmetric(700, 12, 0), 32, 0)
    tctx.use_tune_cache(False)
    _check(E, tctx, O.gen_symmetric(700, 12, 0), 32, 0)
    tctx.use_tune_cache(True)
