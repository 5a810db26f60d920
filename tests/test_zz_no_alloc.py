"""evd_sb2st_f64_noalloc (the bulge chase with its scratch in a fixed device
array instead of the context arena) against evd_sb2st_f64 on the same band:
the two share the chase kernel, so D, E and U must agree bit for bit, on two
calls in a row (the second with a smaller order, so the shared progress words
and ticket must have been cleared by the call), and the first also against
the oracle.  Orders above the scratch limit are rejected before any launch."""
import numpy as np
import pytest

from oracle import evd_oracle as O
from tests.helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _band(n, b, seed):
    A = O.gen_symmetric(n, seed, 0).astype(np.float64)
    return A, O.band_of(A, b)


@pytest.mark.parametrize("b", [32, 64])
def test_noalloc_chase_equals_arena_chase(ctx, b):
    import torch

    import evd4sc2025_amd as E

    for n, seed in [(900, 31), (300, 32)]:  # second call smaller than the first
        print(f"noalloc chase n={n} b={b}", flush=True)
        _, B = _band(n, b, seed)
        B32 = B.astype(np.float32)
        ref_band = E.band_pack(ctx, to_dev(B32), b)
        new_band = ref_band.clone()
        D0, E0, U0 = E.sb2st(ctx, ref_band, n, b)
        D1, E1, U1 = E.sb2st(ctx, new_band, n, b, noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(D0, D1) and torch.equal(E0, E1) and torch.equal(U0, U1)
        if n == 300:
            Do, Eo, _ = O.sb2st(B32.astype(np.float64), b)
            assert np.abs(to_host(D1) - Do).max() < 1e-10 * max(1.0, np.abs(Do).max())
            assert np.abs(np.abs(to_host(E1)) - np.abs(Eo)).max() < 1e-10 * max(1.0, np.abs(Do).max())


def test_noalloc_chase_rejects_orders_above_the_scratch(ctx):
    import ctypes as C

    import torch

    import evd4sc2025_amd as E
    from evd4sc2025_amd import _native

    lib = _native.lib()
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda")
    n, b = 40000, 32
    p = C.c_void_p(dummy.data_ptr())
    rc = lib.evd_sb2st_f64_noalloc(ctx.h, n, b, p, 2 * b, p, n + b, p, p)
    assert rc == 5  # EVD_E_UNSUPPORTED
    assert b"scratch limit" in lib.evd_ctx_last_error(ctx.h)
