"""Batched solves: one persistent chase launch and one BC back-transformation
launch for many same-size problems.  Each problem's result must equal the
single-problem solve bit for bit (the kernels' arithmetic per problem does not
depend on the batch), and the batched chase must equal the single chase."""
import numpy as np
import pytest

from oracle import evd_oracle as O
from tests.helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count,n,b", [(1, 130, 32), (3, 100, 32), (7, 300, 32), (4, 257, 64)])
def test_sb2st_batched_equals_single(ctx, count, n, b):
    import evd4sc2025_amd as E

    torch = pytest.importorskip("torch")
    packed = []
    for q in range(count):
        A = O.gen_symmetric(n, 70 + q)
        Ab, _, _ = O.sy2sb(A.astype(np.float64), b, 2 * b)
        B = O.band_of(Ab, b)  # dense band matrix
        packed.append(E.band_pack(ctx, to_dev(B.astype(np.float32)), b))  # 2b x n, column-major
    ldb, ldu = 2 * b, n + b
    band = torch.stack([bp.t() for bp in packed]).contiguous()  # (count, n, 2b): problem q column-major
    U = torch.zeros((count, n, ldu), dtype=torch.float32, device="cuda")
    D = torch.zeros((count, n), dtype=torch.float64, device="cuda")
    Ev = torch.zeros((count, n), dtype=torch.float64, device="cuda")
    import ctypes as C
    ctx.call("evd_sb2st_batched_f64", count, n, b, C.c_void_p(band.data_ptr()), ldb, ldb * n,
             C.c_void_p(U.data_ptr()), ldu, ldu * n, C.c_void_p(D.data_ptr()), n, C.c_void_p(Ev.data_ptr()), n)
    torch.cuda.synchronize()
    for q in range(count):
        Dq, Eq, Uq = E.sb2st(ctx, packed[q], n, b)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(D[q]), to_host(Dq))
        assert np.array_equal(to_host(Ev[q])[: n - 1], to_host(Eq))
        assert np.array_equal(to_host(U[q]).T, to_host(Uq))


@pytest.mark.parametrize("count,n,b,jobz", [(5, 200, 32, "V"), (3, 513, 32, "V"), (2, 300, 64, "V"),
                                            (4, 256, 32, "N"), (2, 4096, 32, "V")])
def test_syevd_batched_equals_single(ctx, count, n, b, jobz):
    import evd4sc2025_amd as E

    torch = pytest.importorskip("torch")
    mats = [to_dev(O.gen_symmetric(n, 90 + q, q % 2)) for q in range(count)]
    w, Z = E.syevd_batched(ctx, [m.clone() for m in mats], b=b, nb=256, jobz=jobz)
    torch.cuda.synchronize()
    for q in range(count):
        r = E.syevd(ctx, mats[q].clone(), b=b, nb=256, jobz=jobz)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(w[q]), to_host(r.w))
        if jobz == "V":
            assert np.array_equal(to_host(Z[q]), to_host(r.Z))
            A = to_host(mats[q]).astype(np.float64)
            assert np.abs(np.sort(to_host(w[q])) - np.linalg.eigvalsh(A)).max() < 1e-5 * np.abs(A).sum(0).max()


def test_syevd_batched_fp16_equals_single(ctx):
    import evd4sc2025_amd as E

    torch = pytest.importorskip("torch")
    ctx.set_precision(E.PREC_FP16)
    try:
        mats = [to_dev(O.gen_symmetric(400, 40 + q, 0)) for q in range(3)]
        w, Z = E.syevd_batched(ctx, [m.clone() for m in mats], b=32, nb=256)
        for q in range(3):
            r = E.syevd(ctx, mats[q].clone(), b=32, nb=256)
            torch.cuda.synchronize()
            assert np.array_equal(to_host(Z[q]), to_host(r.Z))
    finally:
        ctx.set_precision(E.PREC_FP32)


def test_syevd_batched_large_count_chunks_the_dc():
    """More merges per D&C level than one launch dimension holds (300 x
    n = 512: 76800 bottom-level merges) run as chunks of problems; sampled
    problems must still equal their single solves bit for bit."""
    import evd4sc2025_amd as E

    torch = pytest.importorskip("torch")
    count, n = 300, 512
    print(f"batched count={count} n={n}", flush=True)
    ctx = E.Context(0)
    try:
        mats = [E.gen_symmetric(ctx, n, 9000 + q, 0) for q in range(count)]
        w, Z = E.syevd_batched(ctx, mats, b=32, nb=256, jobz="V")
        torch.cuda.synchronize()
        for q in (0, 254, 255, 299):
            r = E.syevd(ctx, mats[q], b=32, nb=256, jobz="V")
            assert np.array_equal(to_host(w[q]), to_host(r.w)), q
            assert np.array_equal(to_host(Z[q].contiguous()), to_host(r.Z.contiguous())), q
    finally:
        ctx.close()


def test_syevd_batched_headline_size_equals_single():
    """Two n = 16384 problems as one batch (every stage one launch for both)
    equal their single solves bit for bit (compared on the device)."""
    import evd4sc2025_amd as E

    torch = pytest.importorskip("torch")
    n = 16384
    print(f"batched count=2 n={n}", flush=True)
    ctx = E.Context(0)
    try:
        mats = [E.gen_symmetric(ctx, n, 700 + q, 0) for q in range(2)]
        w, Z = E.syevd_batched(ctx, mats, b=32, nb=1024, jobz="V")
        for q in range(2):
            r = E.syevd(ctx, mats[q], b=32, nb=1024, jobz="V")
            torch.cuda.synchronize()
            assert torch.equal(w[q], r.w[:n]), q
            assert torch.equal(Z[q], r.Z), q
            del r
    finally:
        ctx.close()
