"""Writes tests/golden/headline_n16384_s0.npz: host LAPACK dsyevd (FP64,
eigenvalues only) of the headline input, A = (X + X^T)/2 with X i.i.d.
uniform[-1,1] from the Philox4x32-10 stream, n = 16384, seed 0, rounded to
FP32 exactly as the device generator rounds it.

Run from the repository root:  python tests/golden/gen_golden_headline.py
(about 10 minutes on 8 cores; 6 GiB of host memory).  The input is generated
in column blocks with oracle.rand_entries, the same counter mapping
(r + c n) as oracle.gen_symmetric, whose n = 2048 output the script checks
first, so the blocked form is the oracle's matrix.  Stored: eigenvalues,
||A||_2 (max |eigenvalue|), the SHA-256 of the FP32 input and its first
column, which the GPU test compares with the device generator's output.
"""
import hashlib
import sys
import time
from pathlib import Path

import numpy as np
import scipy.linalg as sl

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import evd_oracle as O  # noqa: E402


def gen_symmetric_blocked(n: int, seed: int, kind: int = 0, block: int = 512) -> np.ndarray:
    """oracle.gen_symmetric in column blocks (bounded memory)."""
    X = np.empty((n, n), dtype=np.float64, order="F")
    r = np.arange(n, dtype=np.uint64)
    for c0 in range(0, n, block):
        c1 = min(n, c0 + block)
        c = np.arange(c0, c1, dtype=np.uint64)
        idx = (r[:, None] + c[None, :] * np.uint64(n)).ravel(order="F")
        X[:, c0:c1] = O.rand_entries(idx, 0, seed, kind).reshape(n, c1 - c0, order="F")
    A = np.empty((n, n), dtype=np.float32, order="F")
    for c0 in range(0, n, block):
        c1 = min(n, c0 + block)
        A[:, c0:c1] = (0.5 * (X[:, c0:c1] + X[c0:c1, :].T)).astype(np.float32)
    return A


def main():
    small = gen_symmetric_blocked(2048, 0, 0, block=300)
    assert np.array_equal(small, O.gen_symmetric(2048, 0, 0)), "blocked generator differs from the oracle"
    n, seed = 16384, 0
    t0 = time.time()
    A = gen_symmetric_blocked(n, seed, 0)
    # the device generator's buffer is column major: hash the column-major bytes
    sha = hashlib.sha256(A.tobytes(order="F")).digest()
    col0 = A[:, 0].copy()
    print(f"generated in {time.time() - t0:.0f} s", flush=True)
    A64 = A.astype(np.float64)
    del A
    t0 = time.time()
    w = sl.eigh(A64, driver="evd", eigvals_only=True, overwrite_a=True, check_finite=False)
    print(f"dsyevd (jobz=N) in {time.time() - t0:.0f} s", flush=True)
    out = {"w": w, "norm2": np.array(np.abs(w).max()), "input_sha256_colmajor": np.frombuffer(sha, dtype=np.uint8),
           "col0": col0}
    np.savez_compressed(Path(__file__).with_name(f"headline_n{n}_s{seed}.npz"), **out)
    print("wrote", Path(__file__).with_name(f"headline_n{n}_s{seed}.npz"))


if __name__ == "__main__":
    main()
