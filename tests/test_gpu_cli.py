"""The myevd command line (the reference's myEVD driver, src/EVD/myEVD.cu:251-706):
one solve of the uniform workload on the GPU, its stage lines, the
orthogonality and backward-error lines, and eigenvalue bounds that agree with
host LAPACK on the same matrix."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import evd_oracle as O

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "evd4sc2025_amd", "myevd")


def test_myevd_cli_solves_and_reports():
    if not os.path.exists(CLI):
        pytest.fail("the myevd binary is not built (run the package build)")
    n, b, nb, seed = 512, 32, 64, 3
    print(f"myevd {n} {b} {nb} {seed}", flush=True)
    out = subprocess.run([CLI, str(n), str(b), str(nb), str(seed)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    text = out.stdout
    assert f"myevd: n={n}, b={b}, nb={nb}, seed={seed}, jobz=V" in text
    assert re.search(r"EVD total 512x512 takes [0-9.]+ ms", text)
    orth = float(re.search(r"The Orthogonality metrix is (\S+)", text).group(1))
    be = float(re.search(r"Backforward err: (\S+)", text).group(1))
    assert orth <= 1e-6 and be <= 1e-6, (orth, be)
    lo, hi = (float(x) for x in re.search(r"eigenvalues: min (\S+) max (\S+)", text).groups())
    w = np.linalg.eigvalsh(O.gen_symmetric(n, seed, 0).astype(np.float64))
    tol = 1e-5 * np.abs(w).max()
    assert abs(lo - w[0]) <= tol and abs(hi - w[-1]) <= tol, (lo, hi, w[0], w[-1])
