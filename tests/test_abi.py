"""The C ABI: the library builds for gfx950, loads, and exports every entry
point include/evd4sc.h declares (no compute calls - CPU only)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "evd4sc.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(evd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_reference_counterparts():
    names = _declared()
    for must in ["evd_panel_qr_f32", "evd_sy2sb_f32", "evd_syr2k_f32", "evd_sb2st_f64", "evd_stedc_f64",
                 "evd_wy_from_slide_f32", "evd_q_from_wy_f32", "evd_bc_back_f32", "evd_sbr_back_f32",
                 "evd_syevd_f32", "evd_orthogonality_f32", "evd_backward_error_f32"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from evd4sc2025_amd import _native

    lib = _native.lib()
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding covers exactly the declared surface
    assert sorted(_native.SIGNATURES) == _declared()


def test_library_is_gfx950_code_object():
    """Every device code object bundled in the library targets gfx950 and
    nothing else, and no NVIDIA fat binary or PTX section is present."""
    so = ROOT / "evd4sc2025_amd" / "libevd4sc.so"
    data = so.read_bytes()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", data))
    assert targets == {b"gfx950"}, targets
    for marker in (b"__nv_fatbin", b".nv_fatbin", b"nvptx64", b"__nv_module_id"):
        assert marker not in data, marker


def test_workspace_query_is_host_only():
    from evd4sc2025_amd import _native

    lib = _native.lib()
    ws = lib.evd_syevd_workspace_f32(16384, 32, 1024, ctypes.c_char(b"V"))
    # W, Y (n x 16384 FP32 each), Z_T and D&C scratch in FP64: tens of GiB at most
    assert 8 << 30 < ws < 64 << 30
    assert lib.evd_syevd_workspace_f32(0, 32, 1024, ctypes.c_char(b"V")) == 0


def test_ctx_create_without_gpu_reports_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from evd4sc2025_amd import _native

    lib = _native.lib()
    h = ctypes.c_void_p()
    rc = lib.evd_ctx_create(0, None, ctypes.byref(h))
    assert rc != 0 and not h.value
