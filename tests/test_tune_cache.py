"""Tuning cache of the autotuner (host-only entry points of the C ABI): keys by
(next power of two >= n, b, precision), text round trip, rejection of
malformed files without touching the cache."""
import pytest


@pytest.fixture()
def E():
    import evd4sc2025_amd as E

    E.tune_cache_clear()
    yield E
    E.tune_cache_clear()


def _full(**kw):
    import evd4sc2025_amd as E

    d = {k: 0 for k in E.KNOBS}
    d.update(kw)
    return d


def test_put_get_buckets(E):
    want = _full(nb=512, chase_groups=128, bc_valu=1, sbr_back_block=2048, chase_depth=3, panel_rows=256,
                 gemm_tile=2, bc_sweeps=32, bc_waves=4)
    E.tune_cache_put(1000, 32, E.PREC_FP32, want)
    assert E.tune_cache_get(1000, 32, E.PREC_FP32) == want
    assert E.tune_cache_get(1024, 32, E.PREC_FP32) == want  # same bucket
    assert E.tune_cache_get(513, 32, E.PREC_FP32) == want
    assert E.tune_cache_get(512, 32, E.PREC_FP32) is None
    assert E.tune_cache_get(1025, 32, E.PREC_FP32) is None
    assert E.tune_cache_get(1000, 64, E.PREC_FP32) is None
    assert E.tune_cache_get(1000, 32, E.PREC_FP16) is None


def test_invalid_entries_rejected(E):
    for bad in ({"nb": 100}, {"nb": 16}, {"bc_valu": 2}, {"chase_groups": -1}, {"chase_depth": 7},
                {"panel_rows": 300}, {"gemm_tile": 3}, {"bc_sweeps": 65}, {"bc_waves": 17}):
        with pytest.raises(E.EvdError):
            E.tune_cache_put(1000, 32, E.PREC_FP32, bad)
    assert E.tune_cache_get(1000, 32, E.PREC_FP32) is None


def test_save_load_round_trip(E, tmp_path):
    E.tune_cache_put(4096, 32, E.PREC_FP16, {"nb": 1024, "chase_groups": 256})
    E.tune_cache_put(300, 64, E.PREC_BF16, {"nb": 256, "sbr_back_block": 512})
    path = tmp_path / "tune.txt"
    E.tune_cache_save(path)
    E.tune_cache_clear()
    assert E.tune_cache_get(4096, 32, E.PREC_FP16) is None
    E.tune_cache_load(path)
    assert E.tune_cache_get(4000, 32, E.PREC_FP16) == _full(nb=1024, chase_groups=256)
    assert E.tune_cache_get(300, 64, E.PREC_BF16)["sbr_back_block"] == 512


def test_malformed_file_leaves_cache(E, tmp_path):
    E.tune_cache_put(2048, 32, E.PREC_FP32, {"nb": 512})
    bad = tmp_path / "bad.txt"
    bad.write_text("# header\n2048 32 0 1024 0 0 0\n4096 32 0 100 0 0 0\n")  # nb 100 is no multiple of 32
    with pytest.raises(E.EvdError):
        E.tune_cache_load(bad)
    assert E.tune_cache_get(2048, 32, E.PREC_FP32)["nb"] == 512
    with pytest.raises(E.EvdError):
        E.tune_cache_load(tmp_path / "missing.txt")


def test_unknown_knob_name_rejected(E):
    with pytest.raises(ValueError):
        E.tune_cache_put(1000, 32, E.PREC_FP32, {"nb": 512, "panel_width": 64})


def test_seven_field_lines_load_with_new_knobs_zero(E, tmp_path):
    """Cache files written before the knob vector grew (n_bucket b precision
    nb chase_groups bc_valu sbr_back_block) still load."""
    old = tmp_path / "old.txt"
    old.write_text("# n_bucket b precision nb chase_groups bc_valu sbr_back_block\n2048 32 0 512 64 1 2048\n")
    E.tune_cache_load(old)
    assert E.tune_cache_get(2048, 32, E.PREC_FP32) == _full(nb=512, chase_groups=64, bc_valu=1, sbr_back_block=2048)
    new = tmp_path / "new.txt"
    E.tune_cache_save(new)
    rows = [l.split() for l in new.read_text().splitlines() if not l.startswith("#")]
    assert rows == [["2048", "32", "0", "512", "64", "1", "2048", "0", "0", "0", "0", "0"]]
    bad = tmp_path / "bad.txt"
    bad.write_text("2048 32 0 512 64 1 2048 3 256\n")  # neither 7 nor 12 fields
    with pytest.raises(E.EvdError):
        E.tune_cache_load(bad)
