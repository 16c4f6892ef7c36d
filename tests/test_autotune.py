"""The autotuner's table lookup and fusion planning on the CPU (no device needed)."""

import json

import pytest

from zkgpu import autotune


@pytest.fixture()
def table(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "TUNING_DIR", tmp_path)
    autotune._CACHE.clear()
    t = {"msm_window": {"10": 9, "12": 10, "16": 13, "20": 16}, "msm_chunk": {"20": 0}, "ntt_fusion": {"22": 8}}
    (tmp_path / "TestGPU.json").write_text(json.dumps(t))
    yield "TestGPU"
    autotune._CACHE.clear()


def test_lookup_nearest_tuned_size(table):
    assert autotune.lookup("msm_window", 1 << 16, device=table) == 13
    assert autotune.lookup("msm_window", 1 << 20, device=table) == 16
    assert autotune.lookup("msm_window", 3 << 15, device=table) == 13  # log2 = 16.58 -> 16
    assert autotune.lookup("msm_window", 1 << 11, device=table) == 9  # tie 10/12 -> the smaller
    assert autotune.lookup("msm_window", 1 << 30, device=table) == 16


def test_lookup_without_table_is_automatic(table):
    assert autotune.lookup("msm_window", 1 << 16, device="NoSuchGPU") == 0
    assert autotune.lookup("ntt_fusion", 0, device=table) == 0
    assert autotune.lookup("unknown_kind", 1 << 16, device=table) == 0


def test_fusion_candidates_cover_every_pass_count():
    # 2^22: 3 passes at 8..10 stages, 4 at 6..7, 5 at 5, 6 at 4
    plans = autotune._distinct_fusions(22)
    assert sorted(p[0] for p in plans) == [3, 4, 5, 6]
    assert plans[(3,)] == 8
    # small sizes: a single pass once the depth reaches log n
    assert autotune._distinct_fusions(10)[(1,)] == 10
