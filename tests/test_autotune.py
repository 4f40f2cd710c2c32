"""Autotuner selection logic and sampling (CPU; the measurements run in the
GPU batch workload)."""
import numpy as np
import pytest

from cucm_amd.autotune import REFERENCE, choose, stratified_sample


def _t(c, tb, ratio, score):
    return {"chunk_max": c, "table_bits": tb, "ratio": ratio, "score": score}


def test_choose_keeps_ratio_within_tolerance():
    ref = _t(REFERENCE["chunk_max"], REFERENCE["table_bits"], 2.000, 100.0)
    fast_bad = _t(128 * 1024, 16, 1.950, 300.0)   # 2.5% worse ratio
    fast_ok = _t(1024 * 1024, 16, 1.995, 150.0)    # 0.25% worse
    assert choose([ref, fast_bad, fast_ok], 0.01) is fast_ok
    assert choose([ref, fast_bad, fast_ok], 0.05) is fast_bad
    assert choose([ref, fast_bad, fast_ok], 0.0) is ref


def test_choose_requires_the_reference_setting():
    with pytest.raises(ValueError):
        choose([_t(128 * 1024, 16, 2.0, 1.0)], 0.01)


def test_stratified_sample_spans_sizes_within_budget():
    rng = np.random.default_rng(1)
    sizes = np.clip(np.round(65536 * np.exp(1.5 * rng.standard_normal(5000))), 512, 64 << 20)
    idx = stratified_sample(sizes, 32 << 20)
    got = sizes[idx]
    assert got.sum() <= 32 << 20
    assert len(idx) == len(set(idx)) > 50
    # both tails of the distribution are represented
    assert got.min() < np.percentile(sizes, 10) and got.max() > np.percentile(sizes, 90)
