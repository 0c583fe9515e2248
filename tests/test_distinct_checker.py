"""The first-occurrence checker the GPU distinct tests compare against, pinned on the CPU: the
reference's own case (python/tests/test_stream_compaction.py:97-121: rows 0, 2, 3, 4 kept for each
of its key sets) and pandas.DataFrame.duplicated(keep="first"), an independent implementation with
the same key equality (every NaN one key, -0.0 equal to 0.0)."""

from __future__ import annotations

import numpy as np
import pyarrow as pa
import pytest

from ldf_testing import first_rows


def test_reference_case_first_rows():
    df = pa.table({"a": [1, 1, 2, 3, 4], "b": [2, 2, 4, 4, None], "c": [10, 10, 20, 30, 40]})
    for keys in (["b", "a"], ["a", "b"], ["b", "c"], ["a", "b", "c"]):
        assert first_rows(df, keys) == [0, 2, 3, 4]
    assert first_rows(pa.table({"k": [None, None, 1, None]}), ["k"]) == [0, 2]  # nulls are one key


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_first_rows_matches_pandas_duplicated(seed):
    pd = pytest.importorskip("pandas")
    rng = np.random.default_rng(seed)
    n = 20_000
    vals = np.array([0.0, -0.0, np.nan, -np.nan, 1.5, -2.25, np.inf, -np.inf])
    f = vals[rng.integers(0, len(vals), n)]
    f[rng.integers(0, n, 20)] = np.array([0x7FF0000000000123], np.uint64).view(np.float64)[0]  # NaN payload
    i = rng.integers(0, 50, n)
    s = np.array(["", "a", "bb", "ccc"])[rng.integers(0, 4, n)]
    t = pa.table({"f": f, "i": i, "s": s})
    pdf = pd.DataFrame({"f": f, "i": i, "s": s})
    for keys in (["f"], ["i", "s"], ["f", "i"], ["f", "i", "s"]):
        exp = np.flatnonzero(~pdf.duplicated(subset=keys, keep="first").to_numpy()).tolist()
        assert first_rows(t, keys) == exp, keys


def test_first_rows_catches_bitwise_float_keys():
    """A checker comparing float keys bitwise would keep -0.0 and 0.0 apart: this one does not."""
    t = pa.table({"k": [0.0, -0.0, float("nan"), -float("nan"), 1.0]})
    assert first_rows(t, ["k"]) == [0, 2, 4]
