"""The oracle against the committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from numpy brute force)."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as orc

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
OPS = {"SqrDistance": orc.SSD, "AbsDistance": orc.SAD, "CrossCorr": orc.CC,
       "SqrDistance_Norm": orc.SSD_NORM, "CrossCorr_Norm": orc.CC_NORM}


def test_fixtures_present():
    assert len(FIXTURES) >= 5


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_oracle_matches_fixture(path):
    g = np.load(path)  # allow_pickle defaults to False
    src, tpl = g["src"], g["tpl"]
    for name, op in OPS.items():
        np.testing.assert_allclose(orc.prox_f64(op, src, tpl), g[name], rtol=1e-12, atol=1e-12,
                                   err_msg=name)
    if np.issubdtype(src.dtype, np.integer):
        np.testing.assert_array_equal(orc.prox_i64(orc.I64_SSD, src, tpl), g["SqrDistance_i64"])
        np.testing.assert_array_equal(orc.prox_i64(orc.I64_SAD, src, tpl), g["AbsDistance_i64"])
        np.testing.assert_array_equal(orc.prox_i64(orc.I64_CC, src, tpl), g["CrossCorr_i64"])
    np.testing.assert_allclose(orc.integral_exact(src), g["integral"], rtol=1e-13)
    np.testing.assert_allclose(orc.integral_exact(src, sqr=True), g["sqr_integral"], rtol=1e-13)
