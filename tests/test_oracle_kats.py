"""The FP64 oracle against the known-answer vectors of the reference's own
tests (fixtures in tests/golden/reference_kats.json, written by
tests/golden/make_golden.py from friction_contact_test.cu:10-93,
broadphase/main.cpp:353-380 and interval_root_finder.cpp:516-517)."""
import json
import os

import numpy as np
import pytest

import oracle as O

with open(os.path.join(os.path.dirname(__file__), "golden",
                       "reference_kats.json")) as _f:
    KAT = json.load(_f)


def test_ee_uv_scale_independent():
    k = KAT["ee_uv_scale_independent"]
    ok, uv = O.ee_uv(np.array(k["x"], dtype=np.float32).astype(np.float64), k["eps"])
    assert ok
    np.testing.assert_allclose(uv, k["expected_uv"], rtol=1e-6)


@pytest.mark.parametrize("case", KAT["coulomb_exact"], ids=lambda c: c["name"])
def test_coulomb_exact(case):
    out = O.coulomb(case["u"], case["n"], case["mu"])
    np.testing.assert_allclose(out, case["expected"], rtol=0, atol=1e-15)


def test_coulomb_complementarity_and_cone():
    k = KAT["coulomb_sweep"]
    n = np.array(k["n"], dtype=np.float64)
    tol = k["tolerance"]
    for mu in k["mu"]:
        for u in k["u"]:
            s = O.coulomb(u, n, mu)
            react = s - np.asarray(u, dtype=np.float64)
            sn, rn = s @ n, react @ n
            rt = react - rn * n
            assert sn >= -tol and rn >= -tol
            assert abs(sn * rn) <= tol
            assert np.linalg.norm(rt) <= mu * rn + tol


def test_touching_boxes_pair():
    k = KAT["touching_boxes"]
    pairs = O.overlap_pairs(np.array(k["boxes"], dtype=np.float32))
    assert pairs.tolist() == k["expected_pairs"]


def test_ccd_filter_constants():
    k = KAT["ccd_filters"]
    # The reference writes the filters as 7-digit float literals: VF rounds
    # to exactly 34 ulps of 1, EE to two float ulps above 32 ulps of 1.
    vf = float(np.float32(k["vf_literal"]))
    ee = float(np.float32(k["ee_literal"]))
    assert vf == k["vf"]
    assert abs(ee - k["ee"]) == 2 * float(np.spacing(np.float32(k["ee"])))
    e = O.ccd_filter([0.5, 2.0, 3.0], True)
    np.testing.assert_array_equal(e, [vf, 2 * vf, 3 * vf])
    e = O.ccd_filter([0.0, 1.0, 4.0], False)
    np.testing.assert_array_equal(e, [ee, ee, 4 * ee])
