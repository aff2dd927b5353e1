"""The committed golden case (tests/golden/room_small.npz): the oracle still
produces every stored array bit for bit, and the fixture is a meaningful
case (hits, observed voxels, ICP inliers)."""
import os
import sys

import numpy as np

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_golden  # noqa: E402

PATH = os.path.join(ROOT, "tests", "golden", "room_small.npz")


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))
    return bool(np.array_equal(a, b))


def test_golden_fixture_is_a_real_case():
    g = np.load(PATH)  # allow_pickle stays False
    assert (g["raycast1_hit"] >= 0).sum() > 1000
    assert np.isfinite(g["weight01"]).sum() > 10000
    assert np.all(g["icp2_sums"][:, 27] > 1000)
    assert np.all(g["icp2_masks"].sum(axis=(1, 2)) == g["icp2_sums"][:, 27])


def test_oracle_reproduces_the_golden_case():
    g = np.load(PATH)
    fresh = make_golden.build()
    assert sorted(fresh) == sorted(g.files)
    bad = [k for k in g.files if not same(fresh[k], g[k])]
    assert not bad, bad
