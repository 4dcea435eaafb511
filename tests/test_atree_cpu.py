"""CPU tests of the adaptive quadtree restatement (DAPIF2's tree, dapif1.f:740-1680 and
dapif2.f:3-474).

Pins: the reference's own routines (fixtures tests/golden/atree_*.npz written from
oracle/_ref/libref.so, and the live library when it is built here), the scaled coordinates
DAPIF2 leaves in its workspace (which fixes the REAL*4 0.99 of dapif1.f:765), and
brute-force geometric definitions of boxes and lists (tests/atree_rules.py).
"""
import os

import numpy as np
import pytest

import oracle
from tests.atree_rules import check_tree, lists_bruteforce, lists_from_tables
from tests.conftest import GOLDEN

CASES = ["curves2048_napb30", "uniform5000_napb40", "cluster3000_napb40", "curves16384_napb40"]


def fixture(name):
    return np.load(os.path.join(GOLDEN, "atree_" + name + ".npz"))


@pytest.mark.parametrize("name", CASES)
def test_atree_oracle_matches_reference_fixture(name):
    f = fixture(name)
    t = oracle.atree(f["x"], f["y"], int(f["napb"]), int(f["lmx"]), float(f["eps7"]))
    assert t["ierr"] == 0
    counts = [t[k] for k in ("nbox", "nlev", "len", "len3", "len4", "maxp")]
    assert counts == f["counts"].tolist()
    for k in oracle.ATREE_TABLES:
        np.testing.assert_array_equal(t[k], f[k], err_msg=k)


@pytest.mark.parametrize("name", ["curves2048_napb30", "cluster3000_napb40"])
def test_atree_oracle_geometric_rules(name):
    f = fixture(name)
    napb = int(f["napb"])
    t = oracle.atree(f["x"], f["y"], napb)
    check_tree(t, napb)
    b1, b2, b3, b4 = lists_bruteforce(t)
    L1, L2, L3, L4 = lists_from_tables(t)
    for b in b1:
        assert set(L1[b]) == b1[b] and len(L1[b]) == len(b1[b]), ("list1", b)
        assert set(L2[b]) == b2[b] and len(L2[b]) == len(b2[b]), ("list2", b)
        assert set(L3[b]) == b3[b] and len(L3[b]) == len(b3[b]), ("list3", b)
        assert set(L4[b]) == b4[b] and len(L4[b]) == len(b4[b]), ("list4", b)
    assert sum(len(v) for v in b3.values()) > 0 and sum(len(v) for v in b4.values()) > 0


def test_atree_double_precision_099_would_be_caught():
    """A restatement that scaled with a double 0.99 would miss DAPIF2's coordinates."""
    f = fixture("curves2048_napb30")
    x, y = f["x"], f["y"]
    rs = max(x.max() - x.min(), y.max() - y.min())
    xo = (x.max() + x.min()) / 2.0
    wrong = (x - xo) * (128.0 / rs * 0.99)
    right = (x - xo) * (128.0 / rs * float(np.float32(0.99)))
    np.testing.assert_array_equal(right, f["xs"])
    assert not np.array_equal(wrong, f["xs"])


def test_atree_error_codes():
    rng = np.random.default_rng(3)
    x, y = rng.uniform(size=(2, 30))
    assert oracle.atree(x, y, 30)["ierr"] == 16           # NAT <= NAPB (dapif1.f:711-717)
    x = np.concatenate([rng.uniform(size=50), np.full(40, 0.5)])
    y = np.concatenate([rng.uniform(size=50), np.full(40, 0.5)])
    assert oracle.atree(x, y, 30, lmx=12)["ierr"] == 32   # 40 coincident points never separate
    assert oracle.dapif_lmx(300, 1e-11) == 30              # the parameters of laplace.f:209-213


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built here")
def test_atree_oracle_matches_live_reference():
    rng = np.random.default_rng(11)
    pts = np.concatenate([rng.normal(-0.2, 0.003, (700, 2)), rng.normal(0.4, 0.05, (900, 2)),
                          rng.uniform(-1, 1, (400, 2))])
    for napb in (20, 40):
        w = oracle.ref_dapif2_scaled(pts[:, 0], pts[:, 1], napb)
        r = oracle.ref_atree(w["xs"], w["ys"], napb)
        t = oracle.atree(pts[:, 0], pts[:, 1], napb)
        assert (t["nbox"], t["nlev"]) == (w["nbox"], w["nlev"])
        for k in oracle.ATREE_TABLES:
            np.testing.assert_array_equal(t[k], r[k], err_msg=k)
