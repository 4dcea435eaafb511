"""CPU tests of the DGMRES restatement (oracle/gmres.py) for the settings of the Stokes solve.

Pins: the reference's DGMRES itself (fixtures tests/golden/gmres_*.npz written from
oracle/_ref, and the live library when built here) and numpy's dense solve.
"""
import os

import numpy as np
import pytest

import oracle
from oracle.gmres import dgmres
from tests.conftest import GOLDEN


def g(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def test_gmres_restatement_matches_reference_dense_fixtures():
    f = g("gmres_dense")
    for name in f["names"]:
        a, b, x0 = f[f"{name}_a"], f[f"{name}_b"], f[f"{name}_x0"]
        maxl, nrmax, it, ierr = f[f"{name}_meta"].tolist()
        x, ito, err, iero = dgmres(lambda v: a @ v, b, x0, maxl, 1e-10, nrmax)
        assert (ito, iero) == (it, ierr), name
        assert abs(err - float(f[f"{name}_err"])) <= 1e-6 * float(f[f"{name}_err"]), name
        np.testing.assert_allclose(x, f[f"{name}_x"], rtol=0, atol=1e-12 * np.abs(f[f"{name}_x"]).max())
        if ierr == 0:  # and it solves the system
            xs = np.linalg.solve(a, b)
            assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs), name


def test_gmres_restatement_stagnation_and_zero_rhs_paths():
    rng = np.random.default_rng(2)
    n = 30
    a = np.eye(n)
    b = rng.normal(size=n)
    x, it, err, ierr = dgmres(lambda v: a @ v, b, b.copy(), 10, 1e-10, -1)  # already solved
    assert (it, ierr) == (0, 0) and np.array_equal(x, b)
    x, it, err, ierr = dgmres(lambda v: a @ v, b, np.zeros(n), 10, 1e-10, -1)  # identity: one step
    assert (it, ierr) == (1, 0) and np.allclose(x, b, atol=1e-14)
    # a rotation (cyclic shift): GMRES makes no progress until the full Krylov space
    p = np.roll(np.eye(n), 1, axis=0)
    e1 = np.zeros(n)
    e1[0] = 1.0
    x, it, err, ierr = dgmres(lambda v: p @ v, e1, np.zeros(n), 5, 1e-10, -1)
    assert ierr in (1, 2) and it == 5


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built here")
def test_gmres_restatement_matches_live_reference():
    rng = np.random.default_rng(9)
    for n, maxl, nrmax in ((40, 8, 0), (90, 30, 2), (25, 25, -1)):
        a = np.eye(n) + 0.4 * rng.normal(size=(n, n)) / np.sqrt(n)
        b = rng.normal(size=n)
        xr, itr, er, ierr_r = oracle.ref_dgmres_dense(a, b, np.zeros(n), maxl, 1e-10, nrmax)
        xo, ito, eo, ierr_o = dgmres(lambda v: a @ v, b, np.zeros(n), maxl, 1e-10, nrmax)
        assert (itr, ierr_r) == (ito, ierr_o)
        np.testing.assert_allclose(xo, xr, rtol=0, atol=1e-12 * np.abs(xr).max())
    p = np.roll(np.eye(20), 1, axis=0)
    e1 = np.eye(20)[0]
    xr, itr, er, ierr_r = oracle.ref_dgmres_dense(p, e1, np.zeros(20), 5, 1e-10, -1)
    xo, ito, eo, ierr_o = dgmres(lambda v: p @ v, e1, np.zeros(20), 5, 1e-10, -1)
    assert (itr, ierr_r) == (ito, ierr_o)


def test_stokes_fixture_solutions_recover_the_density():
    """The reference solve of A u = A u_true returns u_true to its tolerance."""
    for tag in ("two_drops", "16drops"):
        f = g("gmres_stokes_" + tag)
        rel = np.linalg.norm(f["x"] - f["u_true"]) / np.linalg.norm(f["u_true"])
        assert rel < 1e-9, (tag, rel)
        assert f["meta"][1] == 0
