"""GPU parity of the device-resident GMRES (include/eb_gmres.h) with the reference's DGMRES:
dense systems (tests/golden/gmres_dense.npz) and SOLVEI's Stokes solve with the reference
MATVEC (gmres_stokes_*.npz: GMRES iteration counts within one, solutions to the tolerance)."""
import os

import numpy as np
import pytest

from electricbubbles_amd.workloads import many_drops_cfg
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no GPU")
    return _t


@pytest.fixture(scope="module")
def mods(torch):
    from electricbubbles_amd import gmres, spectral
    return gmres, spectral


def g(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def test_gmres_dense_matches_reference(torch, mods):
    gm, _ = mods
    f = g("gmres_dense")
    for name in f["names"]:
        print("gmres dense", name, flush=True)
        a, b, x0 = (torch.tensor(f[f"{name}_{k}"], device="cuda") for k in ("a", "b", "x0"))
        maxl, nrmax, it, ierr = f[f"{name}_meta"].tolist()
        solver = gm.GMRES(a.shape[0], maxl)
        x = x0.clone()
        it_g, err_g, ierr_g = solver.solve_dense(a, b, x, 1e-10, nrmax)
        solver.close()
        assert ierr_g == ierr and abs(it_g - it) <= 1, (name, it_g, it)
        ref = f[f"{name}_x"]
        assert np.abs(x.cpu().numpy() - ref).max() <= 1e-9 * np.abs(ref).max(), name
        assert abs(err_g - float(f[f"{name}_err"])) <= 1e-3 * float(f[f"{name}_err"]) or it_g != it, name


def _stokes(torch, sp, nd, rku, levmax=7, nterms=43, mode=0):
    B = sp.Bodies(nd)
    geo = B.extractdat(torch.tensor(rku, device="cuda"))
    mv = sp.MatVec(B, geo["x"], geo["y"], geo["dz"], geo["dsda"], geo["rkappa"], geo["zk"], levmax=levmax,
                   nterms=nterms, mode=mode)
    return B, geo, mv


@pytest.mark.parametrize("tag", ["two_drops", "16drops"])
def test_stokes_solve_matches_reference(torch, mods, tag):
    gm, sp = mods
    f = g("gmres_stokes_" + tag)
    print("stokes solve", tag, flush=True)
    B, geo, mv = _stokes(torch, sp, f["nd"], f["rku"])
    rhs = torch.tensor(f["rhs"], device="cuda")
    solver = gm.GMRES(rhs.numel(), 100)
    soln, za, pk, it, err, ierr = gm.stokes_solve(mv, solver, rhs)
    it_ref, ierr_ref = f["meta"].tolist()
    assert ierr == ierr_ref == 0 and abs(it - it_ref) <= 1, (it, it_ref)
    assert err <= 1e-10
    x = soln.cpu().numpy()
    assert np.linalg.norm(x - f["x"]) <= 1e-9 * np.linalg.norm(f["x"])
    assert np.linalg.norm(x - f["u_true"]) <= 1e-9 * np.linalg.norm(f["u_true"])
    # za = h sum u is a cancelling sum (zero-mean density modes): compare on the scale of h sum |u|
    assert np.abs(za.cpu().numpy() - f["za"]).max() <= 1e-9 * 2 * np.pi * np.abs(x).max()
    assert np.abs(pk.cpu().numpy() - f["pk"]).max() <= 1e-9 * max(1.0, np.abs(f["pk"]).max())
    # the true residual of the returned density, through an independent operator call
    ax = mv.apply(soln).cpu().numpy()
    assert np.linalg.norm(ax - f["rhs"]) <= 1.5e-10 * np.linalg.norm(f["rhs"])
    solver.close()


def test_stokes_solve_64_drops(torch, mods):
    """The many-drop shape (64 drops of 128..1024 points, about 32K unknowns pairs) on the FMM."""
    gm, sp = mods
    nd, rku = many_drops_cfg(64, np.random.default_rng(640))
    print("stokes solve 64 drops, points", int(nd.sum()), flush=True)
    B, geo, mv = _stokes(torch, sp, nd, rku)
    th = np.concatenate([2 * np.pi * np.arange(m) / m for m in nd])
    u_true = torch.tensor(np.stack([np.cos(th), 0.5 * np.sin(2 * th)], 1).ravel(), device="cuda")
    rhs = mv.apply(u_true)
    solver = gm.GMRES(rhs.numel(), 100)
    soln, za, pk, it, err, ierr = gm.stokes_solve(mv, solver, rhs)
    assert ierr == 0 and err <= 1e-10 and 5 <= it <= 100
    res = torch.linalg.norm(mv.apply(soln) - rhs) / torch.linalg.norm(rhs)
    assert float(res) <= 1.5e-10
    assert float(torch.linalg.norm(soln - u_true) / torch.linalg.norm(u_true)) <= 1e-8
    solver.close()
