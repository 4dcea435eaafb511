"""GPU parity of F (include/eb_field.h) with the reference's F, stage by stage.

Checkers: the compiled reference's own stages on three geometries (tests/golden/field_*.npz,
written by tests/golden/make_golden.py field from oracle.ref_field_chain, whose chain is
checked against the reference's whole F there): FASMVP_LAP on a smooth density, SOLVE_LAPL,
DIRICH_NEU_MAP, SETVELF, SOLVEI and INTERFACENEW, each fed the reference's own input to the
stage, then the whole F and three RK2 steps.
"""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

# "underresolved": bodies of 32 and 48 points with content at mode N/2 - 1, where PVINTEV's
# real-data interpolation (FINTER, pv_integral.f:216-252) differs from the complex one
CASES = ["single_drop", "two_drops", "four_circles", "underresolved"]
# "auto" takes the O(N^2) sums at these sizes (EB_DIRECT_MAX); "fmm" forces every layer potential
# of F (the Laplace and PVINTEV Hilbert products, the Stokes matvec, CADAPMOD) onto the FMM at
# the reference's LEVMAX 7, the path F takes above EB_DIRECT_MAX points
MODES = ["auto", "fmm"]


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no GPU")
    return _t


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def setup(torch, tag, mode="auto"):
    import electricbubbles_amd as eb
    from electricbubbles_amd import field, spectral
    f = np.load(os.path.join(GOLDEN, f"field_{tag}.npz"))
    print("field case", tag, mode, list(f["nd"]), flush=True)
    B = spectral.Bodies(f["nd"])
    F = field.Field(B, cap=float(f["cap"]), e0=float(f["e0"]))
    F.set_mode({"auto": eb.MODE_AUTO, "fmm": eb.MODE_FMM}[mode])
    rku = torch.tensor(f["rku"], device="cuda")
    F.set_geometry(rku)
    return f, B, F, rku


def dev(torch, a):
    return torch.tensor(np.ascontiguousarray(a), device="cuda")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tag", CASES)
def test_laplace_operator(torch, tag, mode):
    f, B, F, rku = setup(torch, tag, mode)
    w = F.laplace_apply(dev(torch, f["lap_u"])).cpu().numpy()
    assert rel(w, f["lap_w"]) < 1e-10


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tag", CASES)
def test_laplace_solve(torch, tag, mode):
    f, B, F, rku = setup(torch, tag, mode)
    u, it, ierr = F.laplace_solve(rku)
    assert ierr == 0
    assert rel(u.cpu().numpy(), f["ulap"]) < 1e-8


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tag", CASES)
def test_dirichlet_neumann_map(torch, tag, mode):
    f, B, F, rku = setup(torch, tag, mode)
    phi, psi, dphi = F.dirich_neu(dev(torch, f["ulap"]))
    assert rel(phi.cpu().numpy(), f["phi"]) < 1e-10
    assert rel(psi.cpu().numpy(), f["psi"]) < 1e-10
    assert rel(dphi.cpu().numpy(), f["dphi_dn"]) < 1e-9


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tag", CASES)
def test_setvelf(torch, tag, mode):
    f, B, F, rku = setup(torch, tag, mode)
    rhs = F.setvelf(dev(torch, f["dphi_dn"])).cpu().numpy()
    assert rel(rhs, f["rhs_stokes"]) < 1e-12


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tag", CASES)
def test_stokes_solve(torch, tag, mode):
    f, B, F, rku = setup(torch, tag, mode)
    u, za, pk, it, ierr = F.stokes_solve(dev(torch, f["rhs_stokes"]))
    assert ierr == 0
    assert rel(u.cpu().numpy(), f["u"]) < 1e-9
    assert np.abs(pk.cpu().numpy() - f["pk"]).max() <= 1e-9 * max(1.0, np.abs(f["pk"]).max())


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tag", CASES)
def test_interface_velocity(torch, tag, mode):
    f, B, F, rku = setup(torch, tag, mode)
    un = F.interface(dev(torch, f["u"])).cpu().numpy()
    assert rel(un, f["un"]) < 1e-10


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tag", CASES)
def test_whole_rhs(torch, tag, mode):
    f, B, F, rku = setup(torch, tag, mode)
    rkup = F.rhs(rku).cpu().numpy()
    assert rel(rkup, f["rkup"]) < 1e-8


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tag", CASES)
def test_rk2_steps(torch, tag, mode):
    """Three Heun steps (electric.f:762-786) with F on the GPU against the reference's."""
    from electricbubbles_amd import spectral
    f, B, F, rku = setup(torch, tag, mode)
    dt = float(f["rk2_dt"])
    ybar = rku.clone()
    for _ in range(int(f["rk2_steps"])):
        phi0, y = spectral.rk2_stage1(dt, ybar, F.rhs(ybar))
        spectral.rk2_stage2(dt, phi0, F.rhs(y), ybar)
    assert rel(ybar.cpu().numpy(), f["rk2"]) < 1e-10
    assert np.abs(ybar.cpu().numpy() - f["rk2"]).max() < 1e-8


def test_sweep_field_matches_reference(torch):
    """The three geometries as three simulations of one sweep (each with its own CAP and E0):
    F of each against the reference's F, and three RK2 steps against the reference's RK2."""
    from electricbubbles_amd import field, spectral
    fs = [np.load(os.path.join(GOLDEN, f"field_{t}.npz")) for t in CASES]
    nd = np.concatenate([f["nd"] for f in fs]).astype(np.int32)
    sim_off = np.concatenate([[0], np.cumsum([f["nd"].size for f in fs])]).astype(np.int32)
    B = spectral.Bodies(nd)
    SF = field.SweepField(B, sim_off, [float(f["cap"]) for f in fs], [float(f["e0"]) for f in fs])
    rku = torch.tensor(np.concatenate([f["rku"] for f in fs]), device="cuda")
    rkup = SF.rhs(rku).cpu().numpy()
    o = 0
    for f in fs:
        m = f["rku"].size
        assert rel(rkup[o:o + m], f["rkup"]) < 1e-8
        o += m
    lap_it, lap_ie, st_it, st_ie = SF.iterations()
    assert np.all(lap_ie == 0) and np.all(st_ie == 0)
    dt = float(fs[0]["rk2_dt"])
    ybar = rku.clone()
    for _ in range(int(fs[0]["rk2_steps"])):
        phi0, y = spectral.rk2_stage1(dt, ybar, SF.rhs(ybar))
        spectral.rk2_stage2(dt, phi0, SF.rhs(y), ybar)
    yb = ybar.cpu().numpy()
    o = 0
    for f in fs:
        m = f["rku"].size
        assert np.abs(yb[o:o + m] - f["rk2"]).max() < 1e-8
        o += m


def test_sweep_field_matches_plain_fields(torch):
    """Simulations of one to three drops with uneven lengths and parameters: the sweep's F of
    each equals the plain per-simulation F (both take the O(N^2) sums at these sizes)."""
    from electricbubbles_amd import field, spectral
    from electricbubbles_amd.workloads import perturbed_circle
    rng = np.random.default_rng(21)
    nds, zs, sim_off, caps, e0s = [], [], [0], [], []
    for nb in [1, 2, 1, 3, 1]:
        for b in range(nb):
            n = int(rng.choice([128, 130, 194, 256, 322]))
            zs.append(perturbed_circle(n, [(2, 0.05), (int(rng.integers(3, 7)), rng.uniform(-0.03, 0.03))],
                                       center=3.0 * b + 0.4j * b)[0])
            nds.append(n)
        sim_off.append(sim_off[-1] + nb)
        caps.append(rng.uniform(0.0, 0.2))
        e0s.append(rng.uniform(0.1, 2.0))
    nd = np.array(nds, np.int32)
    B = spectral.Bodies(nd)
    SF = field.SweepField(B, sim_off, caps, e0s)
    rku = torch.tensor(np.concatenate(zs).view(np.float64).copy(), device="cuda")
    got = SF.rhs(rku).cpu().numpy()
    off = np.concatenate([[0], np.cumsum(nd)])
    for s in range(len(caps)):
        b0, b1 = sim_off[s], sim_off[s + 1]
        Bs = spectral.Bodies(nd[b0:b1])
        F = field.Field(Bs, cap=caps[s], e0=e0s[s])
        ref = F.rhs(rku[2 * off[b0]:2 * off[b1]].contiguous()).cpu().numpy()
        assert rel(got[2 * off[b0]:2 * off[b1]], ref) < 1e-10, s
        F.close()


def test_sweep_field_512_single_drops(torch):
    """The batched-sweep shapes (512 single drops, N log-uniform 128..2048 rounded to even,
    E0 = |E|^2 with |E| uniform in 0.1..2 and CAP log-uniform in 0.01..0.5 per simulation):
    every solve converges and F is finite."""
    from electricbubbles_amd import field, spectral
    from electricbubbles_amd.workloads import perturbed_circle
    rng = np.random.default_rng(5120)
    nd = (np.round(np.exp(rng.uniform(np.log(128), np.log(2048), 512)) / 2) * 2).astype(np.int32)
    print("sweep field, 512 drops, points", int(nd.sum()), flush=True)
    B = spectral.Bodies(nd)
    e0 = rng.uniform(0.1, 2.0, 512) ** 2
    cap = np.exp(rng.uniform(np.log(0.01), np.log(0.5), 512))  # Ca log-uniform in 0.01..0.5
    SF = field.SweepField(B, np.arange(513), cap, e0)
    rku = torch.tensor(np.concatenate([perturbed_circle(int(n), [(2, 0.05)])[0] for n in nd]).view(np.float64).copy(),
                       device="cuda")
    out = SF.rhs(rku)
    assert bool(torch.isfinite(out).all())
    lap_it, lap_ie, st_it, st_ie = SF.iterations()
    assert np.all(lap_ie == 0) and np.all(st_ie == 0)


@pytest.mark.parametrize("mode", MODES)
def test_many_drops_rhs_and_rk2(torch, mode):
    """Eight drops of the many-drop construction (256..1024 points each, E0 = 2.25): the whole
    F and two RK2 steps at the configuration's dt against the reference's."""
    import electricbubbles_amd as eb
    from electricbubbles_amd import field, spectral
    f = np.load(os.path.join(GOLDEN, "field_many_drops_8.npz"))
    print("many drops", mode, list(f["nd"]), flush=True)
    B = spectral.Bodies(f["nd"])
    F = field.Field(B, cap=float(f["cap"]), e0=float(f["e0"]))
    F.set_mode({"auto": eb.MODE_AUTO, "fmm": eb.MODE_FMM}[mode])
    rku = torch.tensor(f["rku"], device="cuda")
    assert rel(F.rhs(rku).cpu().numpy(), f["rkup"]) < 1e-8
    dt = float(f["rk2_dt"])
    ybar = rku.clone()
    for _ in range(int(f["rk2_steps"])):
        phi0, y = spectral.rk2_stage1(dt, ybar, F.rhs(ybar))
        spectral.rk2_stage2(dt, phi0, F.rhs(y), ybar)
    assert np.abs(ybar.cpu().numpy() - f["rk2"]).max() < 1e-8
