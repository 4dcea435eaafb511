"""The unchanged reference drivers on the GPU FMM.

oracle/_ref/libref_on_eb.so holds the reference's own objects with cadapdrop.o, cadapmod.o and
cadap.o left out, linked against the ISO_C_BINDING shim (electricbubbles_amd/libeb_fortran.so):
FASMVP, INTERFACENEW, FASMVP_LAP, PVINTEV, F and RK2 run exactly as written and reach the GPU
through the CADAPDR / CADAPMOD names and, with the reference's own DAPIF2 made local to its
object, through the DAPIF2 name.  The same calls on oracle/_ref/libref.so (the reference as written)
are the checker.
"""
import os

import numpy as np
import pytest

import oracle
from electricbubbles_amd.workloads import perturbed_circle
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _need_libs():
    for p in (oracle.LIBREF, oracle.LIBREF_ON_EB):
        if not os.path.exists(p):
            pytest.skip(f"{p} not built (it is compiled from the reference sources)")


@pytest.fixture(scope="module")
def sims(tmp_path_factory):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _need_libs()
    cwd = os.getcwd()
    os.chdir(tmp_path_factory.mktemp("fortran_units"))  # the reference writes fort.13 here
    try:
        yield oracle.RefSim(oracle.LIBREF), oracle.RefSim(oracle.LIBREF_ON_EB)
    finally:
        os.chdir(cwd)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(params=["auto", "fmm"])
def dropin_mode(request):
    """The drop-in layer's mode: "auto" takes the O(N^2) sum at these sizes (N <= EB_DIRECT_MAX);
    "fmm" sends every CADAPDR / CADAPMOD call of the unchanged drivers to the GPU FMM at the
    LEVMAX and TOL they pass (stokesbubb.f:564-571, electric.f:835-837)."""
    import electricbubbles_amd as eb
    mode = {"auto": eb.MODE_AUTO, "fmm": eb.MODE_FMM}[request.param]
    eb.set_mode(mode)
    print("drop-in mode", request.param, flush=True)
    try:
        yield mode
    finally:
        eb.set_mode(eb.MODE_AUTO)


def test_fasmvp_through_shim_matches_reference(sims, dropin_mode):
    ref, gpu = sims
    g = np.load(os.path.join(GOLDEN, "geometry_two_drops.npz"))
    args = (g["nd"], g["x"], g["y"], g["dz"], g["dsda"], g["rkappa"], g["zk"], g["u"])
    w_ref, s_ref = ref.fasmvp(*args)
    w_gpu, s_gpu = gpu.fasmvp(*args)
    assert rel(w_gpu, w_ref) < 1e-10
    assert rel(w_gpu, g["w"]) < 1e-10


def test_single_drop_trajectory_matches_reference(sims, dropin_mode):
    """Config-1 drop r = 1 + 0.05 cos 2 theta, 256 points, 10 Heun steps of dt 1e-3: the
    reference's F and RK2 with the GPU FMM against the reference as written (shapes 1e-8)."""
    ref, gpu = sims
    z, _, nd, _ = perturbed_circle(256, [(2, 0.05)])
    y0 = z.view(np.float64).copy()
    out = []
    for s in (ref, gpu):
        s.setup(nd, cap=0.0, e0=1.0)
        out.append(s.rk2(y0, 1e-3, 10))
    y_ref, y_gpu = out
    assert np.abs(y_gpu - y_ref).max() < 1e-8
    r_ref = np.abs(y_ref.view(np.complex128))
    r_gpu = np.abs(y_gpu.view(np.complex128))
    d_ref = (r_ref.max() - r_ref.min()) / (r_ref.max() + r_ref.min())
    d_gpu = (r_gpu.max() - r_gpu.min()) / (r_gpu.max() + r_gpu.min())
    assert abs(d_gpu - d_ref) < 1e-8


@pytest.mark.parametrize("tag", ["two_drops", "four_circles", "underresolved"])
def test_laplace_operator_through_shim_matches_reference(sims, dropin_mode, tag):
    """FASMVP_LAP (laplace.f:159-262) as written, its DAPIF2 call (laplace.f:214-215) on the GPU,
    against the reference's own output on the same smooth density."""
    ref, gpu = sims
    f = np.load(os.path.join(GOLDEN, f"field_{tag}.npz"))
    nd = f["nd"]
    geo = oracle.extractdat(nd, f["rku"])
    outs = []
    for s in (ref, gpu):
        outs.append(s.fasmvp_lap(nd, geo["zk"], geo["x"] + 1j * geo["y"], geo["dz"], geo["rkappa"], geo["dsda"],
                                 f["lap_u"]))
    assert rel(outs[1], outs[0]) < 1e-10
    assert rel(outs[1], f["lap_w"]) < 1e-10


@pytest.mark.parametrize("tag", ["two_drops", "underresolved"])
def test_whole_rhs_through_shim_matches_reference(sims, dropin_mode, tag):
    """The reference's F as written, with every CADAPDR, CADAPMOD and DAPIF2 call on the GPU."""
    ref, gpu = sims
    f = np.load(os.path.join(GOLDEN, f"field_{tag}.npz"))
    for s in (ref, gpu):
        s.setup(f["nd"], cap=float(f["cap"]), e0=float(f["e0"]))
    r, g = ref.rhs(f["rku"]), gpu.rhs(f["rku"])
    assert rel(g, r) < 1e-8
    assert rel(g, f["rkup"]) < 1e-8


# ------------------------------------------------------------------------------------------
# The single drop r = 1 + 0.05 cos 2 theta, 256 points, |E| = 1 (E0 = 1), no straining flow,
# over 100 steps of dt 1e-3, with the deformation D = (L - B) / (L + B) after every step.
# ------------------------------------------------------------------------------------------

def _cfg1():
    z, _, nd, _ = perturbed_circle(256, [(2, 0.05)])
    return nd, z.view(np.float64).copy()


def test_single_drop_100_heun_steps_match_reference(sims):
    """The reference's RK2 driver as written, its CADAPDR / CADAPMOD / DAPIF2 calls on the GPU,
    against the reference as written: D after every one of 100 steps within 1e-8, the final
    shape within 1e-8."""
    ref, gpu = sims
    nd, y0 = _cfg1()
    series = []
    for s in (ref, gpu):
        s.setup(nd, cap=0.0, e0=1.0)
        y = y0.copy()
        D = [oracle.deformation(nd, y)[0]]
        for step in range(100):
            y = s.rk2(y, 1e-3, 1, t0=step * 1e-3)
            D.append(oracle.deformation(nd, y)[0])
        series.append((y, np.array(D)))
        print("D after 100 Heun steps", D[-1], flush=True)
    (y_ref, d_ref), (y_gpu, d_gpu) = series
    assert np.abs(d_gpu - d_ref).max() < 1e-8
    assert np.abs(y_gpu - y_ref).max() < 1e-8
    assert d_ref[-1] > d_ref[0]  # the field stretches the drop along x


def test_single_drop_100_rk4_steps_match_reference_f(sims):
    """The device field's captured RK4 step (eb_field_rk4_graph) against classic RK4 around the
    reference's own F: D after every one of 100 steps (eb_deformation on the device) within 1e-8,
    the final shape within 1e-8; eb_deformation against the oracle's D of the same state."""
    import torch

    from electricbubbles_amd import field, spectral
    ref, _ = sims
    nd, y0 = _cfg1()
    ref.setup(nd, cap=0.0, e0=1.0)
    d_ref = [oracle.deformation(nd, y0)[0]]
    y_ref = oracle.rk4(ref.rhs, y0, 1e-3, 100, record=lambda k, y: d_ref.append(oracle.deformation(nd, y)[0]))
    st = torch.cuda.Stream()
    B = spectral.Bodies(nd, stream=st)
    F = field.Field(B, cap=0.0, e0=1.0, stream=st)
    y = torch.tensor(y0, device="cuda")
    d_gpu = [float(B.deformation(y)[0])]
    with torch.cuda.stream(st):
        for step in range(100):
            assert not F.rk4_graph(y, 1e-3, budget=40)
            d_gpu.append(float(B.deformation(y)[0]))
    st.synchronize()
    d_gpu = np.array(d_gpu)
    yg = y.cpu().numpy()
    print("D after 100 RK4 steps", d_gpu[-1], "reference F", d_ref[-1], flush=True)
    assert abs(d_gpu[-1] - oracle.deformation(nd, yg)[0]) < 1e-13
    assert np.abs(d_gpu - np.array(d_ref)).max() < 1e-8
    assert np.abs(yg - y_ref).max() < 1e-8
