"""GPU checks of the adjoint (include/eb_adjoint.h, electricbubbles_amd/adjoint.py).

Checkers:
* the forward operators against the reference's own FASMVP_LAP output and SOLVEI solution
  (tests/golden/field_*.npz) and against the library's Stokes matvec (direct mode);
* the transposes by <A x, y> = <x, A^T y>;
* every pair-sum reverse against central differences of the forward pair sum;
* the differentiable F against the reference's F and its Heun steps (field_*.npz);
* the gradient of the summed deformation against central differences of the same GPU forward,
  and against central differences of the reference's own F (tests/golden/adjoint_*.npz).
"""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no GPU")
    return _t


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def field_case(torch, tag):
    f = np.load(os.path.join(GOLDEN, f"field_{tag}.npz"))
    nd = [int(v) for v in f["nd"]]
    print("adjoint case", tag, nd, flush=True)
    z = f["rku"][0::2] + 1j * f["rku"][1::2]
    return f, nd, z


def system_on(torch, nd, z):
    """An AdjointSystem with the geometry of z from the library's EXTRACTDAT."""
    from electricbubbles_amd import adjoint, spectral
    B = spectral.Bodies(nd)
    rku = torch.tensor(np.stack([z.real, z.imag], 1).ravel(), device="cuda")
    g = B.extractdat(rku)
    sys = adjoint.AdjointSystem(nd)
    geo = dict(z=torch.tensor(z, device="cuda"), dz=g["dz"], s=g["dsda"], kap=g["rkappa"], zk=g["zk"])
    sys.set_geometry(geo["z"], geo["dz"], geo["s"], geo["kap"], geo["zk"])
    return sys, B, g, geo


@pytest.mark.parametrize("tag", ["single_drop", "two_drops", "four_circles"])
def test_operators_match_reference(torch, tag):
    from electricbubbles_amd import spectral
    f, nd, z = field_case(torch, tag)
    sys, B, g, geo = system_on(torch, nd, z)
    # FASMVP_LAP on the reference's density
    w = sys.laplace_apply(torch.tensor(f["lap_u"], device="cuda")).cpu().numpy()
    assert rel(w, f["lap_w"]) < 1e-10
    # SOLVE_LAPL and SOLVEI solutions against the reference's
    u = sys.laplace_solve(torch.tensor(f["rhs_lap"], device="cuda"), 1e-10)
    assert rel(u.cpu().numpy(), f["ulap"]) < 1e-8
    s = sys.stokes_solve(torch.tensor(f["rhs_stokes"], device="cuda"), 1e-10)
    assert rel(s.cpu().numpy().view(np.complex128), f["u"]) < 1e-8
    # the Stokes operator against the library's matvec (O(N^2) mode) on a random density
    rng = np.random.default_rng(3)
    x = torch.tensor(rng.standard_normal(2 * sum(nd)), device="cuda")
    mv = spectral.MatVec(B, torch.tensor(z.real.copy(), device="cuda"), torch.tensor(z.imag.copy(), device="cuda"),
                         g["dz"], g["dsda"], g["rkappa"], g["zk"], mode=2)
    ref = mv.apply(x).cpu().numpy()
    assert rel(sys.stokes_apply(x).cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("tag", ["two_drops", "four_circles"])
def test_transposes(torch, tag):
    f, nd, z = field_case(torch, tag)
    sys, B, g, geo = system_on(torch, nd, z)
    rng = np.random.default_rng(5)
    n = sum(nd)
    for m, ap in ((n, sys.laplace_apply), (2 * n, sys.stokes_apply)):
        x = torch.tensor(rng.standard_normal(m), device="cuda")
        y = torch.tensor(rng.standard_normal(m), device="cuda")
        lhs = float(torch.dot(ap(x), y))
        rhs_ = float(torch.dot(x, ap(y, transpose=True)))
        assert abs(lhs - rhs_) < 1e-12 * abs(lhs) + 1e-13 * float(x.norm() * y.norm()), (lhs, rhs_)
    # the transposed solves invert the transposed operators
    b = torch.tensor(rng.standard_normal(2 * n), device="cuda")
    lam = sys.stokes_solve(b, 1e-13, transpose=True)
    assert rel(sys.stokes_apply(lam, transpose=True).cpu().numpy(), b.cpu().numpy()) < 1e-12
    b = torch.tensor(rng.standard_normal(n), device="cuda")
    lam = sys.laplace_solve(b, 1e-13, transpose=True)
    assert rel(sys.laplace_apply(lam, transpose=True).cpu().numpy(), b.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("tag", ["two_drops", "four_circles"])
def test_fmm_mode_operators(torch, tag):
    """Operators on FMM plans (eb_adj_set_mode(FMM): the transposed ones through the transposed
    FMM product) against the O(N^2) pair sums: both directions, the transposes by
    <A x, y> = <x, A^T y>, and the transposed solves."""
    f, nd, z = field_case(torch, tag)
    dsys = system_on(torch, nd, z)[0]
    fsys = system_on(torch, nd, z)[0]
    fsys.set_mode(1)
    rng = np.random.default_rng(8)
    n = sum(nd)
    for m, dap, fap in ((n, dsys.laplace_apply, fsys.laplace_apply), (2 * n, dsys.stokes_apply, fsys.stokes_apply)):
        x = torch.tensor(rng.standard_normal(m), device="cuda")
        y = torch.tensor(rng.standard_normal(m), device="cuda")
        for tr in (False, True):
            e = rel(fap(x, transpose=tr).cpu().numpy(), dap(x, transpose=tr).cpu().numpy())
            print(f"{tag} m={m} transpose={tr}: FMM against pairs {e:.2e}", flush=True)
            assert e < 1e-11
        lhs = float(torch.dot(fap(x), y))
        rhs_ = float(torch.dot(x, fap(y, transpose=True)))
        assert abs(lhs - rhs_) < 1e-11 * float(x.norm() * y.norm()), (lhs, rhs_)
    b = torch.tensor(rng.standard_normal(2 * n), device="cuda")
    lam_f = fsys.stokes_solve(b, 1e-13, transpose=True)
    lam_d = dsys.stokes_solve(b, 1e-13, transpose=True)
    assert rel(lam_f.cpu().numpy(), lam_d.cpu().numpy()) < 1e-10


def test_auto_mode_takes_the_fmm_above_the_direct_limit(torch):
    """Eight drops of 2048 points (16384 > 10240): the default mode gives the FMM mode's bits,
    within 1e-11 of the pair sums, in both directions."""
    nd = [2048] * 8
    th = 2 * np.pi * np.arange(2048) / 2048
    z = np.concatenate([3 * (b % 4) + 3j * (b // 4) + (1.0 + 0.05 * np.cos(2 * th)) * np.exp(1j * th)
                        for b in range(8)])
    asys = system_on(torch, nd, z)[0]
    fsys = system_on(torch, nd, z)[0]
    fsys.set_mode(1)
    dsys = system_on(torch, nd, z)[0]
    dsys.set_mode(2)
    x = torch.tensor(np.random.default_rng(2).standard_normal(2 * z.size), device="cuda")
    for tr in (False, True):
        a = asys.stokes_apply(x, transpose=tr).cpu().numpy()
        np.testing.assert_array_equal(a, fsys.stokes_apply(x, transpose=tr).cpu().numpy())
        e = rel(a, dsys.stokes_apply(x, transpose=tr).cpu().numpy())
        print(f"16384 points, transpose={tr}: FMM against pairs {e:.2e}", flush=True)
        assert e < 1e-11


def _cplx(torch, rng, n, scale=1.0):
    return torch.tensor(scale * (rng.standard_normal(n) + 1j * rng.standard_normal(n)), device="cuda")


def _spread(torch, rng, n, shift=0.0, jitter=0.3):
    """n points of a jittered unit grid (pairs at least 0.4 apart): central differences with
    h = 1e-6 then resolve the pair kernels' derivatives to about 1e-9."""
    side = int(np.ceil(np.sqrt(n)))
    ij = np.arange(n)
    p = (ij % side) + 1j * (ij // side) + jitter * (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)) + shift
    return torch.tensor(p, device="cuda")


def _check_vjp(torch, f, inputs, grads, rng, h=1e-6):
    """<grads, v> against the central difference of L(inputs + h v) along random directions v."""
    vs = [_cplx(torch, rng, x.numel(), 0.1) for x in inputs]
    plus = f(*[x + h * v for x, v in zip(inputs, vs)])
    minus = f(*[x - h * v for x, v in zip(inputs, vs)])
    fd = (plus - minus) / (2 * h)
    an = sum(float((gr.real * v.real + gr.imag * v.imag).sum()) for gr, v in zip(grads, vs))
    assert abs(an - fd) < 1e-7 * max(1.0, abs(fd)), (an, fd)


@pytest.mark.parametrize("self_sum", [True, False])
def test_cauchy_reverse(torch, self_sum):
    from electricbubbles_amd import adjoint
    rng = np.random.default_rng(7)
    nt = 300
    ns = nt if self_sum else 517
    sys = adjoint.AdjointSystem([nt])
    t = _spread(torch, rng, nt, jitter=0.3 if self_sum else 0.1)
    s = t.clone() if self_sum else _spread(torch, rng, ns, shift=0.5 + 0.5j, jitter=0.1)
    c = _cplx(torch, rng, ns)
    W = _cplx(torch, rng, nt)

    if self_sum:
        def L(tt, cc):
            g = sys.cauchy(tt, tt, cc, True)
            return float((W.conj() * g).real.sum())
        tb, sb, cb = sys.cauchy_vjp(t, t, c, True, W)
        _check_vjp(torch, L, [t, c], [tb + sb, cb], rng)
    else:
        def L(tt, ss, cc):
            g = sys.cauchy(tt, ss, cc, False)
            return float((W.conj() * g).real.sum())
        tb, sb, cb = sys.cauchy_vjp(t, s, c, False, W)
        _check_vjp(torch, L, [t, s, c], [tb, sb, cb], rng)


@pytest.mark.parametrize("s2", [-1.0, 1.0])
def test_stokes_pairs_reverse(torch, s2):
    from electricbubbles_amd import adjoint
    import electricbubbles_amd as eb
    rng = np.random.default_rng(11)
    n = 400
    sys = adjoint.AdjointSystem([n])
    z = _spread(torch, rng, n)
    qa, qb, W = (_cplx(torch, rng, n) for _ in range(3))

    def L(zz, a, b):
        return float((W.conj() * sys.stokes_pairs(zz, a, b, s2)).real.sum())
    zb, qab, qbb = sys.stokes_pairs_vjp(z, qa, qb, s2, W)
    _check_vjp(torch, L, [z, qa, qb], [zb, qab, qbb], rng)


def test_stokes_pairs_match_library_kernel(torch):
    """The pair sum equals the library's O(N^2) CADAPDR / CADAPMOD evaluation."""
    from electricbubbles_amd import adjoint
    import electricbubbles_amd as eb
    rng = np.random.default_rng(13)
    n = 700
    z = rng.uniform(0, 1, n) + 1j * rng.uniform(0, 1, n)
    dens = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    wdz = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    sys = adjoint.AdjointSystem([n])
    zt, dt_, wt = (torch.tensor(a, device="cuda") for a in (z, dens, wdz))
    qa, qb = adjoint.stokes_charges(dt_, wt, 1.0)
    for variant, s2 in ((eb.VARIANT_DR, -1.0), (eb.VARIANT_MOD, 1.0)):
        plan = eb.FMMPlan(torch.tensor(z.real.copy(), device="cuda"), torch.tensor(z.imag.copy(), device="cuda"), wt,
                          5, 43, variant=variant, mode=eb.MODE_DIRECT)
        ref = plan.apply(dt_).cpu().numpy()
        got = sys.stokes_pairs(zt, qa, qb, s2).cpu().numpy()
        assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("tag", ["single_drop", "four_circles"])
def test_rhs_and_heun_match_reference(torch, tag):
    """The differentiable F against the reference's F and its three Heun steps."""
    from electricbubbles_amd import adjoint
    f, nd, z = field_case(torch, tag)
    P = adjoint.Problem(len(nd), nd[0], tol=1e-10)
    zt = torch.tensor(z.reshape(len(nd), nd[0]), device="cuda")
    F = adjoint.rhs(P, zt, float(f["e0"]), float(f["cap"])).reshape(-1).cpu().numpy()
    ref = f["rkup"][0::2] + 1j * f["rkup"][1::2]
    assert rel(F, ref) < 1e-8
    dt = float(f["rk2_dt"])
    for _ in range(int(f["rk2_steps"])):
        zt = adjoint.step(P, zt, dt, float(f["e0"]), float(f["cap"]), "rk2")
    zr = f["rk2"][0::2] + 1j * f["rk2"][1::2]
    assert rel(zt.reshape(-1).cpu().numpy(), zr) < 1e-10


def _small_problem(torch, k_side=2, N=64, nmodes=8, tol=1e-13):
    from electricbubbles_amd import adjoint
    centres, modes = adjoint.grid_problem(k_side=k_side, N=N, spacing=3.0, nmodes=nmodes, amp=0.03, seed=2)
    P = adjoint.Problem(k_side * k_side, N, tol=tol)
    return P, centres, modes


def test_gradient_matches_gpu_forward_differences(torch):
    from electricbubbles_amd import adjoint
    P, centres, modes = _small_problem(torch)
    E, Ca, steps, dt = 1.0, 0.1, 4, 1e-3
    loss, gE, gCa, gm, _ = adjoint.gradient(P, centres, modes, E, Ca, steps, dt, checkpoint=3)
    lf = adjoint.forward_loss(P, centres, modes, E, Ca, steps, dt)
    assert abs(loss - lf) < 1e-14 * abs(lf) + 1e-15
    h = 1e-5

    def rich(fn):
        # central differences at h and h/2, Richardson-extrapolated (error O(h^4))
        return (4 * (fn(h / 2) - fn(-h / 2)) / h - (fn(h) - fn(-h)) / (2 * h)) / 3
    fd_E = rich(lambda e: adjoint.forward_loss(P, centres, modes, E + e, Ca, steps, dt))
    fd_Ca = rich(lambda e: adjoint.forward_loss(P, centres, modes, E, Ca + e, steps, dt))
    rng = np.random.default_rng(1)
    v = torch.tensor(rng.standard_normal(modes.shape), device="cuda")
    v /= v.norm()
    fd_m = rich(lambda e: adjoint.forward_loss(P, centres, modes + e * v, E, Ca, steps, dt))
    an_m = float((gm * v).sum())
    print("grad E", gE, fd_E, "Ca", gCa, fd_Ca, "modes", an_m, fd_m, flush=True)
    assert abs(gE - fd_E) < 1e-6 * abs(fd_E)
    assert abs(gCa - fd_Ca) < 1e-6 * abs(fd_Ca)
    assert abs(an_m - fd_m) < 1e-6 * max(abs(fd_m), 1e-3 * float(gm.norm()))


def test_checkpointing_does_not_change_the_gradient(torch):
    """Checkpoints every step, every two steps (uneven last segment) and one segment for the
    whole run give the same gradient."""
    from electricbubbles_amd import adjoint
    P, centres, modes = _small_problem(torch, nmodes=4)
    a = adjoint.gradient(P, centres, modes, 1.0, 0.1, 5, 1e-3, checkpoint=1)
    for ck in (2, 5):
        b = adjoint.gradient(P, centres, modes, 1.0, 0.1, 5, 1e-3, checkpoint=ck)
        assert abs(a[0] - b[0]) < 1e-14 * abs(a[0]), ck
        assert abs(a[1] - b[1]) < 1e-13 * abs(a[1]), ck
        assert abs(a[2] - b[2]) < 1e-13 * abs(a[2]), ck
        assert rel(a[3].cpu().numpy(), b[3].cpu().numpy()) < 1e-13, ck


def test_gradient_matches_reference_differences(torch):
    """dL/d|E|, dL/dCa and three directional derivatives in the shape modes against central
    differences (h = 1e-4) of the same loss whose F is the reference's own."""
    from electricbubbles_amd import adjoint
    f = np.load(os.path.join(GOLDEN, "adjoint_fd_four_drops.npz"))
    N, steps, dt = int(f["N"]), int(f["steps"]), float(f["dt"])
    centres = torch.tensor(f["centres"], device="cuda")
    modes = torch.tensor(f["modes"], device="cuda")
    P = adjoint.Problem(centres.numel(), N, tol=1e-13)
    loss, gE, gCa, gm, _ = adjoint.gradient(P, centres, modes, float(f["E"]), float(f["Ca"]), steps, dt, checkpoint=2)
    print("loss", loss, float(f["loss"]), "dE", gE, float(f["fd_E"]), "dCa", gCa, float(f["fd_Ca"]), flush=True)
    assert abs(loss - float(f["loss"])) < 1e-9 * abs(float(f["loss"]))
    assert abs(gE - float(f["fd_E"])) < 1e-6 * abs(float(f["fd_E"]))
    assert abs(gCa - float(f["fd_Ca"])) < 1e-6 * abs(float(f["fd_Ca"]))
    gmn = gm.cpu().numpy()
    for v, fd in zip(f["dirs"], f["fd_dirs"]):
        an = float((gmn * v).sum())
        assert abs(an - fd) < 1e-6 * max(abs(fd), 1e-2 * np.linalg.norm(gmn)), (an, fd)


def test_trajectory_loss_autograd_entry(torch):
    """loss.backward() through the autograd entry point hands over the adjoint's gradient."""
    from electricbubbles_amd import adjoint
    P, centres, modes = _small_problem(torch, nmodes=4)
    ref = adjoint.gradient(P, centres, modes, 1.0, 0.1, 3, 1e-3, checkpoint=2)
    E = torch.tensor(1.0, dtype=torch.float64, device="cuda", requires_grad=True)
    Ca = torch.tensor(0.1, dtype=torch.float64, device="cuda", requires_grad=True)
    m = modes.clone().requires_grad_()
    loss = adjoint.trajectory_loss(E, Ca, m, P, centres, 3, 1e-3, checkpoint=2)
    (2.0 * loss).backward()
    assert abs(float(loss) - ref[0]) < 1e-15 * abs(ref[0]) + 1e-16
    assert abs(float(E.grad) - 2 * ref[1]) < 1e-13 * abs(ref[1])
    assert abs(float(Ca.grad) - 2 * ref[2]) < 1e-13 * abs(ref[2])
    assert rel(m.grad.cpu().numpy(), 2 * ref[3].cpu().numpy()) < 1e-13


# ------------------------------------------------------------------------------------------
# The adjoint configuration at its size: 16 drops x 512 points (4 x 4 grid, spacing 4), 64
# shape modes per drop of amplitude U[0, 0.02] weighted 1/m^2, |E| = 1, Ca = 0.1, RK4 at
# dt = 1e-3, GMRES 1e-12.
# ------------------------------------------------------------------------------------------

def _config16(torch):
    from electricbubbles_amd import adjoint
    centres, modes = adjoint.grid_problem(4, 512, 4.0, 64, 0.02, seed=5, device="cuda")
    return adjoint.Problem(16, 512, tol=1e-12), centres, modes


def test_gradient_at_size_matches_gpu_central_differences(torch):
    """20 steps: the gradient along two random unit directions of all 1026 parameters against
    Richardson-extrapolated central differences (h = 1e-5 and 5e-6) of the same forward."""
    from electricbubbles_amd import adjoint
    P, centres, modes = _config16(torch)
    steps, dt, E, Ca, decay = 20, 1e-3, 1.0, 0.1, 2.0
    print("adjoint 16 x 512, 20 steps", flush=True)
    loss, gE, gCa, gm, _ = adjoint.gradient(P, centres, modes, E, Ca, steps, dt, checkpoint=5, decay=decay)
    rng = np.random.default_rng(20)
    h = 1e-5
    for _ in range(2):
        v = rng.standard_normal(2 + modes.numel())
        v /= np.linalg.norm(v)
        vm = torch.tensor(v[2:].reshape(modes.shape), device="cuda")

        def L(e):
            return adjoint.forward_loss(P, centres, modes + e * vm, E + e * v[0], Ca + e * v[1], steps, dt,
                                        decay=decay)
        fd = (4 * (L(h / 2) - L(-h / 2)) / h - (L(h) - L(-h)) / (2 * h)) / 3
        an = gE * v[0] + gCa * v[1] + float((gm * vm).sum())
        print("directional derivative: adjoint", an, "central differences", fd, flush=True)
        assert abs(an - fd) < 1e-6 * abs(fd)


def test_gradient_at_size_matches_reference_differences(torch):
    """10 steps: dL/d|E|, dL/dCa and two directional derivatives against central differences of
    the same loss whose F is the reference's own (tests/golden/adjoint_fd_16drops.npz)."""
    from electricbubbles_amd import adjoint
    f = np.load(os.path.join(GOLDEN, "adjoint_fd_16drops.npz"))
    P = adjoint.Problem(16, int(f["N"]), tol=1e-12)
    centres = torch.tensor(f["centres"], device="cuda")
    modes = torch.tensor(f["modes"], device="cuda")
    loss, gE, gCa, gm, _ = adjoint.gradient(P, centres, modes, float(f["E"]), float(f["Ca"]), int(f["steps"]),
                                            float(f["dt"]), checkpoint=5, decay=float(f["decay"]))
    print("loss", loss, float(f["loss"]), "dE", gE, float(f["fd_E"]), "dCa", gCa, float(f["fd_Ca"]), flush=True)
    assert abs(loss - float(f["loss"])) < 1e-9 * abs(float(f["loss"]))
    assert abs(gE - float(f["fd_E"])) < 1e-6 * abs(float(f["fd_E"]))
    assert abs(gCa - float(f["fd_Ca"])) < 1e-6 * abs(float(f["fd_Ca"]))
    gmn = gm.cpu().numpy().ravel()
    for v, fd in zip(f["dirs"], f["fd_dirs"]):
        an = gE * v[0] + gCa * v[1] + float((gmn * v[2:]).sum())
        assert abs(an - fd) < 1e-6 * abs(fd), (an, fd)


def test_checkpoint_interval_does_not_change_the_200_step_gradient(torch):
    """The 200-step run: checkpoints every 20 steps and every 25 give the same gradient."""
    from electricbubbles_amd import adjoint
    P, centres, modes = _config16(torch)
    print("adjoint 16 x 512, 200 steps, checkpoints 20 and 25", flush=True)
    a = adjoint.gradient(P, centres, modes, 1.0, 0.1, 200, 1e-3, checkpoint=20, decay=2.0)
    b = adjoint.gradient(P, centres, modes, 1.0, 0.1, 200, 1e-3, checkpoint=25, decay=2.0)
    assert abs(a[0] - b[0]) < 1e-13 * abs(a[0])
    assert abs(a[1] - b[1]) < 1e-9 * abs(a[1])
    assert abs(a[2] - b[2]) < 1e-9 * abs(a[2])
    assert rel(a[3].cpu().numpy(), b[3].cpu().numpy()) < 1e-9
