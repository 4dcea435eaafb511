"""CPU tests: pin the oracle (C restatement) to the reference's outputs and to the mathematics.

The golden fixtures were written by the compiled reference (tests/golden/make_golden.py).
Independent pins: the O(N^2) closed form (long double), numpy's FFT, trigonometric
polynomials for the spectral derivative, ellipse closed forms for the geometry, brute-force
integer rules for the tree and lists.  Plausible mistakes (a dropped conj, a wrong sign, a
transposed index) are planted in the checks below to show each pin bites.
"""
import os

import numpy as np
import pytest

import oracle
from electricbubbles_amd.workloads import complex_normal, curves_cfg, perturbed_circle, uniform_cfg
from tests.conftest import GOLDEN


def g(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


# ------------------------------------------------------------------------------------------
# tree and lists
# ------------------------------------------------------------------------------------------

def test_scale_and_assign1_match_reference():
    t = g("assign1_curves2000")
    zs, rs, rmin = oracle.scale(t["x"], t["y"])
    np.testing.assert_array_equal(zs, t["zs"])
    assert rs == float(t["rscal"])
    np.testing.assert_array_equal(rmin, t["rmin"])
    for lev in range(1, 8):
        o = oracle.assign1(lev, zs)
        for k in ("ibox", "ioffst", "iatadr", "iloc"):
            np.testing.assert_array_equal(o[k], t[f"{k}_{lev}"], err_msg=f"{k} {lev}")
        np.testing.assert_array_equal(o["icnt"], np.diff(t[f"ioffst_{lev}"]))
        assert o["ntbox"] == int(t[f"ntbox_{lev}"]) and o["nsmax"] == int(t[f"nsmax_{lev}"])


def _assign1_bruteforce(lev, zs):
    nside = 2 ** lev
    h = 64.0 / nside
    ix = np.clip((zs.real / h).astype(int), 0, nside - 1)
    iy = np.clip((zs.imag / h).astype(int), 0, nside - 1)
    ibox = iy * nside + ix + 1
    order = np.lexsort((np.arange(zs.size), ibox))  # stable by box, then source index
    cnt = np.bincount(ibox - 1, minlength=nside * nside)
    first_seen = {}
    for j, b in enumerate(ibox):
        first_seen.setdefault(b, len(first_seen) + 1)
    iloc = np.zeros(nside * nside, int)
    for b, r in first_seen.items():
        iloc[b - 1] = r
    return ibox, order + 1, np.concatenate([[0], np.cumsum(cnt)]), iloc


def test_assign1_bruteforce_rules():
    rng = np.random.default_rng(0)
    z, _, _, _ = curves_cfg(10, 50, rng)
    zs, _, _ = oracle.scale(z.real, z.imag)
    for lev in range(1, 7):
        o = oracle.assign1(lev, zs)
        ibox, iat, ioff, iloc = _assign1_bruteforce(lev, zs)
        np.testing.assert_array_equal(o["ibox"], ibox)
        np.testing.assert_array_equal(o["iatadr"], iat)
        np.testing.assert_array_equal(o["ioffst"], ioff)
        np.testing.assert_array_equal(o["iloc"], iloc)


def _intl_bruteforce(lev, b):
    """Children of the parent's 3x3 neighbourhood that are not adjacent to the box."""
    nside = 2 ** lev
    iy, ix = divmod(b - 1, nside)
    out = []
    for jx in range(nside):          # column-outer, row-inner
        for jy in range(nside):
            if abs(jx // 2 - ix // 2) <= 1 and abs(jy // 2 - iy // 2) <= 1 and not (
                    abs(jx - ix) <= 1 and abs(jy - iy) <= 1):
                out.append(jy * nside + jx + 1)
    return np.array(out, np.int32)


def test_lists_match_reference_and_rules():
    t = g("lists_l1to5")
    for lev in range(1, 6):
        for b in range(1, 4 ** lev + 1):
            gi = t[f"intl_{lev}"][b - 1]
            gn = t[f"nbor_{lev}"][b - 1]
            np.testing.assert_array_equal(oracle.mkintl(lev, b), gi[gi > 0])
            np.testing.assert_array_equal(oracle.mknbor(lev, b), gn[gn > 0])
            np.testing.assert_array_equal(oracle.mkintl(lev, b), _intl_bruteforce(lev, b))


# ------------------------------------------------------------------------------------------
# the FMM restatement
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["cadap_drop256", "cadap_curves1024", "cadap_curves4096", "cadap_uniform4096"])
def test_cadap_restatement_matches_reference(name):
    t = g(name)
    for var, tag in ((oracle.DR, "dr"), (oracle.MOD, "mod"), (oracle.PLUS, "plus")):
        gw, inform, ier, _ = oracle.cadap(var, int(t["levmax"]), t["nd"], t["x"], t["y"], t["dzdt"], t["dens"],
                                          t["h"], float(t["tol"]))
        assert ier[0] == 0
        np.testing.assert_array_equal(inform, t["inform"])
        assert rel(gw, t["gw_" + tag]) < 1e-11, (name, tag)


@pytest.mark.parametrize("name", ["cadap_drop256", "cadap_curves1024", "cadap_uniform4096"])
def test_reference_fmm_matches_direct_sum(name):
    """The closed form (Appendix-C kernel, long double) against the reference's own output."""
    t = g(name)
    for var, tag in ((oracle.DR, "dr"), (oracle.MOD, "mod"), (oracle.PLUS, "plus")):
        d = oracle.direct(var, t["nd"], t["x"], t["y"], t["dzdt"], t["dens"], t["h"])
        assert rel(d, t["gw_" + tag]) < 1e-10, (name, tag)


def test_direct_sum_pins_signs():
    """A flipped variant sign or a dropped conj term must be visible against the reference."""
    t = g("cadap_curves1024")
    dr = t["gw_dr"]
    mod = oracle.direct(oracle.MOD, t["nd"], t["x"], t["y"], t["dzdt"], t["dens"], t["h"])
    plus = oracle.direct(oracle.PLUS, t["nd"], t["x"], t["y"], t["dzdt"], t["dens"], t["h"])
    assert rel(mod, dr) > 1e-2 and rel(plus, t["gw_mod"]) > 1e-2
    # conj(dens) instead of dens is a plausible slip; it changes the answer
    bad = oracle.direct(oracle.DR, t["nd"], t["x"], t["y"], t["dzdt"], np.conj(t["dens"]), t["h"])
    assert rel(bad, dr) > 1e-2


def test_direct_sum_two_points_closed_form():
    """Two sources: GW_1 written out by hand from the kernel definition."""
    z = np.array([0.2 + 0.1j, 0.9 - 0.4j])
    dz = np.array([1.0 + 0.5j, -0.3 + 0.8j])
    dens = np.array([0.7 - 0.2j, -0.4 + 1.1j])
    nd = np.array([2], np.int32)
    h = np.array([0.3])
    z2pi = 1.0 / (2 * np.pi * 1j)
    qa = z2pi * dens * dz * h[0]
    qb = (-np.conj(dens) * dz + dens * np.conj(dz)) * z2pi * h[0]
    qc = -np.conj(z) * qa
    d = z[1] - z[0]
    phi = qa[1] / d
    phip = qa[1] / d ** 2
    psi = qb[1] / d + qc[1] / d ** 2
    expect0 = phi - z[0] * np.conj(phip) - np.conj(psi)
    got = oracle.direct(oracle.DR, nd, z.real, z.imag, dz, dens, h)
    assert abs(got[0] - expect0) < 1e-14 * abs(expect0)


@pytest.mark.parametrize("tag", ["lap", "pv", "uniform"])
def test_cauchy_sum_matches_reference_dapif2_and_numpy(tag):
    """oracle.cauchy (DAPIF2's IFLAG7 = 3 product as a long-double sum) against the reference
    DAPIF2's FIELD at its TOL and against a brute-force numpy sum; a flipped sign (z_j - z_i)
    and a dropped self-exclusion both fail the numpy pin."""
    f = g("dapif2_hilbert")
    x, y, qa = f[f"x_{tag}"], f[f"y_{tag}"], f[f"qa_{tag}"]
    got = oracle.cauchy(x, y, qa)
    assert rel(got, f[f"field_{tag}"]) < 10 * float(f[f"tol_{tag}"])
    z = x + 1j * y
    d = z[:, None] - z[None, :]
    np.fill_diagonal(d, np.inf)
    ref = (qa[None, :] / d).sum(1)
    assert rel(got, ref) < 1e-14
    assert rel(-got, ref) > 1.0                      # the sign of z_i - z_j
    dd = d.copy()
    np.fill_diagonal(dd, 1.0)                        # a self term 1 * qa_i left in
    assert rel(got, (qa[None, :] / dd).sum(1)) > 1e-6
    # a subset of targets equals the same rows of the whole product
    tgt = np.array([0, 5, x.size - 1], np.int32)
    np.testing.assert_array_equal(oracle.cauchy(x, y, qa, tgt), got[tgt])


def test_direct_sum_is_independent_of_thread_count(monkeypatch):
    """The threaded long-double sums give the same bits on one and on several threads."""
    rng = np.random.default_rng(4)
    z, dz, nd, h = curves_cfg(16, 256, rng)
    dens = complex_normal(z.size, rng)
    outs = []
    for nth in ("1", "5"):
        monkeypatch.setenv("ORA_THREADS", nth)
        outs.append((oracle.direct(oracle.DR, nd, z.real, z.imag, dz, dens, h),
                     oracle.cauchy(z.real, z.imag, dens)))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def _real_matrix(apply, n):
    """The 2n x 2n real matrix of a real-linear map of C^n, on the coordinates
    (Re v_0..Re v_{n-1}, Im v_0..Im v_{n-1}): under them Re sum conj(a) b is the dot product,
    so the transpose of the map is this matrix's transpose."""
    cols = []
    for u in (1.0, 1j):
        for k in range(n):
            e = np.zeros(n, complex)
            e[k] = u
            v = apply(e)
            cols.append(np.concatenate([v.real, v.imag]))
    return np.array(cols).T


@pytest.mark.parametrize("variant", [oracle.DR, oracle.MOD, oracle.PLUS, 3])
def test_direct_transpose_is_the_dense_transpose(variant):
    """oracle.direct_transpose against the transposed real matrix of the forward sums
    (oracle.direct for the CADAP family, oracle.cauchy with QA = WDZ dens for the Hilbert
    product), built column by column on two small bodies; the untransposed matrix and a
    conj(lam) slip both fail."""
    rng = np.random.default_rng(11)
    z, dz, nd, h = curves_cfg(2, 24, rng)
    n = z.size
    hp = np.repeat(h, nd)
    wdz = dz * hp
    if variant == 3:
        fwd = lambda d: oracle.cauchy(z.real, z.imag, wdz * d)
    else:
        fwd = lambda d: oracle.direct(variant, nd, z.real, z.imag, dz, d, h)
    M = _real_matrix(fwd, n)
    lam = complex_normal(n, rng)
    lr = np.concatenate([lam.real, lam.imag])
    ref = M.T @ lr
    ref = ref[:n] + 1j * ref[n:]
    got = oracle.direct_transpose(variant, z.real, z.imag, wdz, lam)
    assert rel(got, ref) < 1e-13
    wrong = M @ lr
    assert rel(got, wrong[:n] + 1j * wrong[n:]) > 1e-2          # not the map itself
    assert rel(oracle.direct_transpose(variant, z.real, z.imag, wdz, np.conj(lam)), ref) > 1e-2
    tgt = np.array([0, 7, n - 1], np.int32)
    np.testing.assert_array_equal(oracle.direct_transpose(variant, z.real, z.imag, wdz, lam, tgt), got[tgt])


def test_cadap_error_codes_match_reference():
    t = g("cadap_errors")
    for tag in ("tol",):
        gw, inform, ier, _ = oracle.cadap(oracle.DR, 5, t["nd"], t["x"], t["y"], t["dzdt"], t["dens"], t["h"],
                                          float(t["tol_tol"]))
        np.testing.assert_array_equal(ier, t["ier_tol"])


# ------------------------------------------------------------------------------------------
# spectral, geometry, FASMVP epilogue, RK2
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [8, 12, 30, 64, 100, 256, 1234])
def test_fft_and_fdifff_match_reference_and_numpy(n):
    t = g("spectral")
    c = t[f"c_{n}"]
    np.testing.assert_allclose(oracle.dft(c, -1), t[f"fwd_{n}"], rtol=0, atol=1e-12 * np.abs(t[f"fwd_{n}"]).max())
    np.testing.assert_allclose(oracle.dft(c, -1), np.fft.fft(c), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(oracle.dft(c, +1), n * np.fft.ifft(c), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(oracle.dft(c, +1), t[f"bwd_{n}"], rtol=0, atol=1e-12 * np.abs(t[f"bwd_{n}"]).max())
    d = oracle.fdifff(c)
    assert rel(d, t[f"dif_{n}"]) < 1e-12


def test_fdifff_exact_on_trig_polynomials():
    for n in (16, 30, 64):
        a = 2 * np.pi * np.arange(n) / n
        f = np.exp(3j * a) + 0.5 * np.cos(5 * a) - 0.25j * np.sin(2 * a)
        df = 3j * np.exp(3j * a) - 2.5 * np.sin(5 * a) - 0.5j * np.cos(2 * a)
        assert np.abs(oracle.fdifff(f) - df).max() < 1e-12


@pytest.mark.parametrize("n", [64, 256])
def test_filter_and_equit_match_reference(n):
    t = g("spectral")
    np.testing.assert_allclose(oracle.filter_(t[f"filt_in_{n}"]), t[f"filt_out_{n}"], rtol=1e-14, atol=1e-300)
    te = oracle.equit(t[f"equit_tha_{n}"], t[f"equit_un_{n}"])
    assert rel(te, t[f"equit_t_{n}"]) < 1e-12


def test_extractdat_matches_reference():
    t = g("geometry_two_drops")
    o = oracle.extractdat(t["nd"], t["rku"])
    for k in ("x", "y", "zk", "dz", "tha", "dsda", "rkappa", "ckern"):
        assert rel(o[k], t[k]) < 1e-12, k
    assert o["xleft"] == float(t["xleft"]) and abs(o["slength"] - float(t["slength"])) < 1e-14


def test_extractdat_ellipse_closed_forms():
    n, a, b = 128, 1.5, 0.8
    th = 2 * np.pi * np.arange(n) / n
    z = a * np.cos(th) + 1j * b * np.sin(th) + (0.3 - 0.2j)
    o = oracle.extractdat(np.array([n], np.int32), z.view(np.float64).copy())
    assert abs(o["area"][0] - np.pi * a * b) < 1e-12
    kappa = a * b / (a * a * np.sin(th) ** 2 + b * b * np.cos(th) ** 2) ** 1.5
    # the reference's sign convention: rkappa = -(x'y'' - y'x'')/|z'|^3
    assert np.abs(o["rkappa"] + kappa).max() < 1e-10
    assert abs(o["zk"][0] - (0.3 - 0.2j)) < 1e-14


def test_fasmvp_epilogue_matches_reference():
    t = g("geometry_two_drops")
    w, za, pk = oracle.fasmvp_epilogue(t["nd"], t["x"], t["y"], t["dz"], t["dsda"], t["rkappa"], t["zk"], t["u"],
                                       t["stress"] * 0 + _stress_in(t))
    assert rel(w, t["w"]) < 1e-12
    assert rel(za, t["za"]) < 1e-13


def _stress_in(t):
    # the reference's STRESS on exit already holds GW + self + pressure terms; recompute GW with
    # the same CADAPDR call FASMVP makes (LEVMAX 7, TOL 1e-13) from the restatement
    nd = t["nd"]
    dens = t["u"].view(np.complex128)
    h = 2 * np.pi / nd
    gw, _, ier, _ = oracle.cadap(oracle.DR, 7, nd, t["x"], t["y"], t["dz"], dens, h, 1e-13)
    assert ier[0] == 0
    return gw


def test_ftail_matches_oracle_pieces():
    """F tail = EQUIT + pointwise frame + FFT/FILTER/iFFT; pieces pinned above."""
    rng = np.random.default_rng(4)
    t = g("geometry_two_drops")
    un = rng.normal(size=t["x"].size) * 1e-2
    rk = oracle.ftail(t["nd"], t["tha"], un, t["dz"], t["dsda"])
    # per body: a band-limited velocity must survive the filter almost unchanged when Un = 0
    rk0 = oracle.ftail(t["nd"], t["tha"], 0 * un, t["dz"], t["dsda"])
    assert np.abs(rk0).max() < 1e-12
    assert np.isfinite(rk).all()


def test_rk4_on_linear_ode():
    """One step of classic RK4 on y' = A y is the degree-4 Taylor polynomial of exp(dt A)
    applied to y; a dropped factor 2 on k2 or k3 fails it."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 6))
    y0 = rng.standard_normal(6)
    dt = 0.05
    hA = dt * A
    P = np.eye(6) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    got = oracle.rk4(lambda y: A @ y, y0, dt, 1)
    assert rel(got, P @ y0) < 1e-14
    y2 = oracle.rk4(lambda y: A @ y, y0, dt, 3)
    assert rel(y2, np.linalg.matrix_power(P, 3) @ y0) < 1e-14
    bad = y0 + dt / 6 * (A @ y0 + A @ (y0 + 0.5 * dt * A @ y0) + 2 * A @ (y0 + 0.5 * dt * A @ (y0 + 0.5 * dt * A @ y0)))
    assert rel(got, bad) > 1e-6


def test_deformation_of_ellipses_is_the_closed_form():
    """D of an ellipse with semi-axes a > b is (a - b) / (a + b) whatever its centre, rotation
    and parametrisation origin; two bodies at once; a circle gives 0."""
    th = 2 * np.pi * np.arange(256) / 256
    bodies, expect = [], []
    for a, b, rot, c, ph in ((1.3, 0.8, 0.4, 20 - 7j, 0.3), (1.0, 1.0, 0.0, 0.5, 0.0), (2.0, 0.5, -1.1, 3j, 1.0)):
        e = (a * np.cos(th + ph) + 1j * b * np.sin(th + ph)) * np.exp(1j * rot) + c
        bodies.append(e)
        expect.append((a - b) / (a + b))
    z = np.concatenate(bodies)
    D = oracle.deformation([256] * 3, z.view(np.float64))
    np.testing.assert_allclose(D, expect, atol=1e-12)


def test_rk2_heun_on_linear_ode():
    lam = -0.7
    y0 = np.array([1.0, 2.0])
    dt = 0.01
    phi0, y = oracle.rk2_stage1(dt, y0, lam * y0)
    y1 = oracle.rk2_stage2(dt, phi0, lam * y, y0)
    exact_heun = y0 * (1 + lam * dt + 0.5 * (lam * dt) ** 2)
    np.testing.assert_allclose(y1, exact_heun, rtol=1e-15)
