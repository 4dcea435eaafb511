"""GPU parity of the Stokes-matvec FMM (CADAPDR family) through the C-ABI.

Checkers: the reference's own outputs (tests/golden/*.npz, written by the compiled reference),
the C restatement of the reference algorithm (oracle.cadap, oracle.assign1, ...) and the
long-double O(N^2) closed-form sum (oracle.direct).

Bars: tree tables and lists bit-exact; GW within 1e-10 relative (2-norm over all points) of
the reference FMM and of the direct sum.
"""
import os

import numpy as np
import pytest

import electricbubbles_amd as eb
import oracle
from electricbubbles_amd.workloads import complex_normal, curves_cfg, perturbed_circle, uniform_cfg, weights
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL_REL = 1e-10


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no GPU")
    return _t


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


# ------------------------------------------------------------------------------------------
# tree and lists: bit-exact
# ------------------------------------------------------------------------------------------

def test_tables_match_reference_golden(torch):
    g = golden("assign1_curves2000")
    x = torch.tensor(g["x"], device="cuda")
    y = torch.tensor(g["y"], device="cuda")
    for lev in range(1, 8):
        print(f"tables golden level {lev}", flush=True)
        t = eb.tables(lev, x, y)
        for k in ("ibox", "ioffst", "iatadr", "iloc"):
            np.testing.assert_array_equal(t[k], g[f"{k}_{lev}"], err_msg=f"{k} level {lev}")
        np.testing.assert_array_equal(t["icnt"], np.diff(g[f"ioffst_{lev}"]))
        assert t["ntbox"] == int(g[f"ntbox_{lev}"])
        assert t["nsmax"] == int(g[f"nsmax_{lev}"])


@pytest.mark.parametrize("kind", ["curves", "uniform", "grid_edges"])
def test_tables_match_oracle(torch, kind):
    rng = np.random.default_rng(7)
    if kind == "curves":
        z, _, _, _ = curves_cfg(60, 100, rng)
    elif kind == "uniform":
        z, _, _, _ = uniform_cfg(20000, rng)
    else:
        # points exactly on box edges of every level, and the top/right boundary (X = 64)
        k = rng.integers(0, 257, size=(3000, 2))
        z = k[:, 0] / 256.0 + 1j * k[:, 1] / 256.0
        z[:4] = [0.0, 1.0, 1j, 1 + 1j]
    x = torch.tensor(z.real.copy(), device="cuda")
    y = torch.tensor(z.imag.copy(), device="cuda")
    zs, _, _ = oracle.scale(z.real, z.imag)
    for lev in range(1, 9):
        print(f"tables {kind} level {lev}", flush=True)
        t = eb.tables(lev, x, y)
        o = oracle.assign1(lev, zs)
        for key in ("ibox", "icnt", "ioffst", "iatadr", "iloc"):
            np.testing.assert_array_equal(t[key], o[key], err_msg=f"{key} level {lev}")
        assert t["ntbox"] == o["ntbox"] and t["nsmax"] == o["nsmax"]


def test_lists_match_reference(torch):
    g = golden("lists_l1to5")
    for lev in range(1, 8):
        intl, nbor = eb.lists(lev)
        for b in range(4 ** lev):
            if lev <= 5:
                gi = g[f"intl_{lev}"][b]
                gn = g[f"nbor_{lev}"][b]
                gi, gn = gi[gi > 0], gn[gn > 0]
            else:
                gi, gn = oracle.mkintl(lev, b + 1), oracle.mknbor(lev, b + 1)
            np.testing.assert_array_equal(intl[b], gi)
            np.testing.assert_array_equal(nbor[b], gn)


# ------------------------------------------------------------------------------------------
# drop-in interface vs the reference's GRADW
# ------------------------------------------------------------------------------------------

GOLDEN_FMM = ["cadap_drop256", "cadap_curves1024", "cadap_curves4096", "cadap_uniform4096"]
VARIANTS = [("dr", eb.cadapdr), ("mod", eb.cadapmod), ("plus", eb.cadap)]


@pytest.mark.parametrize("name", GOLDEN_FMM)
@pytest.mark.parametrize("mode", [eb.MODE_FMM, eb.MODE_DIRECT, eb.MODE_AUTO])
def test_dropin_matches_reference(torch, name, mode):
    g = golden(name)
    eb.set_mode(mode)
    try:
        for tag, fn in VARIANTS:
            print(f"dropin {name} mode {mode} {tag}", flush=True)
            gw, inform, ier = fn(int(g["levmax"]), g["x"], g["y"], g["dzdt"], g["nd"], g["dens"], g["h"],
                                 float(g["tol"]))
            assert ier[0] == 0, ier
            np.testing.assert_array_equal(inform, g["inform"])
            e = rel(gw, g["gw_" + tag])
            assert e < TOL_REL, (name, mode, tag, e)
    finally:
        eb.set_mode(eb.MODE_AUTO)


def test_dropin_repeated_calls_reuse_geometry(torch):
    """Same geometry, new densities (one GMRES solve): every call matches the oracle."""
    g = golden("cadap_curves4096")
    rng = np.random.default_rng(3)
    eb.set_mode(eb.MODE_FMM)
    try:
        for it in range(3):
            dens = complex_normal(g["x"].size, rng)
            gw, _, ier = eb.cadapdr(int(g["levmax"]), g["x"], g["y"], g["dzdt"], g["nd"], dens, g["h"], 1e-13)
            ref, _, _, _ = oracle.cadap(oracle.DR, int(g["levmax"]), g["nd"], g["x"], g["y"], g["dzdt"], dens,
                                        g["h"], 1e-13)
            assert ier[0] == 0 and rel(gw, ref) < TOL_REL, it
    finally:
        eb.set_mode(eb.MODE_AUTO)


def test_dropin_error_codes_match_reference(torch):
    g = golden("cadap_errors")
    for tag in ("tol", "lenw_p", "leniw", "lenw_c"):
        lev, lenw, leniw = (int(v) for v in g[f"args_{tag}"])
        _, inform, ier = eb.cadapdr(lev, g["x"], g["y"], g["dzdt"], g["nd"], g["dens"], g["h"],
                                    float(g[f"tol_{tag}"]), lenw=lenw, leniw=leniw)
        np.testing.assert_array_equal(ier, g[f"ier_{tag}"], err_msg=tag)
        assert inform[1] == g[f"inform_{tag}"][1]
        if tag != "tol":
            assert inform[0] == g[f"inform_{tag}"][0]


def test_dropin_rejects_unbuilt_branches(torch):
    g = golden("cadap_drop256")
    for f1, f2 in ((1, 0), (0, 1)):
        with pytest.raises(eb.EBError) as ei:
            eb.cadapdr(7, g["x"], g["y"], g["dzdt"], g["nd"], g["dens"], g["h"], 1e-13, iflag1=f1, iflag2=f2)
        assert ei.value.code == eb.EB_ERR_UNSUPPORTED


# ------------------------------------------------------------------------------------------
# device plan vs the direct sum
# ------------------------------------------------------------------------------------------

def _plan_case(torch, z, dz, nd, h, dens, levmax, p, variant=eb.VARIANT_DR, mode=eb.MODE_FMM):
    x = torch.tensor(z.real.copy(), device="cuda")
    y = torch.tensor(z.imag.copy(), device="cuda")
    w = torch.tensor(weights(dz, nd, h), device="cuda")
    plan = eb.FMMPlan(x, y, w, levmax, p, variant, mode)
    out = plan.apply(torch.tensor(dens, device="cuda")).cpu().numpy()
    return plan, out


@pytest.mark.parametrize("kind,levmax,p", [("curves", 5, 43), ("curves", 6, 43), ("uniform", 6, 43),
                                           ("uniform", 7, 43), ("curves", 6, 33), ("uniform", 4, 30)])
@pytest.mark.parametrize("variant", [eb.VARIANT_DR, eb.VARIANT_MOD, eb.VARIANT_PLUS])
def test_plan_matches_direct_sum(torch, kind, levmax, p, variant):
    rng = np.random.default_rng(11 + levmax + p)
    if kind == "curves":
        z, dz, nd, h = curves_cfg(64, 256, rng)
    else:
        z, dz, nd, h = uniform_cfg(16384, rng)
    dens = complex_normal(z.size, rng)
    print(f"plan {kind} L={levmax} p={p} variant={variant}", flush=True)
    plan, out = _plan_case(torch, z, dz, nd, h, dens, levmax, p, variant)
    tgt = np.sort(rng.choice(z.size, 2048, replace=False)).astype(np.int32)
    ref = oracle.direct(variant, nd, z.real, z.imag, dz, dens, h, tgt)
    e = rel(out[tgt], ref)
    bound = TOL_REL if p >= 33 else 1e-8
    assert e < bound, (kind, levmax, p, variant, e)


def test_expansion_length_sweep_converges(torch):
    """Every kernel variant the expansion length selects (K = 16, 32, 48, 64; the M2L with a
    vector tail at p = 33..36, three MFMA tiles at 37..48; the wide P2M from K = 48) against
    the long-double direct sum.  A wrong row or coefficient in one variant leaves its error
    stuck at that row's size, so besides a bound per p the error must keep falling by the
    series' rate from one p to the next (about 2.2 per term on these curves)."""
    rng = np.random.default_rng(77)
    z, dz, nd, h = curves_cfg(64, 256, rng)
    dens = complex_normal(z.size, rng)
    tgt = np.sort(rng.choice(z.size, 2048, replace=False)).astype(np.int32)
    ref = oracle.direct(oracle.DR, nd, z.real, z.imag, dz, dens, h, tgt)
    errs = {}
    for p in (16, 17, 24, 30, 31, 32, 33, 34, 35, 36, 37, 38, 40, 41, 44, 48, 49, 52, 60):
        print(f"p sweep: p={p}", flush=True)
        plan, out = _plan_case(torch, z, dz, nd, h, dens, 6, p)
        plan.close()
        errs[p] = rel(out[tgt], ref)
    print({p: f"{e:.2e}" for p, e in errs.items()}, flush=True)
    for p, e in errs.items():
        assert e < max(5e-13, 1e-3 * 2.19 ** -(p - 16)), (p, e)
    for p0, p1 in ((30, 32), (31, 33), (32, 34), (33, 35), (34, 36), (35, 37), (36, 38)):
        assert errs[p1] < 0.6 * errs[p0], (p0, p1, errs[p0], errs[p1])


@pytest.mark.parametrize("n", [2, 3, 63, 64, 65, 1000, 4097, 10240])
def test_direct_mode_chunk_boundaries(torch, n):
    """The direct sum runs as target tiles of 64 x source chunks with a fixed-order chunk
    sum; sizes at and around the tile and chunk boundaries against the long-double sum."""
    rng = np.random.default_rng(n)
    z = rng.uniform(size=n) + 1j * rng.uniform(size=n)
    dzv = np.exp(1j * rng.uniform(0, 6.28, n))
    nd = np.array([n], np.int32)
    h = np.array([0.01])
    dens = complex_normal(n, rng)
    tgt = np.arange(n, dtype=np.int32) if n <= 4097 else np.sort(rng.choice(n, 1024, replace=False)).astype(np.int32)
    ref = oracle.direct(oracle.DR, nd, z.real, z.imag, dzv, dens, h, tgt)
    _, out = _plan_case(torch, z, dzv, nd, h, dens, 4, 43, mode=eb.MODE_DIRECT)
    assert rel(out[tgt], ref) < 1e-12, n


def test_plan_is_real_linear(torch):
    rng = np.random.default_rng(5)
    z, dz, nd, h = curves_cfg(32, 128, rng)
    x = torch.tensor(z.real.copy(), device="cuda")
    y = torch.tensor(z.imag.copy(), device="cuda")
    w = torch.tensor(weights(dz, nd, h), device="cuda")
    plan = eb.FMMPlan(x, y, w, 6, 40)
    d1 = torch.tensor(complex_normal(z.size, rng), device="cuda")
    d2 = torch.tensor(complex_normal(z.size, rng), device="cuda")
    a = 0.37
    lhs = plan.apply(a * d1 + d2).cpu().numpy()
    rhs = a * plan.apply(d1).cpu().numpy() + plan.apply(d2).cpu().numpy()
    assert rel(lhs, rhs) < 1e-13


def test_plan_edge_sizes(torch):
    rng = np.random.default_rng(9)
    # two points; all points on one vertical line (RSX = 0); a single dense leaf
    cases = [np.array([0.1 + 0.2j, 0.7 + 0.9j]),
             0.5 + 1j * np.linspace(0, 1, 300),
             0.5 + 0.5j + 1e-3 * (rng.uniform(size=500) + 1j * rng.uniform(size=500))]
    cases[2][0] = 0.0
    cases[2][1] = 1.0 + 1.0j
    for z in cases:
        dz = np.exp(1j * rng.uniform(0, 6.28, z.size))
        nd = np.array([z.size], np.int32)
        h = np.array([0.01])
        dens = complex_normal(z.size, rng)
        ref = oracle.direct(oracle.DR, nd, z.real, z.imag, dz, dens, h)
        for lev in (1, 3, 6):
            _, out = _plan_case(torch, z, dz, nd, h, dens, lev, 43)
            assert rel(out, ref) < TOL_REL, (z.size, lev)


def test_plan_ntbox_matches_reference_counts(torch):
    g = golden("assign1_curves2000")
    z = g["x"] + 1j * g["y"]
    dz = np.ones(z.size, complex)
    x = torch.tensor(g["x"], device="cuda")
    y = torch.tensor(g["y"], device="cuda")
    w = torch.tensor(dz, device="cuda")
    plan = eb.FMMPlan(x, y, w, 7, 43)
    nt = plan.ntbox()
    for lev in range(1, 8):
        assert nt[lev] == int(g[f"ntbox_{lev}"])
    info = plan.info()
    assert info["rscal"] == float(g["rscal"])


def test_large_curves_against_direct_subset(torch):
    """Config-3a shape at 2^18 points: 1024 curves x 256 points, checked on 256 targets."""
    rng = np.random.default_rng(2025)
    z, dz, nd, h = curves_cfg(1024, 256, rng)
    dens = complex_normal(z.size, rng)
    print("large curves 2^18", flush=True)
    plan, out = _plan_case(torch, z, dz, nd, h, dens, 7, 36)
    tgt = np.sort(rng.choice(z.size, 256, replace=False)).astype(np.int32)
    ref = oracle.direct(oracle.DR, nd, z.real, z.imag, dz, dens, h, tgt)
    assert rel(out[tgt], ref) < TOL_REL


# ------------------------------------------------------------------------------------------
# The Hilbert-matrix product (EB_VARIANT_HILBERT) on the FMM: the product DAPIF2 forms with
# IFLAG7 = 3 for the Laplace side (laplace.f:204-215, pv_integral.f:171-182), against the
# long-double sum of oracle.cauchy (itself pinned to the reference DAPIF2's output and to a
# brute-force numpy sum in tests/test_oracle_cpu.py).
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("kind,n,levmax,p", [("curves", 16384, 7, 37), ("curves", 65536, 7, 37),
                                             ("uniform", 16384, 6, 37), ("curves", 65536, 8, 33)])
def test_hilbert_fmm_against_cauchy_sum(torch, kind, n, levmax, p):
    rng = np.random.default_rng(n + levmax + p)
    if kind == "curves":
        z, _, _, _ = curves_cfg(n // 128, 128, rng)
    else:
        z, _, _, _ = uniform_cfg(n, rng)
    qa = complex_normal(n, rng)
    print("hilbert fmm", kind, n, levmax, p, flush=True)
    x = torch.tensor(z.real.copy(), device="cuda")
    y = torch.tensor(z.imag.copy(), device="cuda")
    ones = torch.ones(n, dtype=torch.complex128, device="cuda")
    plan = eb.FMMPlan(x, y, ones, levmax, p, eb.VARIANT_HILBERT, eb.MODE_FMM)
    out = plan.apply(torch.tensor(qa, device="cuda")).cpu().numpy()
    tgt = np.sort(rng.choice(n, 2048, replace=False)).astype(np.int32)
    ref = oracle.cauchy(z.real, z.imag, qa, tgt)
    # p = 37 is the term count of TOL 1e-11 (laplace.f:212); p = 33 that of 1e-10 (pv_integral.f:175)
    assert rel(out[tgt], ref) < (1e-10 if p >= 37 else 1e-9)
    # the direct mode of the same plan type at the same targets
    dplan = eb.FMMPlan(x, y, ones, levmax, p, eb.VARIANT_HILBERT, eb.MODE_DIRECT) if n <= 16384 else None
    if dplan is not None:
        dout = dplan.apply(torch.tensor(qa, device="cuda")).cpu().numpy()
        assert rel(dout[tgt], ref) < 1e-13


# ------------------------------------------------------------------------------------------
# The 2^20-point workloads of the headline (circles and uniform points, the bench's seeds) at
# the tuned depth and expansion length, against the long-double O(N^2) sum over all 2^20
# sources at 1024 random targets.
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("kind", ["curves", "uniform"])
def test_two_to_the_20_against_direct_sum(torch, kind):
    rng = np.random.default_rng(20260303)
    n = 1 << 20
    if kind == "curves":
        z, dz, nd, h = curves_cfg(n // 256, 256, rng)
    else:
        z, dz, nd, h = uniform_cfg(n, rng)
    dens = complex_normal(n, rng)
    print("2^20", kind, "LEVMAX 8, p 36", flush=True)
    plan, out = _plan_case(torch, z, dz, nd, h, dens, 8, 36)
    info = plan.info()
    assert info["levmax"] == 8 and info["nterms"] == 36
    tgt = np.sort(np.random.default_rng(7).choice(n, 1024, replace=False)).astype(np.int32)
    ref = oracle.direct(oracle.DR, nd, z.real, z.imag, dz, dens, h, tgt)
    err = rel(out[tgt], ref)
    print(f"2^20 {kind}: rel err {err:.2e} at 1024 targets, max leaf {info['max_leaf']}", flush=True)
    assert err < TOL_REL
    # the same targets from a plan re-binned in place (the bench's timed loop)
    plan.rebuild(plan._keep[0], plan._keep[1], plan._keep[2])
    out2 = plan.apply(torch.tensor(dens, device="cuda")).cpu().numpy()
    assert rel(out2[tgt], ref) < TOL_REL
    plan.close()


# ------------------------------------------------------------------------------------------
# The tree-free O(N^2) checker at a subset of targets (eb_direct_subset), used by the
# autotuner and the benchmark to measure accuracy: against the long-double sums.
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("variant", [eb.VARIANT_DR, eb.VARIANT_MOD, eb.VARIANT_PLUS, eb.VARIANT_HILBERT])
def test_direct_subset_against_long_double_sums(torch, variant):
    rng = np.random.default_rng(40 + variant)
    z, dz, nd, h = curves_cfg(40, 256, rng)
    z = np.concatenate([z, z[:3] + 1e-7])  # near-coincident points
    dz = np.concatenate([dz, dz[:3]])
    nd = np.append(nd, 3).astype(np.int32)
    h = np.append(h, h[0])
    dens = complex_normal(z.size, rng)
    tgt = np.sort(rng.choice(z.size, 300, replace=False)).astype(np.int32)
    tgt[-1] = z.size - 1
    x = torch.tensor(z.real.copy(), device="cuda")
    y = torch.tensor(z.imag.copy(), device="cuda")
    w = torch.tensor(weights(dz, nd, h), device="cuda")
    got = eb.direct_subset(x, y, w, torch.tensor(dens, device="cuda"), torch.tensor(tgt, device="cuda"),
                           variant).cpu().numpy()
    if variant == eb.VARIANT_HILBERT:
        ref = oracle.cauchy(z.real, z.imag, dens * weights(dz, nd, h), tgt)
    else:
        ref = oracle.direct(variant, nd, z.real, z.imag, dz, dens, h, tgt)
    assert rel(got, ref) < 1e-13


# ------------------------------------------------------------------------------------------
# The DAPIF2 drop-in (IFLAG7 = 3): FIELD, IERR and INFORM against the reference DAPIF2
# (tests/golden/dapif2_hilbert.npz), in the O(N^2) and the FMM modes.
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("mode", [eb.MODE_AUTO, eb.MODE_FMM])
@pytest.mark.parametrize("tag", ["lap", "pv", "uniform"])
def test_dapif2_dropin_matches_reference(torch, tag, mode):
    g = golden("dapif2_hilbert")
    eb.set_mode(mode)
    try:
        print("dapif2 drop-in", tag, mode, flush=True)
        field, ierr, inform = eb.dapif2(g[f"x_{tag}"], g[f"y_{tag}"], g[f"qa_{tag}"], napb=30,
                                        tol=float(g[f"tol_{tag}"]), nsp=int(g[f"ierr_{tag}"][2]))
    finally:
        eb.set_mode(eb.MODE_AUTO)
    np.testing.assert_array_equal(ierr, g[f"ierr_{tag}"])
    np.testing.assert_array_equal(inform, g[f"inform_{tag}"])  # NTERMS, NBOX, levels, workspace used
    assert rel(field, g[f"field_{tag}"]) < TOL_REL


@pytest.mark.parametrize("tag", ["tol", "few", "nsp", "nsp_last", "nsp_half", "nsp_tables"])
def test_dapif2_dropin_error_codes_match_reference(torch, tag):
    g = golden("dapif2_hilbert")
    napb, nsp = (int(v) for v in g[f"err_{tag}_args"])
    _, ierr, inform = eb.dapif2(g["x_lap"], g["y_lap"], g["qa_lap"], napb=napb, tol=float(g[f"err_{tag}_tol"]),
                                nsp=nsp if nsp > 0 else None)
    np.testing.assert_array_equal(ierr, g[f"err_{tag}_ierr"], err_msg=tag)
    np.testing.assert_array_equal(inform, g[f"err_{tag}_inform"], err_msg=tag)
