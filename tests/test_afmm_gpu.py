"""GPU checks of the FMM on the adaptive tree (eb_afmm_*, electricbubbles_amd.AFMMPlan): the
lists carry every pair of childless boxes exactly once (host check), the sums agree with the
O(N^2) sum of the same kernel for every variant on curves, uniform points and clustered points
whose tree is strongly graded (lists W and X in use), and at 2^20 points against the tree-free
direct sum at 1,024 seeded targets.  The adaptive tree is DAPIF2's (dapif1.f:1165-1680); the
sums are the layer potentials of CADAPDR / CADAPMOD / CADAP and DAPIF2's Hilbert product."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no GPU")
    return _t


def rel(a, b):
    return float((a - b).norm() / b.norm())


def _clusters(n, rng):
    """Gaussian clusters of widths 1e-4 .. 1e-1 (a tree ten and more levels deep beside
    nearly empty regions), weights of unit size."""
    k = 12
    cen = rng.uniform(0.0, 1.0, (k, 2))
    wid = np.exp(rng.uniform(np.log(1e-4), np.log(1e-1), k))
    lab = rng.integers(0, k, n)
    z = cen[lab, 0] + 1j * cen[lab, 1] + wid[lab] * (rng.normal(size=n) + 1j * rng.normal(size=n))
    return z, np.ones(n, np.complex128), np.array([n], np.int32), np.array([1.0 / n])


def _case(kind, n, seed):
    from electricbubbles_amd.workloads import curves_cfg, uniform_cfg, weights
    rng = np.random.default_rng(seed)
    if kind == "curves":
        z, dz, nd, h = curves_cfg(n // 256, 256, rng)
    elif kind == "uniform":
        z, dz, nd, h = uniform_cfg(n, rng)
    else:
        z, dz, nd, h = _clusters(n, rng)
    return z, weights(dz, nd, h), rng


def _tensors(torch, z, wdz):
    return (torch.tensor(z.real.copy(), device="cuda"), torch.tensor(z.imag.copy(), device="cuda"),
            torch.tensor(wdz, device="cuda"))


@pytest.mark.parametrize("kind", ["curves", "uniform", "clusters"])
def test_lists_carry_every_pair_once(torch, kind):
    import electricbubbles_amd as eb
    z, wdz, rng = _case(kind, 8192, 11)
    print("lists:", kind, z.size, "points", flush=True)
    x, y, w = _tensors(torch, z, wdz)
    P = eb.AFMMPlan(x, y, w, 20)
    info = P.info()
    print(info, flush=True)
    assert info["leaves"] > 100
    assert P.check_lists() == 0
    if kind == "clusters":
        assert info["m2p_evals"] > 0 and info["p2l_terms"] > 0
    P.close()


def _bound(p):
    """Truncation at one-box separation, 0.55^p (sqrt 2 / (4 - sqrt 2) = 0.547), with a factor
    20, above a rounding floor of 5e-13 (the QC charges -conj(z) QA carry the coordinates'
    size into the psi series)."""
    return max(20 * 0.55 ** p, 5e-13)


def _direct(eb, P, w, d, targets, variant=0):
    """The O(N^2) sum (tree-free, at the targets) on the plan's own scaled coordinates, times
    RSCAL: the plan's result up to its expansions.  Two scalings of the same points round
    differently, and where points lie 1e-6 apart (the clusters) that alone moves the sum by
    1e-9 relative."""
    xs, ys = P.scaled_points()
    return eb.direct_subset(xs, ys, w, d, targets, variant=variant) * P.info()["rscal"]


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", ["curves", "uniform", "clusters"])
def test_against_direct_sum(torch, kind, variant):
    """p = 36 (two MFMA tiles and a vector tail in the translations) at every point, for two
    densities on one plan: 1e-10 relative on curves and uniform points; on the clusters, whose
    graded tree puts many boxes at the worst separation, the truncation bound of
    test_expansion_lengths (9e-9 at p = 36; about 5e-10 is what they reach)."""
    import electricbubbles_amd as eb
    from electricbubbles_amd.workloads import complex_normal
    z, wdz, rng = _case(kind, 16384, 5 + variant)
    print("against direct:", kind, "variant", variant, z.size, "points", flush=True)
    x, y, w = _tensors(torch, z, wdz)
    P = eb.AFMMPlan(x, y, w, 36, variant=variant)
    every = torch.arange(z.size, dtype=torch.int32, device="cuda")
    for rep in range(2):
        d = torch.tensor(complex_normal(z.size, rng), device="cuda")
        e = rel(P.apply(d), _direct(eb, P, w, d, every, variant))
        print("  info", P.info(), "density", rep, "rel err", e, flush=True)
        assert e < (_bound(36) if kind == "clusters" else 1e-10)
    P.close()


@pytest.mark.parametrize("nterms", [16, 28, 44, 56])
def test_expansion_lengths(torch, nterms):
    """Every expansion length class (KMAX 16, 32, 48, 64) on clustered points, against the
    truncation bound of one-box separation (_bound)."""
    import electricbubbles_amd as eb
    from electricbubbles_amd.workloads import complex_normal
    z, wdz, rng = _case("clusters", 12000, 3)
    print("expansion length", nterms, flush=True)
    x, y, w = _tensors(torch, z, wdz)
    d = torch.tensor(complex_normal(z.size, rng), device="cuda")
    P = eb.AFMMPlan(x, y, w, nterms)
    e = rel(P.apply(d), _direct(eb, P, w, d, torch.arange(z.size, dtype=torch.int32, device="cuda")))
    bar = _bound(nterms)
    print("  rel err", e, "bar", bar, flush=True)
    assert e < bar


@pytest.mark.parametrize("kind", ["curves", "uniform"])
def test_two_to_the_20_against_direct_sum(torch, kind):
    """The 2^20-point workloads (4096 circles of 256 points; uniform points) at p = 36 and
    NAPB = 40 against the tree-free O(N^2) sum at 1,024 seeded targets."""
    import electricbubbles_amd as eb
    from electricbubbles_amd.workloads import complex_normal
    z, wdz, rng = _case(kind, 1 << 20, 20260512)
    print("2^20:", kind, flush=True)
    x, y, w = _tensors(torch, z, wdz)
    d = torch.tensor(complex_normal(z.size, rng), device="cuda")
    P = eb.AFMMPlan(x, y, w, 36)
    got = P.apply(d)
    tg = torch.tensor(np.sort(rng.choice(z.size, 1024, replace=False)).astype(np.int32), device="cuda")
    e = rel(got[tg.long()], _direct(eb, P, w, d, tg))
    e0 = rel(got[tg.long()], eb.direct_subset(x, y, w, d, tg))
    print("  against the sum on the caller's coordinates:", e0, flush=True)
    assert e0 < 1e-10
    print("  info", P.info(), "rel err at 1024 targets", e, flush=True)
    assert e < 1e-10
