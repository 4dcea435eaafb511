"""GPU parity of the transposed products (eb_fmm_apply_transpose) through the C-ABI.

The transpose is of the real-linear map dens -> GW under Re sum conj(a) b: the same layer
potential with sources and targets swapped, run by the plan's own passes (FMM or O(N^2)).
Checkers: the long-double transposed sum oracle.direct_transpose (pinned on the CPU to the dense
transpose of the forward sums) and the defining identity Re <lam, A dens> = Re <A^T lam, dens>.

Bars: FMM within 1e-10 relative of the long-double sum (the forward's bar at the same depth and
expansion length), the O(N^2) mode within 1e-11: its double-precision sums over uniform random
points, whose closest pairs dominate the 1/conj(d)^2 term, measured 1.9e-12 against the long
double (the forward's O(N^2) sums 1e-13 .. 1e-12 at the same points).
"""
import numpy as np
import pytest

import electricbubbles_amd as eb
import oracle
from electricbubbles_amd.workloads import complex_normal, curves_cfg, uniform_cfg, weights

pytestmark = pytest.mark.gpu

VARIANTS = [eb.VARIANT_DR, eb.VARIANT_MOD, eb.VARIANT_PLUS, eb.VARIANT_HILBERT]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no GPU")
    return _t


def _case(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "curves":
        z, dz, nd, h = curves_cfg(n // 256, 256, rng)
    else:
        z, dz, nd, h = uniform_cfg(n, rng)
    return rng, z, weights(dz, nd, h)


def _plan(torch, z, wdz, levmax, p, variant, mode):
    x = torch.tensor(z.real.copy(), device="cuda")
    y = torch.tensor(z.imag.copy(), device="cuda")
    w = torch.tensor(wdz, device="cuda")
    return eb.FMMPlan(x, y, w, levmax, p, variant, mode)


@pytest.mark.parametrize("mode", [eb.MODE_FMM, eb.MODE_DIRECT])
@pytest.mark.parametrize("kind", ["curves", "uniform"])
@pytest.mark.parametrize("variant", VARIANTS)
def test_transpose_against_long_double_sum(torch, variant, kind, mode):
    rng, z, wdz = _case(kind, 16384, 31 + variant)
    n = z.size
    lam = complex_normal(n, rng)
    print(f"transpose variant={variant} {kind} mode={mode}", flush=True)
    plan = _plan(torch, z, wdz, 6, 43, variant, mode)
    mu = plan.apply_transpose(torch.tensor(lam, device="cuda")).cpu().numpy()
    tgt = np.sort(rng.choice(n, 2048, replace=False)).astype(np.int32)
    ref = oracle.direct_transpose(variant, z.real, z.imag, wdz, lam, tgt)
    e = rel(mu[tgt], ref)
    print(f"rel err {e:.2e}", flush=True)
    assert e < (1e-10 if mode == eb.MODE_FMM else 1e-11), e
    # the forward product is not its own transpose (a transposed-by-mistake path fails here)
    fwd = plan.apply(torch.tensor(lam, device="cuda")).cpu().numpy()
    assert rel(fwd[tgt], ref) > 1e-3
    plan.close()


@pytest.mark.parametrize("mode", [eb.MODE_FMM, eb.MODE_DIRECT])
@pytest.mark.parametrize("variant", VARIANTS)
def test_transpose_identity(torch, variant, mode):
    """Re <lam, A dens> = Re <A^T lam, dens> for random lam and dens, through both entry
    points of one plan; after a rebuild to new coordinates it still holds."""
    rng, z, wdz = _case("curves", 8192, 5 + variant)
    n = z.size
    plan = _plan(torch, z, wdz, 5, 43, variant, mode)
    for rebuild in (False, True):
        if rebuild:
            z2 = z * (1.0 + 0.01j) + 0.003
            plan.rebuild(torch.tensor(z2.real.copy(), device="cuda"), torch.tensor(z2.imag.copy(), device="cuda"),
                         torch.tensor(wdz * 1.01, device="cuda"))
        lam = complex_normal(n, rng)
        dens = complex_normal(n, rng)
        a = plan.apply(torch.tensor(dens, device="cuda")).cpu().numpy()
        at = plan.apply_transpose(torch.tensor(lam, device="cuda")).cpu().numpy()
        lhs = float(np.real(np.vdot(lam, a)))
        rhs = float(np.real(np.vdot(at, dens)))
        scale = np.linalg.norm(lam) * np.linalg.norm(a)
        assert abs(lhs - rhs) < (1e-10 if mode == eb.MODE_FMM else 1e-13) * scale, (rebuild, lhs, rhs)
    plan.close()


def test_transpose_two_to_the_20(torch):
    """The 2^20 curves of the headline at its depth and expansion length (LEVMAX 8, p = 36):
    the transposed product at 512 random targets against the long-double sum over all 2^20
    sources."""
    rng, z, wdz = _case("curves", 1 << 20, 20260303)
    n = z.size
    lam = complex_normal(n, rng)
    print("transpose 2^20 curves, LEVMAX 8, p 36", flush=True)
    plan = _plan(torch, z, wdz, 8, 36, eb.VARIANT_DR, eb.MODE_FMM)
    mu = plan.apply_transpose(torch.tensor(lam, device="cuda")).cpu().numpy()
    tgt = np.sort(np.random.default_rng(9).choice(n, 512, replace=False)).astype(np.int32)
    ref = oracle.direct_transpose(eb.VARIANT_DR, z.real, z.imag, wdz, lam, tgt)
    e = rel(mu[tgt], ref)
    print(f"transpose 2^20: rel err {e:.2e}", flush=True)
    assert e < 1e-10
    plan.close()


def test_adaptive_plan_has_no_transpose(torch):
    rng, z, wdz = _case("curves", 4096, 3)
    x = torch.tensor(z.real.copy(), device="cuda")
    y = torch.tensor(z.imag.copy(), device="cuda")
    plan = eb.AFMMPlan(x, y, torch.tensor(wdz, device="cuda"), 36, napb=40)
    with pytest.raises(eb.EBError):
        plan.apply_transpose(torch.tensor(complex_normal(z.size, rng), device="cuda"))
    plan.close()
