"""GPU parity of segmented FMM plans (eb_fmm_plan_create_segmented) through the C-ABI.

A segmented plan evaluates independent point sets (the simulations of a batch) in one tree:
each segment is binned on its own (own bounding box and scale) inside its own box of the grid,
and its sums run over its own points only.  Checker: the long-double O(N^2) sums of the oracle
run on each segment alone (oracle.direct, oracle.cauchy, oracle.direct_transpose).

Bars: 1e-10 relative per segment (2-norm over the segment's points) at 43 terms; a segment of
one point has GW = 0.
"""
import numpy as np
import pytest

import electricbubbles_amd as eb
import oracle
from electricbubbles_amd.workloads import complex_normal, curves_cfg, perturbed_circle, uniform_cfg, weights

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no GPU")
    return _t


def _segments(seed, shift=0j):
    """Seven segments of different sizes and shapes, all in overlapping parts of the plane (the
    sums must not see each other): circles, a perturbed drop, an empty segment, uniform points,
    a single point, a small drop, many circles.  Returns per segment (z, dzdt, nd, h)."""
    rng = np.random.default_rng(seed)
    segs = [curves_cfg(4, 256, rng)]
    z, dz, nd, h = perturbed_circle(300, [(2, 0.05), (3, 0.02)], center=0.5 + 0.5j, radius=0.4)
    segs.append((z, dz, nd, h))
    segs.append((np.zeros(0, complex), np.zeros(0, complex), np.zeros(0, np.int32), np.zeros(0)))
    segs.append(uniform_cfg(2048, rng))
    segs.append((np.array([0.3 + 0.3j]), np.array([1j]), np.array([1], np.int32), np.array([0.1])))
    segs.append(perturbed_circle(64, [(2, 0.1)], center=0.2 + 0.7j, radius=0.05))
    segs.append(curves_cfg(16, 256, rng))
    return [(s[0] + shift, s[1], s[2], s[3]) for s in segs]


def _pack(segs):
    off = np.concatenate([[0], np.cumsum([s[0].size for s in segs])]).astype(np.int32)
    z = np.concatenate([s[0] for s in segs])
    w = np.concatenate([weights(s[1], s[2], s[3]) if s[0].size else np.zeros(0, complex) for s in segs])
    return off, z, w


def _expect(variant, segs, vals, off, transpose=False):
    out = []
    for k, (z, dz, nd, h) in enumerate(segs):
        v = vals[off[k]:off[k + 1]]
        if z.size == 0:
            continue
        if z.size == 1:
            out.append((k, np.zeros(1, complex)))
            continue
        w = weights(dz, nd, h)
        if transpose:
            out.append((k, oracle.direct_transpose(variant, z.real, z.imag, w, v)))
        elif variant == eb.VARIANT_HILBERT:
            out.append((k, oracle.cauchy(z.real, z.imag, w * v)))
        else:
            out.append((k, oracle.direct(variant, nd, z.real, z.imag, dz, v, h)))
    return out


def _check(got, expect, off, bar):
    for k, e in expect:
        g = got[off[k]:off[k + 1]]
        if not np.any(e):
            assert np.abs(g).max() == 0.0, k
            continue
        err = rel(g, e)
        print(f"segment {k} ({off[k + 1] - off[k]} points): rel err {err:.2e}", flush=True)
        assert err < bar, (k, err)


@pytest.mark.parametrize("variant", [eb.VARIANT_DR, eb.VARIANT_MOD, eb.VARIANT_PLUS, eb.VARIANT_HILBERT])
def test_segmented_plan_matches_each_segment_alone(torch, variant):
    segs = _segments(3)
    off, z, w = _pack(segs)
    rng = np.random.default_rng(7 + variant)
    dens = complex_normal(z.size, rng)
    print("segmented plan, variant", variant, "segments", np.diff(off).tolist(), flush=True)
    t = lambda a: torch.tensor(a, device="cuda")
    plan = eb.FMMPlan.segmented(t(z.real.copy()), t(z.imag.copy()), t(w), off, 5, 43, variant)
    got = plan.apply(t(dens)).cpu().numpy()
    _check(got, _expect(variant, segs, dens, off), off, 1e-10)
    # the transpose, by segment
    lam = complex_normal(z.size, rng)
    got_t = plan.apply_transpose(t(lam)).cpu().numpy()
    _check(got_t, _expect(variant, segs, lam, off, transpose=True), off, 1e-10)
    # moved points, same segments: the rebuilt plan against the moved segments
    segs2 = _segments(3, shift=0.37 - 1.2j)
    segs2 = [(s[0] * (1.0 + 0.3 * k), s[1] * (1.0 + 0.3 * k), s[2], s[3]) for k, s in enumerate(segs2)]
    off2, z2, w2 = _pack(segs2)
    plan.rebuild(t(z2.real.copy()), t(z2.imag.copy()), t(w2))
    got2 = plan.apply(t(dens)).cpu().numpy()
    _check(got2, _expect(variant, segs2, dens, off2), off2, 1e-10)
    plan.close()


def test_segmented_plan_of_many_equal_segments(torch):
    """70 single drops (three levels of segment boxes, most of the last one empty) against each
    drop alone; and one segment holding all points of one drop gives the ordinary plan's result
    to rounding."""
    rng = np.random.default_rng(11)
    segs = []
    for k in range(70):
        m = int(rng.integers(32, 160)) * 2
        segs.append(perturbed_circle(m, [(2, 0.05 * rng.uniform())], center=complex(*rng.uniform(-1, 1, 2)),
                                     radius=float(rng.uniform(0.5, 2.0))))
    off, z, w = _pack(segs)
    dens = complex_normal(z.size, rng)
    t = lambda a: torch.tensor(a, device="cuda")
    plan = eb.FMMPlan.segmented(t(z.real.copy()), t(z.imag.copy()), t(w), off, 4, 43)
    got = plan.apply(t(dens)).cpu().numpy()
    _check(got, _expect(eb.VARIANT_DR, segs, dens, off), off, 1e-10)
    plan.close()
    z1, dz1, nd1, h1 = segs[0]
    w1 = weights(dz1, nd1, h1)
    d1 = dens[:z1.size]
    one = eb.FMMPlan.segmented(t(z1.real.copy()), t(z1.imag.copy()), t(w1), [0, z1.size], 4, 43)
    plain = eb.FMMPlan(t(z1.real.copy()), t(z1.imag.copy()), t(w1), 4, 43)
    a = one.apply(t(d1)).cpu().numpy()
    b = plain.apply(t(d1)).cpu().numpy()
    assert rel(a, b) < 1e-12


def test_segmented_plan_rejects_bad_offsets(torch):
    t = lambda a: torch.tensor(a, device="cuda")
    z = np.exp(2j * np.pi * np.arange(64) / 64)
    x, y, w = t(z.real.copy()), t(z.imag.copy()), t(z * 0.1)
    for off in ([0, 40, 30, 64], [0, 32, 63], [1, 32, 64]):
        with pytest.raises(eb.EBError):
            eb.FMMPlan.segmented(x, y, w, off, 3, 30)
    with pytest.raises(eb.EBError):  # 4^5 segments' level plus 6 below exceeds 10 levels
        eb.FMMPlan.segmented(x, y, w, np.linspace(0, 64, 1000).astype(np.int32), 6, 30)
