"""Autotuner for the FMM plan: the expansion length p and the tree depth (leaf size).

The reference fixes both by hand: NTERMS = -NINT(ln TOL / ln 2) (cadapdrop.f:126) and LEVMAX
from its caller (stokesbubb.f:564).  For a given point set and accuracy target this picks

* p: the shortest expansion whose result is within `target` (relative, 2-norm over up to 1024
  seeded random targets) of the O(N^2) sum over all sources at those targets (eb_direct_subset,
  which shares nothing with the FMM's tree, lists or expansions);
* LEVMAX: the depth with the shortest measured apply (tree kept, as inside a GMRES solve)
  among depths whose mean occupancy of the non-empty leaves lies in [lo, hi] points, each
  checked against the same accuracy target;
* the near-field tile: the number of targets per P2P work item (eb_fmm_set_p2p_tile) with
  the shortest measured apply at that depth and length.

Every candidate is timed on the device with the real data; results are cached per
(n, variant, target, geometry fingerprint), in memory and optionally in a JSON file.
"""
from __future__ import annotations

import json
import math
import os
import time

from . import FMMPlan, MODE_FMM, VARIANT_DR, _torch, direct_subset

P_MAX = 60  # the longest expansion the reference allows (cadapdrop.f:129-133)
N_CHECK = 1024  # targets of the accuracy check
_CACHE: dict = {}


def _fingerprint(x, y):
    # bounding box and first moments at float32 resolution: same data -> same key
    torch = _torch()
    v = torch.stack([x.min(), x.max(), y.min(), y.max(), x.mean(), y.mean()]).cpu().tolist()
    return ",".join(f"{a:.6e}" for a in v)


def _time_apply(plan, dens, reps):
    torch = _torch()
    out = plan.apply(dens)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.apply(dens, out)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, out


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def tune_fmm(x, y, wdz, dens=None, target=1e-10, variant=VARIANT_DR, nterms=(16, 24, 28, 32, 36, 40, 48),
             occupancy=(4.0, 64.0), tiles=(16, 24, 32, 42, 64, 96, 128), reps=3, cache_file=None, seed=0):
    """Returns {"levmax", "nterms", "p2p_tile", "ms", "err", "candidates": [...]} for the point set (x, y)
    with weights wdz (device tensors as FMMPlan takes them).  dens: the density to time and
    check with (complex, default: a fixed pseudo-random one).  err is the relative 2-norm
    difference from the O(N^2) sum at the check targets."""
    torch = _torch()
    n = int(x.numel())
    key = f"{n}|{variant}|{target:.3e}|{_fingerprint(x, y)}"
    if key in _CACHE:
        return _CACHE[key]
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            disk = json.load(f)
        if key in disk:
            _CACHE[key] = disk[key]
            return disk[key]
    if dens is None:
        g = torch.Generator(device=x.device).manual_seed(seed)
        dens = torch.complex(torch.randn(n, generator=g, device=x.device, dtype=torch.float64),
                             torch.randn(n, generator=g, device=x.device, dtype=torch.float64))
    # candidate depths: mean points per non-empty leaf within the occupancy window
    levels = []
    for L in range(2, 11):
        plan = FMMPlan(x, y, wdz, L, 16, variant=variant, mode=MODE_FMM)
        occ = n / max(1, plan.info()["leaves"])
        plan.close()
        if occupancy[0] <= occ <= occupancy[1]:
            levels.append((L, occ))
    if not levels:
        L = max(2, min(10, int(round(math.log(max(n, 4) / 16.0, 4)))))
        levels = [(L, float("nan"))]
    # the reference values: the O(N^2) sum at seeded random targets
    g = torch.Generator(device="cpu").manual_seed(seed + 1)
    tgt = torch.randperm(n, generator=g)[:min(n, N_CHECK)].sort().values.to(torch.int32).to(x.device)
    ref = direct_subset(x, y, wdz, dens, tgt, variant)
    tl = tgt.long()
    # p: shortest expansion within target of the direct sum, at the middle candidate depth
    Lm = levels[len(levels) // 2][0]
    p_sel = P_MAX
    p_scan = []
    for p in sorted(nterms):
        plan = FMMPlan(x, y, wdz, Lm, p, variant=variant, mode=MODE_FMM)
        err = _rel(plan.apply(dens)[tl], ref)
        plan.close()
        p_scan.append({"nterms": p, "err": err})
        if err <= target:
            p_sel = p
            break
    # depth: fastest apply at p_sel that still meets the target
    cands = []
    for L, occ in levels:
        plan = FMMPlan(x, y, wdz, L, p_sel, variant=variant, mode=MODE_FMM)
        t, out = _time_apply(plan, dens, reps)
        err = _rel(out[tl], ref)
        plan.close()
        cands.append({"levmax": L, "points_per_leaf": round(occ, 2), "ms": round(t * 1e3, 4), "err": err})
    ok = [c for c in cands if c["err"] <= target] or cands
    best = min(ok, key=lambda c: c["ms"])
    # near-field tile at the chosen depth and length
    plan = FMMPlan(x, y, wdz, best["levmax"], p_sel, variant=variant, mode=MODE_FMM)
    tile_scan = []
    for tm in tiles:
        plan.set_p2p_tile(tm)
        t, _ = _time_apply(plan, dens, reps)
        tile_scan.append({"p2p_tile": tm, "ms": round(t * 1e3, 4)})
    plan.close()
    bt = min(tile_scan, key=lambda c: c["ms"])
    res = {"levmax": best["levmax"], "nterms": p_sel, "p2p_tile": bt["p2p_tile"], "ms": bt["ms"], "err": best["err"],
           "target": target, "check": f"O(N^2) sum at {int(tgt.numel())} targets", "candidates": cands, "p_scan": p_scan, "p_scan_levmax": Lm, "tile_scan": tile_scan}
    _CACHE[key] = res
    if cache_file:
        disk = {}
        if os.path.exists(cache_file):
            with open(cache_file) as f:
                disk = json.load(f)
        disk[key] = res
        with open(cache_file, "w") as f:
            json.dump(disk, f, indent=1)
    return res
