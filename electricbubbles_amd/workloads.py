"""Synthetic boundary geometries of the benchmark configurations (numpy, seeded).

Every generator returns ``(z, dzdt, nd, h)``: points z (complex128), the parametric tangent
dz/dt (complex128), points per body nd (int32) and the quadrature spacing per body h
(float64) -- the arguments XAT+iYAT, DZDT, ND and H of the reference's CADAPDR
(cadapdrop.f:14-27).
"""
from __future__ import annotations

import numpy as np


def perturbed_circle(n, modes=((2, 0.05),), center=0j, radius=1.0):
    """r(theta) = R (1 + sum_k a_k cos k theta), n equispaced theta (config 1: 1 + 0.05 cos 2 theta)."""
    th = 2 * np.pi * np.arange(n) / n
    r = np.ones(n)
    dr = np.zeros(n)
    for k, a in modes:
        r = r + a * np.cos(k * th)
        dr = dr - a * k * np.sin(k * th)
    r *= radius
    dr *= radius
    e = np.exp(1j * th)
    z = center + r * e
    dz = (dr + 1j * r) * e
    return z, dz, np.array([n], np.int32), np.array([2 * np.pi / n])


def curves_cfg(ncurves, m, rng, rmin=0.01, rmax=0.05):
    """Closed circles with centres U[0,1]^2 and radii U[rmin, rmax], m points each (config 3a)."""
    c = rng.uniform(0.0, 1.0, (ncurves, 2))
    r = rng.uniform(rmin, rmax, ncurves)
    th = 2 * np.pi * np.arange(m) / m
    e = np.exp(1j * th)[None, :]
    z = (c[:, 0] + 1j * c[:, 1])[:, None] + r[:, None] * e
    dz = 1j * r[:, None] * e
    return (z.ravel(), dz.ravel(), np.full(ncurves, m, np.int32), np.full(ncurves, 2 * np.pi / m))


def uniform_cfg(n, rng):
    """n points uniform in the unit square with DZDT = 1 and H = 1 (config 3b)."""
    z = rng.uniform(0.0, 1.0, n) + 1j * rng.uniform(0.0, 1.0, n)
    return z, np.ones(n, np.complex128), np.array([n], np.int32), np.array([1.0])


def complex_normal(n, rng):
    """CN(0,1): real and imaginary parts i.i.d. N(0, 1/2)."""
    return (rng.normal(size=n) + 1j * rng.normal(size=n)) / np.sqrt(2.0)


def weights(dzdt, nd, h):
    """Per-point DZDT * H(body), the quadrature-weighted tangent of the plan API."""
    return dzdt * np.repeat(np.asarray(h, np.float64), np.asarray(nd))


def many_drops_cfg(ndrops, rng, box=20.0, gap=0.5):
    """The many-drop configuration: centres U[-box, box]^2 with rejection for a surface gap of
    at least `gap`, radii U[0.5, 1.5], three random Fourier shape modes of amplitude <= 0.05,
    N per drop = 2^round(log2(512 r)) clipped to 128..1024.  Returns (nd, rku): points per
    drop and the interleaved (x, y) state of all points, drop after drop."""
    zs, nds, placed = [], [], []
    while len(zs) < ndrops:
        r = rng.uniform(0.5, 1.5)
        c = rng.uniform(-box, box) + 1j * rng.uniform(-box, box)
        if any(abs(c - c2) < r + r2 + gap for c2, r2 in placed):
            continue
        placed.append((c, r))
        n = int(min(1024, max(128, 2 ** round(np.log2(512 * r)))))
        modes = [(int(m), rng.uniform(-0.05, 0.05)) for m in rng.choice(np.arange(2, 8), 3, replace=False)]
        z, _, _, _ = perturbed_circle(n, modes, center=c, radius=r)
        zs.append(z)
        nds.append(n)
    return np.array(nds, np.int32), np.concatenate(zs).view(np.float64).copy()
