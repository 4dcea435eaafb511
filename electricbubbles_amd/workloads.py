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
