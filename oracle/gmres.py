"""CPU restatement of DGMRES for the settings on the Stokes path (test infrastructure only).

DGMRES (dgmres.f:2-553) with ITOL = 0, JSCAL = 0, JPRE = 0 and KMP = MAXL, in the operation
order of DPIGMR (dpigmr.f:295-437), DORTH (dorth.f:90-121), DHEQR (dheqr.f:82-176), DHELS
(dhels.f), DRLCAL (drlcal.f) and the ITOL = 0 branch of ISDGMR (isdgmr.f:289-394).  Inner
products use numpy (the reference's DDOT/DNRM2 sum in another order: the iterates agree to
rounding, the iteration counts exactly unless a test sits on the tolerance).

Pinned by tests/test_gmres_cpu.py against the reference's DGMRES itself (oracle/_ref and the
fixtures tests/golden/gmres_*.npz).
"""
from __future__ import annotations

import math

import numpy as np


def _dheqr_column(hes, ll, q):
    """One new Hessenberg column ll (1-based): previous rotations, then a new one."""
    for k in range(1, ll):
        c, s = q[2 * (k - 1)], q[2 * (k - 1) + 1]
        t1, t2 = hes[k - 1, ll - 1], hes[k, ll - 1]
        hes[k - 1, ll - 1] = c * t1 - s * t2
        hes[k, ll - 1] = s * t1 + c * t2
    t1, t2 = hes[ll - 1, ll - 1], hes[ll, ll - 1]
    if t2 == 0.0:
        c, s = 1.0, 0.0
    elif abs(t2) >= abs(t1):
        t = t1 / t2
        s = -1.0 / math.sqrt(1.0 + t * t)
        c = -s * t
    else:
        t = t2 / t1
        c = 1.0 / math.sqrt(1.0 + t * t)
        s = -c * t
    q[2 * (ll - 1)], q[2 * (ll - 1) + 1] = c, s
    hes[ll - 1, ll - 1] = c * t1 - s * t2
    return ll if hes[ll - 1, ll - 1] == 0.0 else 0


def _dhels(hes, n, q, b):
    for k in range(1, n + 1):
        c, s = q[2 * (k - 1)], q[2 * (k - 1) + 1]
        t1, t2 = b[k - 1], b[k]
        b[k - 1] = c * t1 - s * t2
        b[k] = s * t1 + c * t2
    for kb in range(1, n + 1):
        k = n + 1 - kb
        b[k - 1] = b[k - 1] / hes[k - 1, k - 1]
        t = -b[k - 1]
        b[:k - 1] += t * hes[:k - 1, k - 1]
    return b


def dgmres(apply, b, x, maxl=10, tol=0.0, nrmax=0):
    """Solve A x = b from the initial guess x.  `apply(v)` returns A v.  maxl, nrmax follow
    IGWK(1) and IGWK(5).  Returns (x, iter, err, ierr)."""
    b = np.asarray(b, np.float64)
    x = np.array(x, np.float64)
    n = b.size
    if maxl == 0:
        maxl = 10
    maxl = min(maxl, n)
    if nrmax == 0:
        nrmax = 10
    elif nrmax == -1:
        nrmax = 0
    if tol == 0.0:
        tol = 500 * 2.0 ** -53
    bnrm = float(np.linalg.norm(b))
    r = b - apply(x)
    it = 0
    err = 0.0
    nrsts = 0
    while True:
        if nrsts > nrmax:
            return x, it, err, 1
        # DPIGMR
        r0nrm = float(np.linalg.norm(r))
        err = r0nrm / bnrm
        if err <= tol:
            return x, it, err, 0
        V = np.zeros((n, maxl + 1))
        V[:, 0] = r * (1.0 / r0nrm)
        hes = np.zeros((maxl + 1, maxl))
        q = np.zeros(2 * maxl)
        prod = 1.0
        rho = r0nrm
        snormw = 0.0
        lgmr = 0
        iflag = None
        for ll in range(1, maxl + 1):
            lgmr = ll
            w = apply(V[:, ll - 1])
            # DORTH
            vnrm = float(np.linalg.norm(w))
            for i in range(1, ll + 1):
                hes[i - 1, ll - 1] = float(V[:, i - 1] @ w)
                w = w + (-hes[i - 1, ll - 1]) * V[:, i - 1]
            snormw = float(np.linalg.norm(w))
            if not (vnrm + 0.001 * snormw != vnrm):
                sumdsq = 0.0
                for i in range(1, ll + 1):
                    tem = -float(V[:, i - 1] @ w)
                    if hes[i - 1, ll - 1] + 0.001 * tem == hes[i - 1, ll - 1]:
                        continue
                    hes[i - 1, ll - 1] -= tem
                    w = w + tem * V[:, i - 1]
                    sumdsq += tem * tem
                if sumdsq != 0.0:
                    snormw = math.sqrt(max(0.0, snormw ** 2 - sumdsq))
            V[:, ll] = w
            hes[ll, ll - 1] = snormw
            if _dheqr_column(hes, ll, q) == ll:
                iflag = 2
                break
            prod = prod * q[2 * ll - 1]
            rho = abs(prod * r0nrm)
            err = rho / bnrm
            if err <= tol:
                iflag = 0
                break
            if ll == maxl:
                break
            V[:, ll] = V[:, ll] * (1.0 / snormw)
        if iflag is None:
            iflag = 1 if rho < r0nrm else 2
        it += lgmr
        if iflag == 2:
            return x, it, err, 2
        if iflag == 1 and nrmax > 0:
            # DRLCAL: residual of the updated iterate for the restart
            rl = V[:, 0].copy()
            for i in range(1, lgmr):
                rl = q[2 * i - 1] * rl + q[2 * i - 2] * V[:, i]
            rl = q[2 * lgmr - 1] * rl + (q[2 * lgmr - 2] / snormw) * V[:, lgmr]
            r = rl * (r0nrm * prod)
        bb = np.zeros(lgmr + 1)
        bb[0] = r0nrm
        _dhels(hes, lgmr, q, bb)
        z = np.zeros(n)
        for i in range(lgmr):
            z = z + bb[i] * V[:, i]
        x = x + z
        if iflag == 0:
            return x, it, err, 0
        nrsts += 1
