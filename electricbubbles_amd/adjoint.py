"""Gradients of the drop evolution: a hand-written adjoint over the C-ABI of include/eb_adjoint.h.

The reference (electric.f, stokesbubb.f, laplace.f) has no derivatives.  This module gives
gradients of the summed drop deformation after K time steps with respect to the field
strength |E| (E0 = |E|^2, the reference's Maxwell-stress coefficient, stokesbubb.f:114-127),
the capillary number Ca of the imposed straining flow (the reference's CAP, psi = -i CAP z / 2,
stokesbubb.f:63-67, electric.f:869) and the Fourier modes of the initial shapes.

How the gradient is formed:

* F (electric.f:968-1109) is written as a chain of differentiable steps.  The steps that carry
  the arithmetic run in HIP through the C-ABI, each with a hand-written reverse:
    - the layer potentials (the Laplace Hilbert product, PVINTEV's alternating-point product,
      CADAPMOD's kernel in INTERFACENEW): O(N^2) pair sums; their reverse is a pair sum with
      sources and targets swapped (charges) plus the kernel's derivative (positions);
    - the Laplace solve (SOLVE_LAPL) and the Stokes solve (SOLVEI): GMRES on the device; their
      reverse is one GMRES solve of the transposed operator, A^T lam = ubar, and the geometry
      term -lam^T d(A u)/dgeometry (the pair-sum reverse again);
  the per-body spectral and pointwise steps between them (EXTRACTDAT's derivatives and
  curvature, PVINTEV's interpolation, SETVELF's FOURINT, INTERFACENEW's corrections, EQUIT,
  FILTER) are FFTs and elementwise formulas on device tensors, differentiated by torch.
* Time stepping (RK4, or the reference's Heun RK2, electric.f:762-786) runs forward once
  without a graph, keeping the state every `checkpoint` steps; the reverse sweep recomputes
  one segment at a time from its checkpoint with a graph and pulls the adjoint state back
  through it, so memory holds one segment's graph at a time.

Restrictions: bodies of equal length (reshaped to (k, N) for the FFTs), no gravity, no wall;
the step is fixed (the reference's arclength law, electric.f:239-244, is not used here) and
bodies are not doubled.  Everything runs on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _check, _ptr, _stream_ptr, _torch, lib

ADJ_EXPORTS = ("eb_adj_create", "eb_adj_destroy", "eb_adj_set_geometry", "eb_adj_laplace_apply",
               "eb_adj_stokes_apply", "eb_adj_laplace_solve", "eb_adj_stokes_solve", "eb_adj_cauchy",
               "eb_adj_cauchy_vjp", "eb_adj_stokes_pairs", "eb_adj_stokes_pairs_vjp", "eb_adj_set_mode")

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_f64 = ctypes.c_double
_bound = False
PI = math.pi


def _bind():
    global _bound
    L = lib()
    if not _bound:
        L.eb_adj_create.argtypes = [ctypes.POINTER(_vp), _i32, ctypes.POINTER(_i32), _i32, _vp]
        L.eb_adj_destroy.argtypes = [_vp]
        L.eb_adj_set_geometry.argtypes = [_vp] * 6
        L.eb_adj_set_mode.argtypes = [_vp, _i32]
        L.eb_adj_laplace_apply.argtypes = [_vp, _i32, _vp, _vp]
        L.eb_adj_stokes_apply.argtypes = [_vp, _i32, _vp, _vp]
        solve = [_vp, _i32, _vp, _vp, _f64, ctypes.POINTER(_i32), ctypes.POINTER(_f64), ctypes.POINTER(_i32)]
        L.eb_adj_laplace_solve.argtypes = solve
        L.eb_adj_stokes_solve.argtypes = solve
        L.eb_adj_cauchy.argtypes = [_vp, _i32, _vp, _i32, _vp, _vp, _i32, _vp]
        L.eb_adj_cauchy_vjp.argtypes = [_vp, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp]
        L.eb_adj_stokes_pairs.argtypes = [_vp, _i32, _vp, _vp, _vp, _f64, _vp]
        L.eb_adj_stokes_pairs_vjp.argtypes = [_vp, _i32, _vp, _vp, _vp, _f64, _vp, _vp, _vp, _vp]
        _bound = True
    return L


def _p(t):
    return _ptr(t) if t is not None else None


class AdjointSystem:
    """One simulation's operators on the device (an eb_adj handle): k bodies of nd[b] points."""

    def __init__(self, nd, maxl=100, stream=None):
        self._torch = _torch()
        self.nd = [int(v) for v in nd]
        self.k = len(self.nd)
        self.n = int(sum(self.nd))
        h = _vp()
        arr = (_i32 * self.k)(*self.nd)
        _check(_bind().eb_adj_create(ctypes.byref(h), self.k, arr, int(maxl), _stream_ptr(stream)))
        self._h = h
        self.iters = {"laplace": [], "stokes": [], "laplace_t": [], "stokes_t": []}

    def close(self):
        if getattr(self, "_h", None):
            _bind().eb_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, n):
        return self._torch.empty(n, dtype=self._torch.complex128, device="cuda")

    def set_mode(self, mode):
        """How the operators evaluate their layer potentials (eb_adj_set_mode): 0 auto (the
        O(N^2) sums up to 10240 points, FMM plans above), 1 FMM, 2 the O(N^2) sums.  In FMM
        mode the transposed operators run the transposed FMM product."""
        _check(_bind().eb_adj_set_mode(self._h, int(mode)))

    def set_geometry(self, z, dz, s, kap, zk):
        self._geom = [t.detach().contiguous() for t in (z, dz, s, kap, zk)]
        _check(_bind().eb_adj_set_geometry(self._h, *[_ptr(t) for t in self._geom]))

    def laplace_apply(self, x, transpose=False):
        y = self._torch.empty_like(x)
        _check(_bind().eb_adj_laplace_apply(self._h, int(transpose), _ptr(x.contiguous()), _ptr(y)))
        return y

    def stokes_apply(self, x, transpose=False):
        x = x.contiguous()
        y = self._torch.empty_like(x)
        _check(_bind().eb_adj_stokes_apply(self._h, int(transpose), _ptr(x), _ptr(y)))
        return y

    def _solve(self, fn, key, b, x, tol, transpose):
        it, ie, er = _i32(), _i32(), _f64()
        _check(fn(self._h, int(transpose), _ptr(b), _ptr(x), float(tol), ctypes.byref(it), ctypes.byref(er),
                  ctypes.byref(ie)))
        self.iters[key + ("_t" if transpose else "")].append((it.value, ie.value))
        return x

    def laplace_solve(self, b, tol, transpose=False, x0=None):
        b = b.detach().contiguous()
        x = (b if x0 is None else x0).detach().clone()
        return self._solve(_bind().eb_adj_laplace_solve, "laplace", b, x, tol, transpose)

    def stokes_solve(self, b, tol, transpose=False, x0=None):
        b = b.detach().contiguous()
        x = (b if x0 is None else x0).detach().clone()
        return self._solve(_bind().eb_adj_stokes_solve, "stokes", b, x, tol, transpose)

    def cauchy(self, t, s, c, self_sum):
        g = self._c(t.numel())
        _check(_bind().eb_adj_cauchy(self._h, t.numel(), _ptr(t), s.numel(), _ptr(s), _ptr(c), int(self_sum), _ptr(g)))
        return g

    def cauchy_vjp(self, t, s, c, self_sum, gbar, need=(True, True, True)):
        tb = self._c(t.numel()) if need[0] else None
        sb = self._c(s.numel()) if need[1] else None
        cb = self._c(s.numel()) if need[2] else None
        _check(_bind().eb_adj_cauchy_vjp(self._h, t.numel(), _ptr(t), s.numel(), _ptr(s), _ptr(c), int(self_sum),
                                         _ptr(gbar), _p(tb), _p(sb), _p(cb)))
        return tb, sb, cb

    def stokes_pairs(self, z, qa, qb, s2):
        g = self._c(z.numel())
        _check(_bind().eb_adj_stokes_pairs(self._h, z.numel(), _ptr(z), _ptr(qa), _ptr(qb), float(s2), _ptr(g)))
        return g

    def stokes_pairs_vjp(self, z, qa, qb, s2, gbar, need_z=True):
        n = z.numel()
        zb = self._c(n) if need_z else None
        qab, qbb = self._c(n), self._c(n)
        _check(_bind().eb_adj_stokes_pairs_vjp(self._h, n, _ptr(z), _ptr(qa), _ptr(qb), float(s2), _ptr(gbar), _p(zb),
                                               _ptr(qab), _ptr(qbb)))
        return zb, qab, qbb


# ------------------------------------------------------------------------------------------
# autograd steps whose arithmetic is HIP
# ------------------------------------------------------------------------------------------

def _functions():
    torch = _torch()

    class Cauchy(torch.autograd.Function):
        """g_i = sum_j c_j / (t_i - s_j) (j = i left out when self_sum)."""

        @staticmethod
        def forward(ctx, sys, t, s, c, self_sum):
            t_, s_, c_ = (x.detach().contiguous() for x in (t, s, c))
            ctx.sys, ctx.self_sum = sys, self_sum
            ctx.save_for_backward(t_, s_, c_)
            return sys.cauchy(t_, s_, c_, self_sum)

        @staticmethod
        def backward(ctx, gbar):
            t, s, c = ctx.saved_tensors
            need = ctx.needs_input_grad[1:4]
            tb, sb, cb = ctx.sys.cauchy_vjp(t, s, c, ctx.self_sum, gbar.contiguous(), need)
            return None, tb, sb, cb, None

    class StokesPairs(torch.autograd.Function):
        """CADAPDR (s2 = -1) / CADAPMOD (s2 = +1) kernel summed over all pairs j != i."""

        @staticmethod
        def forward(ctx, sys, z, qa, qb, s2):
            z_, qa_, qb_ = (x.detach().contiguous() for x in (z, qa, qb))
            ctx.sys, ctx.s2 = sys, s2
            ctx.save_for_backward(z_, qa_, qb_)
            return sys.stokes_pairs(z_, qa_, qb_, s2)

        @staticmethod
        def backward(ctx, gbar):
            z, qa, qb = ctx.saved_tensors
            zb, qab, qbb = ctx.sys.stokes_pairs_vjp(z, qa, qb, ctx.s2, gbar.contiguous(), ctx.needs_input_grad[1])
            return None, zb, qab, qbb, None

    class LaplaceSolve(torch.autograd.Function):
        """u = A(geometry)^-1 rhs for the Laplace operator (SOLVE_LAPL, laplace.f:335-425)."""

        @staticmethod
        def forward(ctx, P, rhs, z, dz, s, kap, zk):
            sys = P.sys
            sys.set_geometry(z, dz, s, kap, zk)
            x0, ctx.slot = P.cache.before_solve("lap")
            u = sys.laplace_solve(rhs, P.tol, x0=x0)
            P.cache.after_solve("lap", u)
            ctx.P = P
            ctx.save_for_backward(u, z.detach(), dz.detach(), s.detach(), kap.detach(), zk.detach())
            return u

        @staticmethod
        def backward(ctx, ubar):
            u, z, dz, s, kap, zk = ctx.saved_tensors
            P = ctx.P
            sys = P.sys
            sys.set_geometry(z, dz, s, kap, zk)
            key = ("lap", ctx.slot)
            lam = sys.laplace_solve(ubar.contiguous(), P.tol, transpose=True, x0=P.cache.adj_guess(key))
            P.cache.adj_store(key, lam)
            with torch.enable_grad():
                g = [x.detach().requires_grad_() for x in (z, dz, s, kap)]
                val = (lam * laplace_operator(sys, u, *g)).sum()
                gz, gdz, gs, gk = torch.autograd.grad(val, g)
            return None, lam, -gz, -gdz, -gs, -gk, None

    class StokesSolve(torch.autograd.Function):
        """sigma = A(geometry)^-1 rhs for the Stokes operator (SOLVEI, stokesbubb.f:294-415)."""

        @staticmethod
        def forward(ctx, P, rhs, z, dz, s, kap, zk):
            sys = P.sys
            sys.set_geometry(z, dz, s, kap, zk)
            b = torch.view_as_real(rhs.detach().contiguous()).reshape(-1)
            x0, ctx.slot = P.cache.before_solve("st")
            sg = torch.view_as_complex(sys.stokes_solve(b, P.tol, x0=x0).view(-1, 2))
            P.cache.after_solve("st", torch.view_as_real(sg).reshape(-1))
            ctx.P = P
            ctx.save_for_backward(sg, z.detach(), dz.detach(), s.detach(), kap.detach(), zk.detach())
            return sg

        @staticmethod
        def backward(ctx, sbar):
            sg, z, dz, s, kap, zk = ctx.saved_tensors
            P = ctx.P
            sys = P.sys
            sys.set_geometry(z, dz, s, kap, zk)
            b = torch.view_as_real(sbar.contiguous()).reshape(-1)
            key = ("st", ctx.slot)
            lam = torch.view_as_complex(sys.stokes_solve(b, P.tol, transpose=True,
                                                         x0=P.cache.adj_guess(key)).view(-1, 2))
            P.cache.adj_store(key, torch.view_as_real(lam).reshape(-1))
            with torch.enable_grad():
                g = [x.detach().requires_grad_() for x in (z, dz, s, kap, zk)]
                av = stokes_operator(sys, sg, *g)
                val = (lam.real * av.real + lam.imag * av.imag).sum()
                gz, gdz, gs, gk, gzk = torch.autograd.grad(val, g)
            return None, lam, -gz, -gdz, -gs, -gk, -gzk

    return Cauchy, StokesPairs, LaplaceSolve, StokesSolve


_FN = None


def fn():
    global _FN
    if _FN is None:
        _FN = _functions()
    return _FN


def _hpt(sys, like):
    """h = 2 pi / N_b of every point (cached on the system)."""
    torch = _torch()
    h = getattr(sys, "_h_pt", None)
    if h is None:
        h = torch.cat([torch.full((m,), 2.0 * PI / m, dtype=torch.float64, device=like.device) for m in sys.nd])
        sys._h_pt = h
    return h


def _body_sum(sys, v):
    """Sum over each body, broadcast back to its points (bodies of equal length)."""
    k = sys.k
    return v.view(k, -1).sum(dim=1, keepdim=True).expand(k, v.numel() // k).reshape(-1)


def laplace_operator(sys, u, z, dz, s, kap):
    """(A u)_i = u_i (1 + h kappa s / (2 pi) - 2 h s) - Im(CF_i) / pi, the operator of
    FASMVP_LAP (laplace.f:225-237) with CF the Hilbert product of the charges u h dz."""
    Cauchy = fn()[0]
    h = _hpt(sys, u)
    cf = Cauchy.apply(sys, z, z, u * dz * h, True)
    return u * (1.0 + h * kap * s / (2.0 * PI) - 2.0 * h * s) - cf.imag / PI


def stokes_charges(sg, dz, h):
    """QA, QB of CADAPDR (cadapdrop.f:218-220) for density sg."""
    alpha = -1j * dz * h / (2.0 * PI)
    gamma = -1j * dz.conj() * h / (2.0 * PI)
    return alpha * sg, gamma * sg - alpha * sg.conj()


def stokes_operator(sys, sg, z, dz, s, kap, zk):
    """FASMVP (stokesbubb.f:527-611): sg + za_b + GW + self + (i/2) pk_b z."""
    StokesPairs = fn()[1]
    h = _hpt(sys, s)
    qa, qb = stokes_charges(sg, dz, h)
    gw = StokesPairs.apply(sys, z, qa, qb, -1.0)
    k = sys.k
    za = h * _body_sum(sys, sg)
    zkp = zk.view(k, 1).expand(k, z.numel() // k).reshape(-1)
    e = dz / (z - zkp) ** 2
    pk = 2.0 * h * _body_sum(sys, (sg * e).real) / PI
    c0 = h * s * kap / (2.0 * PI)
    zn = -1j * dz / s
    return sg + za + gw + c0 * (sg - sg.conj() * zn * zn) + 0.5j * pk * z


# ------------------------------------------------------------------------------------------
# F (electric.f:968-1109) as a differentiable chain
# ------------------------------------------------------------------------------------------

class SolveCache:
    """Initial guesses of the GMRES solves inside one gradient.

    * The forward pass with checkpoints records every Laplace and Stokes solution; when a
      segment is recomputed with a graph, each solve starts from the solution it found the
      first time, so DGMRES stops at its initial residual test with that same solution.
    * In the reverse sweep, each transposed solve starts from the quadratic extrapolation of
      the adjoint solutions of the same stage in the three steps after it (the adjoint state
      is smooth in time); linear, or the one solution, while fewer exist.
    The convergence test (relative residual below the tolerance) is the same in every case.
    """

    def __init__(self, stages=4):
        self.mode = None  # None, "record" or "replay"
        self.stages = stages
        self.store = {"lap": [], "st": []}
        self.pos = {"lap": 0, "st": 0}
        self.count = {"lap": 0, "st": 0}
        self.adj = {}

    def reset(self, stages):
        self.__init__(stages)

    def seek(self, nsolve):
        """Replay from the nsolve-th solve of each kind."""
        self.mode = "replay"
        self.pos = {"lap": nsolve, "st": nsolve}
        self.count = {"lap": nsolve, "st": nsolve}

    def before_solve(self, kind):
        slot = self.count[kind] % self.stages
        self.count[kind] += 1
        if self.mode == "replay":
            x0 = self.store[kind][self.pos[kind]]
            self.pos[kind] += 1
            return x0, slot
        return None, slot

    def after_solve(self, kind, x):
        if self.mode == "record":
            self.store[kind].append(x.detach())

    def adj_guess(self, key):
        h = self.adj.get(key)
        if not h:
            return None
        if len(h) == 1:
            return h[-1]
        if len(h) == 2:
            return 2.0 * h[-1] - h[-2]
        return 3.0 * (h[-1] - h[-2]) + h[-3]

    def adj_store(self, key, lam):
        h = self.adj.setdefault(key, [])
        h.append(lam.detach())
        del h[:-3]


class Problem:
    """Constants of F for k bodies of N points each, and the device operators."""

    def __init__(self, k, N, tol=1e-12, maxl=100, device="cuda", mode=0):
        torch = _torch()
        if N % 2 or N < 8:
            raise ValueError("bodies need an even number of points (at least 8)")
        self.k, self.N, self.tol = int(k), int(N), float(tol)
        self.sys = AdjointSystem([N] * k, maxl)
        self.sys.set_mode(mode)
        n = N // 2
        i = np.arange(N)
        ks = np.where(i <= n, i, i - N).astype(np.float64)       # signed wavenumber, +n at Nyquist
        kw = ks.copy()
        kw[n] = 0.0                                                # FDIFFF zeroes the Nyquist mode
        fint = np.zeros(N, dtype=np.complex128)                    # FOURINT: 1 / (i k), 1 <= |k| <= n-1
        sel = (np.abs(ks) >= 1) & (np.abs(ks) <= n - 1)
        fint[sel] = 1.0 / (1j * ks[sel])
        mask = np.ones(N)
        mask[n] = 0.0
        m = np.where(i <= n, i, N - i)                             # FILTER's mode index
        ratio = (np.float32(1.0) * m.astype(np.float32) / np.float32(n)).astype(np.float64)
        filt = np.exp(-10.0 * ratio ** 25)
        filt[n] = 0.0
        dev = torch.device(device)
        self.h = 2.0 * PI / N
        self.kw = torch.tensor(kw, device=dev)
        self.fint = torch.tensor(fint, device=dev)
        self.mask = torch.tensor(mask, device=dev)
        self.shift = torch.tensor(np.exp(1j * PI * ks / N) * mask, device=dev)
        self.filt = torch.tensor(filt, device=dev)
        self.alph = torch.tensor(2.0 * PI * i / N, device=dev)
        self.alpha_last = 2.0 * PI * (N - 1.0) / N                  # stokesbubb.f:94 (last body)
        self.cache = SolveCache()

    def close(self):
        self.sys.close()


def _fdifff(P, f):
    torch = _torch()
    return torch.fft.ifft(1j * P.kw * torch.fft.fft(f, dim=1), dim=1)


def geometry(P, z):
    """EXTRACTDAT (electric.f:676-699): z_alpha, z_alpha_alpha, kappa, ckern, tha, dz, dsda, zk."""
    torch = _torch()
    Z = torch.fft.fft(z, dim=1)
    za = torch.fft.ifft(1j * P.kw * Z, dim=1)
    zaa = torch.fft.ifft(-(P.kw * P.kw) * Z, dim=1)
    xd, yd, xdd, ydd = za.real, za.imag, zaa.real, zaa.imag
    s = torch.sqrt(xd * xd + yd * yd)
    kap = -(xd * ydd - yd * xdd) / (s * s * s)
    return dict(Z=Z, dz=-za, s=s, kap=kap, ckern=-zaa / za, tha=-kap * s, zk=Z[:, 0] / P.N)


def _finter_curve(Zm):
    """The curve's interpolant as PVINTEV forms it with FINTER for x and for y
    (pv_integral.f:216-252): FINTER clears the imaginary part of coefficient N/2 - 1 (and the
    real part of the Nyquist one) of each real series, which halves the imaginary part of the
    +-(N/2 - 1) pair of x's and y's interpolants; in zhat = FFT(x + i y) only entries N/2 - 1 and
    N/2 + 1 change (Zm has its Nyquist entry cleared already)."""
    torch = _torch()
    nyk = Zm.shape[1] // 2
    zp, zm = Zm[:, nyk - 1], Zm[:, nyk + 1]
    xh = 0.5 * (zp + zm.conj())
    yh = -0.5j * (zp - zm.conj())
    xq = torch.complex(xh.real, 0.5 * xh.imag)
    yq = torch.complex(yh.real, 0.5 * yh.imag)
    out = Zm.clone()
    out[:, nyk - 1] = xq + 1j * yq
    out[:, nyk + 1] = xq.conj() + 1j * yq.conj()
    return out


def rhs(P, z, e0, cap):
    """dz/dt = F(z) for z of shape (k, N) complex; e0, cap real scalars or 0-d tensors."""
    torch = _torch()
    Cauchy, StokesPairs, LaplaceSolve, StokesSolve = fn()
    sys, N, h = P.sys, P.N, P.h
    g = geometry(P, z)
    dz, s, kap, zk = g["dz"], g["s"], g["kap"], g["zk"]
    fz, fdz, fs, fk = z.reshape(-1), dz.reshape(-1), s.reshape(-1), kap.reshape(-1)
    # GETBCs + SOLVE_LAPL (laplace.f:14-18, 355-384)
    u = LaplaceSolve.apply(P, -2.0 * fz.real, fz, fdz, fs, fk, zk).view(P.k, N)
    # DIRICH_NEU_MAP with PVINTEV (laplace.f:266-332, pv_integral.f:53-212): the curve and u dz
    # interpolated to the midpoints, the alternating-point Hilbert product at the original points
    Zm = _finter_curve(g["Z"] * P.mask)
    tgt = torch.fft.ifft(Zm, dim=1)
    mid = torch.fft.ifft(Zm * P.shift, dim=1)
    umid = torch.fft.ifft(torch.fft.fft(u * dz, dim=1) * P.shift, dim=1)
    cf = Cauchy.apply(sys, tgt.reshape(-1), mid.reshape(-1), h * umid.reshape(-1), False).view(P.k, N)
    psi = -cf.real / (2.0 * PI) - z.imag
    dphi = -_fdifff(P, psi).real / s
    # SETVELF (stokesbubb.f:4-135) with no gravity and no wall
    Fh = torch.fft.fft(dz * dphi * dphi, dim=1)
    zf1 = Fh[:, :1] / N
    C = torch.fft.ifft(Fh * P.fint, dim=1)
    zs = P.alpha_last * zf1 + C - C[:, :1]
    strain = 0.5 * cap * (z.imag + 1j * z.real)
    rhs_st = -0.5 * dz / s + strain + 0.25j * e0 * zs
    # SOLVEI (stokesbubb.f:294-415)
    sg = StokesSolve.apply(P, rhs_st.reshape(-1), fz, fdz, fs, fk, zk).view(P.k, N)
    # INTERFACENEW (electric.f:820-878): CADAPMOD of the density, then the corrections
    udot = _fdifff(P, sg)
    qa, qb = stokes_charges(sg, dz, h)
    gw = StokesPairs.apply(sys, fz, qa.reshape(-1), qb.reshape(-1), 1.0).view(P.k, N)
    zcorr = -1j * (h * g["ckern"].real * sg - 2.0 * h * udot) / (2.0 * PI)
    zn = -1j * dz / s
    zself = sg.conj() * (0.5 * h * s * kap) * zn * zn / PI
    gt = gw + zself + zcorr + strain
    un = (-1j * gt * zn.conj()).real
    # EQUIT + FOURINT (electric.f:1270-1331), the frame velocity, FILTER (electric.f:1076-1103)
    G = torch.fft.fft(-g["tha"] * un, dim=1)
    a0 = G[:, :1].real / N
    C = torch.fft.ifft(G * P.fint, dim=1)
    T = -(a0 * P.alph + C.real) + P.alph * a0
    zf = (-1j * un - T) * dz / s
    A = torch.fft.fft(zf, dim=1) * (P.filt / N)
    with torch.no_grad():
        mr = (A.real.abs() >= 1e-13).to(torch.float64)
        mi = (A.imag.abs() >= 1e-13).to(torch.float64)
    A = torch.complex(A.real * mr, A.imag * mi)
    return torch.fft.ifft(A, dim=1) * N


# ------------------------------------------------------------------------------------------
# shapes, deformation, time stepping
# ------------------------------------------------------------------------------------------

def initial_shapes(centres, modes, N, decay=0.0):
    """z_j = c + (1 + sum_m m^-decay (a_m cos(m th_j) + b_m sin(m th_j))) e^{i th_j},
    th_j = 2 pi j / N; modes (k, 2M): a_1..a_M, b_1..b_M per body.  Device-agnostic."""
    import torch
    k, twoM = modes.shape
    M = twoM // 2
    th = torch.arange(N, dtype=torch.float64, device=modes.device) * (2.0 * PI / N)
    m = torch.arange(1, M + 1, dtype=torch.float64, device=modes.device)
    w = m ** (-float(decay))
    cs, sn = w[:, None] * torch.cos(m[:, None] * th[None, :]), w[:, None] * torch.sin(m[:, None] * th[None, :])
    r = 1.0 + modes[:, :M] @ cs + modes[:, M:] @ sn
    return centres.view(k, 1) + r * torch.exp(1j * th)[None, :]


def deformation(P, z):
    """D = (L - B) / (L + B) per body, L and B the semi-axes of the ellipse with the body's area
    and second moments of area (trapezoid sums of Green's-theorem integrands).  z: (k, N)
    complex, counterclockwise; P is unused apart from its type (any object works), so the
    loss can be evaluated on any device."""
    import torch
    N = z.shape[1]
    kw = torch.fft.fftfreq(N, 1.0 / N, dtype=torch.float64, device=z.device)
    if N % 2 == 0:
        kw[N // 2] = 0.0
    za = torch.fft.ifft(1j * kw * torch.fft.fft(z, dim=1), dim=1)
    x, y, xa, ya = z.real, z.imag, za.real, za.imag
    h = 2.0 * PI / N
    A = 0.5 * h * (x * ya - y * xa).sum(1)
    mx = 0.5 * h * (x * x * ya).sum(1)
    my = -0.5 * h * (y * y * xa).sum(1)
    ixx = h * (x ** 3 * ya).sum(1) / 3.0
    iyy = -h * (y ** 3 * xa).sum(1) / 3.0
    ixy = 0.5 * h * (x * x * y * ya).sum(1)
    cx, cy = mx / A, my / A
    a = ixx - A * cx * cx
    c = iyy - A * cy * cy
    b = ixy - A * cx * cy
    rt = torch.sqrt(0.25 * (a - c) ** 2 + b * b)
    lp, lm = 0.5 * (a + c) + rt, 0.5 * (a + c) - rt
    sp, sm = torch.sqrt(lp), torch.sqrt(lm)
    return (sp - sm) / (sp + sm)


def step(P, z, dt, e0, cap, method="rk4"):
    if method == "rk4":
        k1 = rhs(P, z, e0, cap)
        k2 = rhs(P, z + 0.5 * dt * k1, e0, cap)
        k3 = rhs(P, z + 0.5 * dt * k2, e0, cap)
        k4 = rhs(P, z + dt * k3, e0, cap)
        return z + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    if method == "rk2":  # Heun, as RK2 (electric.f:773-783)
        p0 = dt * rhs(P, z, e0, cap)
        p1 = dt * rhs(P, z + p0, e0, cap)
        return z + 0.5 * p0 + 0.5 * p1
    raise ValueError(method)


def _stages(method):
    return 4 if method == "rk4" else 2


def forward_loss(P, centres, modes, E, Ca, steps, dt, method="rk4", decay=0.0):
    """Sum of the final deformations, no graph (the forward that finite differences use)."""
    torch = _torch()
    P.cache.reset(_stages(method))
    with torch.no_grad():
        z = initial_shapes(centres, modes, P.N, decay)
        for _ in range(steps):
            z = step(P, z, dt, E * E, Ca, method)
        return float(deformation(P, z).sum())


def gradient(P, centres, modes, E, Ca, steps, dt, checkpoint=20, method="rk4", decay=0.0):
    """Loss and its gradient with respect to |E|, Ca and the shape modes, by the reverse sweep
    with checkpoints every `checkpoint` steps and recomputation of each segment.
    Returns (loss, dL/d|E|, dL/dCa, dL/dmodes (k, 2M) tensor, timings)."""
    torch = _torch()
    dev = modes.device
    st = _stages(method)
    E_t = torch.tensor(float(E), dtype=torch.float64, device=dev, requires_grad=True)
    Ca_t = torch.tensor(float(Ca), dtype=torch.float64, device=dev, requires_grad=True)
    modes_t = modes.detach().clone().requires_grad_()
    t0 = time.perf_counter()
    z0 = initial_shapes(centres, modes_t, P.N, decay)
    ckpt = [z0.detach()]
    z = z0.detach()
    P.cache.reset(st)
    P.cache.mode = "record"
    with torch.no_grad():
        for s in range(steps):
            z = step(P, z, dt, E_t.detach() ** 2, Ca_t.detach(), method)
            if (s + 1) % checkpoint == 0 and s + 1 < steps:
                ckpt.append(z)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    zf = z.detach().requires_grad_()
    loss = deformation(P, zf).sum()
    (zbar,) = torch.autograd.grad(loss, zf)
    nseg = len(ckpt)
    for seg in range(nseg - 1, -1, -1):
        s0 = seg * checkpoint
        s1 = min(steps, s0 + checkpoint)
        zc = ckpt[seg].detach().requires_grad_()
        P.cache.seek(s0 * st)
        with torch.enable_grad():
            e0 = E_t * E_t
            zz = zc
            for _ in range(s0, s1):
                zz = step(P, zz, dt, e0, Ca_t, method)
            torch.autograd.backward(zz, zbar)
        zbar = zc.grad
    torch.autograd.backward(z0, zbar)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    P.cache.reset(st)
    return (float(loss.detach()), float(E_t.grad), float(Ca_t.grad), modes_t.grad.detach(),
            {"forward_s": t1 - t0, "reverse_s": t2 - t1, "total_s": t2 - t0})


def _trajectory_function():
    torch = _torch()

    class TrajectoryLoss(torch.autograd.Function):
        """loss = sum of final deformations as a function of (|E|, Ca, modes); backward runs
        the checkpointed reverse sweep of `gradient`."""

        @staticmethod
        def forward(ctx, E, Ca, modes, P, centres, steps, dt, checkpoint, method, decay):
            with torch.enable_grad():  # the reverse sweep builds its own graphs
                loss, gE, gCa, gm, _ = gradient(P, centres, modes.detach(), float(E), float(Ca), steps, dt,
                                                checkpoint, method, decay)
            f64 = dict(dtype=torch.float64, device=modes.device)
            ctx.save_for_backward(torch.tensor(gE, **f64), torch.tensor(gCa, **f64), gm)
            return torch.tensor(loss, dtype=torch.float64, device=modes.device)

        @staticmethod
        def backward(ctx, g):
            gE, gCa, gm = ctx.saved_tensors
            return g * gE, g * gCa, g * gm, None, None, None, None, None, None, None

    return TrajectoryLoss


_TL = None


def trajectory_loss(E, Ca, modes, P, centres, steps, dt, checkpoint=20, method="rk4", decay=0.0):
    """Autograd entry point: the summed final deformation as a differentiable function of the
    0-d tensors E (|E|) and Ca and the (k, 2M) tensor of shape modes.  The forward pass already
    runs the reverse sweep (one forward and one reverse in total); loss.backward() then hands
    the stored gradients to E, Ca and modes."""
    global _TL
    if _TL is None:
        _TL = _trajectory_function()
    return _TL.apply(E, Ca, modes, P, centres, steps, dt, checkpoint, method, decay)


def grid_problem(k_side=4, N=512, spacing=4.0, nmodes=64, amp=0.02, seed=0, device="cuda"):
    """The 16-drop adjoint configuration: drops on a k_side x k_side grid with the given
    spacing, radius 1, nmodes real Fourier modes per drop of amplitude U[0, amp]."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    g = (np.arange(k_side) - 0.5 * (k_side - 1)) * spacing
    cx, cy = np.meshgrid(g, g, indexing="xy")
    centres = torch.tensor((cx + 1j * cy).ravel(), device=device)
    modes = torch.tensor(rng.uniform(0.0, amp, (k_side * k_side, nmodes)), device=device)
    return centres, modes
