"""Per-body spectral geometry, RK2 stage updates and the Stokes matvec on the GPU.

ctypes layer over include/eb_spectral.h.  Device memory comes from torch tensors (complex128
for COMPLEX*16 arrays, float64 for REAL*8 ones); nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, _ptr, _stream_ptr, _torch, lib

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_ip = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)

SPECTRAL_EXPORTS = (
    "eb_bodies_create", "eb_bodies_destroy", "eb_bodies_npoints", "eb_dcfft", "eb_fdifff", "eb_filter",
    "eb_equit", "eb_extractdat", "eb_ftail", "eb_rk2_stage1", "eb_rk2_stage2", "eb_matvec_create",
    "eb_matvec_apply", "eb_matvec_destroy", "eb_fasmvp", "eb_sweep_create", "eb_sweep_apply", "eb_sweep_parts",
    "eb_sweep_unknown_offsets", "eb_sweep_destroy", "eb_rk4_stage", "eb_deformation", "eb_sweep_set_mode",
)

_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    L.eb_bodies_create.argtypes = [ctypes.POINTER(_vp), _i32, _ip, _vp]
    L.eb_bodies_destroy.argtypes = [_vp]
    L.eb_bodies_npoints.argtypes = [_vp]
    L.eb_dcfft.argtypes = [_vp, _vp, _i32, _vp]
    L.eb_fdifff.argtypes = [_vp, _vp, _vp, _vp]
    L.eb_filter.argtypes = [_vp, _vp, _vp]
    L.eb_equit.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.eb_extractdat.argtypes = [_vp] * 13
    L.eb_ftail.argtypes = [_vp] * 7
    L.eb_rk2_stage1.argtypes = [_i32, ctypes.c_double, _vp, _vp, _vp, _vp, _vp]
    L.eb_rk2_stage2.argtypes = [_i32, ctypes.c_double, _vp, _vp, _vp, _vp]
    L.eb_rk4_stage.argtypes = [_i32, _i32, ctypes.c_double, _vp, _vp, _vp, _vp, _vp]
    L.eb_deformation.argtypes = [_vp, _vp, _vp, _vp]
    L.eb_sweep_set_mode.argtypes = [_vp, _i32]
    L.eb_matvec_create.argtypes = [ctypes.POINTER(_vp), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32, _vp]
    L.eb_matvec_apply.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.eb_matvec_destroy.argtypes = [_vp]
    L.eb_fasmvp.argtypes = [_ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.eb_fasmvp.restype = _i32
    L.eb_sweep_create.argtypes = [ctypes.POINTER(_vp), _vp, _i32, _ip, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.eb_sweep_apply.argtypes = [_vp, _vp, _vp, _vp]
    L.eb_sweep_parts.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.eb_sweep_unknown_offsets.argtypes = [_vp]
    L.eb_sweep_unknown_offsets.restype = _ip
    L.eb_sweep_destroy.argtypes = [_vp]
    _bound = True
    return L


def _p(t):
    return _vp(0) if t is None else _ptr(t)


class Bodies:
    """nbody bodies with nd[k] points each, stored back to back (per-point device arrays)."""

    def __init__(self, nd, stream=None):
        torch = _torch()
        self._torch = torch
        self.nd = np.ascontiguousarray(nd, np.int32)
        self.n = int(self.nd.sum())
        self.nbody = self.nd.size
        self.stream = stream
        L = _bind()
        h = _vp()
        _check(L.eb_bodies_create(ctypes.byref(h), self.nbody, self.nd.ctypes.data_as(_ip), _stream_ptr(stream)))
        self._h = h

    def _s(self):
        return _stream_ptr(self.stream)

    def dcfft(self, c, sign):
        """DCFFTF (sign -1) / DCFFTB (sign +1) of every body; returns a new tensor."""
        out = c.clone()
        _check(_bind().eb_dcfft(self._h, _ptr(out), sign, self._s()))
        return out

    def deformation(self, rku):
        """D = (L - B) / (L + B) of every body of the state rku (eb_deformation)."""
        D = self._torch.empty(self.nbody, dtype=self._torch.float64, device="cuda")
        _check(_bind().eb_deformation(self._h, _ptr(rku.contiguous()), _ptr(D), self._s()))
        return D

    def fdifff(self, f):
        out = self._torch.empty_like(f)
        _check(_bind().eb_fdifff(self._h, _ptr(f.contiguous()), _ptr(out), self._s()))
        return out

    def filter(self, cn):
        out = cn.clone()
        _check(_bind().eb_filter(self._h, _ptr(out), self._s()))
        return out

    def equit(self, tha, un):
        out = self._torch.empty_like(tha)
        _check(_bind().eb_equit(self._h, _ptr(tha.contiguous()), _ptr(un.contiguous()), _ptr(out), self._s()))
        return out

    def extractdat(self, rku):
        torch = self._torch
        dev = rku.device
        f64 = dict(dtype=torch.float64, device=dev)
        c128 = dict(dtype=torch.complex128, device=dev)
        o = dict(x=torch.empty(self.n, **f64), y=torch.empty(self.n, **f64), zk=torch.empty(self.nbody, **c128),
                 dz=torch.empty(self.n, **c128), tha=torch.empty(self.n, **f64), dsda=torch.empty(self.n, **f64),
                 rkappa=torch.empty(self.n, **f64), ckern=torch.empty(self.n, **c128),
                 area=torch.empty(self.nbody, **f64), xls=torch.empty(2, **f64))
        _check(_bind().eb_extractdat(self._h, _ptr(rku.contiguous()), _ptr(o["x"]), _ptr(o["y"]), _ptr(o["zk"]),
                                     _ptr(o["dz"]), _ptr(o["tha"]), _ptr(o["dsda"]), _ptr(o["rkappa"]),
                                     _ptr(o["ckern"]), _ptr(o["area"]), _ptr(o["xls"]), self._s()))
        return o

    def ftail(self, tha, un, dz, dsda):
        out = self._torch.empty(self.n, dtype=self._torch.complex128, device=tha.device)
        _check(_bind().eb_ftail(self._h, _ptr(tha.contiguous()), _ptr(un.contiguous()), _ptr(dz.contiguous()),
                                _ptr(dsda.contiguous()), _ptr(out), self._s()))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().eb_bodies_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rk2_stage1(dt, ybar, dy, stream=None):
    torch = _torch()
    phi0 = torch.empty_like(ybar)
    y = torch.empty_like(ybar)
    _check(_bind().eb_rk2_stage1(ybar.numel(), dt, _ptr(ybar), _ptr(dy), _ptr(phi0), _ptr(y), _stream_ptr(stream)))
    return phi0, y


def rk2_stage2(dt, phi0, dy, ybar, stream=None):
    """In place on ybar."""
    _check(_bind().eb_rk2_stage2(ybar.numel(), dt, _ptr(phi0), _ptr(dy), _ptr(ybar), _stream_ptr(stream)))
    return ybar


def rk4_stage(stage, dt, y, dy, acc, ys, stream=None):
    """Stage update `stage` (0..3) of classic RK4 (eb_rk4_stage): acc and ys are the caller's
    buffers; stage 3 updates y in place."""
    _check(_bind().eb_rk4_stage(y.numel(), int(stage), float(dt), _ptr(y), _ptr(dy), _ptr(acc), _ptr(ys),
                                _stream_ptr(stream)))
    return y


class MatVec:
    """MATVEC + FASMVP of the Stokes GMRES solve with the GPU FMM (stokesbubb.f:419-614)."""

    def __init__(self, bodies, x, y, dz, dsdth, rkappa, zk, levmax=7, nterms=43, mode=1, stream=None):
        self.bodies = bodies
        self.stream = stream
        self._keep = tuple(t.contiguous() for t in (x, y, dz, dsdth, rkappa, zk))
        h = _vp()
        _check(_bind().eb_matvec_create(ctypes.byref(h), bodies._h, *(_ptr(t) for t in self._keep), levmax, nterms,
                                        mode, _stream_ptr(stream)))
        self._h = h

    def apply(self, xx, with_parts=False):
        torch = _torch()
        n, k = self.bodies.n, self.bodies.nbody
        yy = torch.empty(2 * n, dtype=torch.float64, device=xx.device)
        stress = torch.empty(n, dtype=torch.complex128, device=xx.device) if with_parts else None
        za = torch.empty(k, dtype=torch.complex128, device=xx.device) if with_parts else None
        pk = torch.empty(k, dtype=torch.float64, device=xx.device) if with_parts else None
        _check(_bind().eb_matvec_apply(self._h, _ptr(xx.contiguous()), _ptr(yy), _p(stress), _p(za), _p(pk),
                                       _stream_ptr(self.stream)))
        return (yy, stress, za, pk) if with_parts else yy

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().eb_matvec_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Sweep:
    """Many independent simulations in one operator: simulation s owns bodies
    sim_off[s]..sim_off[s+1]-1 of `bodies`; apply() is MATVEC + FASMVP of all of them, each
    simulation's layer potential summed over its own points only (include/eb_spectral.h)."""

    def __init__(self, bodies, sim_off, x, y, dz, dsdth, rkappa, zk, stream=None):
        self.bodies = bodies
        self.stream = stream
        self.sim_off = np.ascontiguousarray(sim_off, np.int32)
        self.nsim = self.sim_off.size - 1
        self._keep = tuple(t.contiguous() for t in (x, y, dz, dsdth, rkappa, zk))
        h = _vp()
        _check(_bind().eb_sweep_create(ctypes.byref(h), bodies._h, self.nsim, self.sim_off.ctypes.data_as(_ip),
                                       *(_ptr(t) for t in self._keep), _stream_ptr(stream)))
        self._h = h
        uo = _bind().eb_sweep_unknown_offsets(h)
        self.unknown_offsets = np.array([uo[i] for i in range(self.nsim + 1)], np.int32)

    def apply(self, xx):
        yy = _torch().empty(2 * self.bodies.n, dtype=_torch().float64, device=xx.device)
        _check(_bind().eb_sweep_apply(self._h, _ptr(xx.contiguous()), _ptr(yy), _stream_ptr(self.stream)))
        return yy

    def set_mode(self, mode):
        """0 auto (segmented FMM when a simulation has more than 10240 points), 1 segmented FMM,
        2 each simulation's O(N^2) sum (eb_sweep_set_mode)."""
        _check(_bind().eb_sweep_set_mode(self._h, int(mode)))

    def parts(self, xx):
        torch = _torch()
        k = self.bodies.nbody
        za = torch.empty(k, dtype=torch.complex128, device=xx.device)
        pk = torch.empty(k, dtype=torch.float64, device=xx.device)
        _check(_bind().eb_sweep_parts(self._h, _ptr(xx.contiguous()), _ptr(za), _ptr(pk), _stream_ptr(self.stream)))
        return za, pk

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().eb_sweep_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fasmvp(nd, x, y, dz, dsdth, rkappa, zk, u):
    """FASMVP on host arrays with the reference's interface (stokesbubb.f:496-498).
    Returns (w, stress, za, pk)."""
    L = _bind()
    nd = np.ascontiguousarray(nd, np.int32)
    k = nd.size
    nbk = int(nd.sum())
    dp = lambda a: a.ctypes.data_as(_dp)  # noqa: E731
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    dzd = np.ascontiguousarray(dz, np.complex128).view(np.float64)
    ds = np.ascontiguousarray(dsdth, np.float64)
    rk = np.ascontiguousarray(rkappa, np.float64)
    zkd = np.ascontiguousarray(zk, np.complex128).view(np.float64)
    u = np.ascontiguousarray(u, np.float64)
    h = (2 * np.pi / nd).astype(np.float64)
    w = np.zeros(2 * nbk)
    za = np.zeros(2 * k)
    pk = np.zeros(k)
    stress = np.zeros(2 * nbk)
    dens = np.zeros(2 * nbk)
    rc = L.eb_fasmvp(nd.ctypes.data_as(_ip), ctypes.byref(_i32(nbk)), ctypes.byref(_i32(k)), dp(x), dp(y), dp(dzd),
                     dp(ds), dp(rk), dp(zkd), dp(u), dp(w), dp(h), dp(za), dp(pk), dp(stress), dp(dens))
    _check(rc)
    return w, stress.view(np.complex128), za.view(np.complex128), pk
