"""DAPIF2's adaptive quadtree built on the GPU (ctypes layer over include/eb_atree.h).

``AdaptiveTree(x, y, napb)`` builds the tree of the reference's adaptive FMM
(dapif1.f:740-1680, dapif2.f:3-474) from float64 device tensors; ``tables()`` copies the
1-based tables to host numpy arrays under the reference's names.  Nothing here computes on
the CPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import EBError, _check, _ptr, _stream_ptr, _torch, lib

ATREE_EXPORTS = ("eb_atree_build", "eb_atree_build_bounded", "eb_atree_counts", "eb_atree_tables",
                 "eb_atree_destroy")

EB_ATREE_EPS_STOP, EB_ATREE_FEW_POINTS, EB_ATREE_TOO_DEEP = 1, 16, 32

_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int32)


class _View(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("xs", "ys", "geom", "xb", "yb", "nl", "ibox", "nbat", "npar", "nchi",
                                               "indb", "iat", "jat", "l1", "l2", "jl", "l3", "l4", "jl3", "jl4")]


_bound = False


def _bind():
    global _bound
    L = lib()
    if not _bound:
        L.eb_atree_build.argtypes = [ctypes.POINTER(_vp), ctypes.c_int32, _vp, _vp, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_double, _ip, _vp]
        L.eb_atree_counts.argtypes = [_vp, _ip]
        L.eb_atree_tables.argtypes = [_vp, ctypes.POINTER(_View)]
        L.eb_atree_destroy.argtypes = [_vp]
        _bound = True
    return L


_hip = None


def _copy_d2h(dst, src):
    """hipMemcpy of a device table into a host numpy array (the HIP runtime libeb.so uses)."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
    rc = _hip.hipMemcpy(dst.ctypes.data, src, dst.nbytes, 2)  # hipMemcpyDeviceToHost
    if rc != 0:
        raise EBError(18, f"hipMemcpy failed ({rc})")


class AdaptiveTree:
    """x, y: float64 device tensors.  napb: most points per childless box above the finest
    level; lmx: most levels (30 = DAPIF2 with MEX 300, TOL 1e-11); eps7: DAPIF2's EPS7."""

    def __init__(self, x, y, napb, lmx=30, eps7=1e-16, stream=None):
        self._torch = _torch()
        self._keep = (x.contiguous(), y.contiguous())
        self.stream = stream
        L = _bind()
        h = _vp()
        self.ierr = np.zeros(3, np.int32)
        rc = L.eb_atree_build(ctypes.byref(h), x.numel(), _ptr(self._keep[0]), _ptr(self._keep[1]), int(napb),
                              int(lmx), float(eps7), self.ierr.ctypes.data_as(_ip), _stream_ptr(stream))
        if rc in (EB_ATREE_EPS_STOP, EB_ATREE_FEW_POINTS, EB_ATREE_TOO_DEEP):
            raise EBError(rc, f"adaptive tree: DAPIF2 IERR = {self.ierr.tolist()}")
        _check(rc)
        self._h = h
        c = np.zeros(8, np.int32)
        _check(L.eb_atree_counts(h, c.ctypes.data_as(_ip)))
        self.nbox, self.nlev, self.len, self.len3, self.len4, self.maxp, self.jatlen, self.nat = (int(v) for v in c)

    def view(self):
        v = _View()
        _check(_bind().eb_atree_tables(self._h, ctypes.byref(v)))
        return v

    def tables(self):
        """Host copies of every table (1-based values, the reference's names)."""
        torch = self._torch
        torch.cuda.synchronize()
        v = self.view()
        nb = self.nbox
        sizes = dict(xs=(self.nat, np.float64), ys=(self.nat, np.float64), geom=(3, np.float64),
                     nl=(self.nlev + 2, np.int32), ibox=(self.nat, np.int32), xb=(nb, np.float64),
                     yb=(nb, np.float64), nbat=(nb, np.int32), npar=(nb, np.int32), nchi=(nb, np.int32),
                     indb=(nb, np.int32), iat=(nb, np.int32), jat=(self.jatlen, np.int32), l1=(nb + 1, np.int32),
                     l2=(nb + 1, np.int32), jl=(self.len, np.int32), l3=(nb + 1, np.int32), l4=(nb + 1, np.int32),
                     jl3=(self.len3, np.int32), jl4=(self.len4, np.int32))
        out = dict(nbox=nb, nlev=self.nlev, len=self.len, len3=self.len3, len4=self.len4, maxp=self.maxp)
        for k, (n, dt) in sizes.items():
            a = np.zeros(n, dt)
            if n:
                _copy_d2h(a, getattr(v, k))
            out[k] = a
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().eb_atree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
