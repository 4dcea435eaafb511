"""Device-resident GMRES (ctypes layer over include/eb_gmres.h).

``GMRES(n, maxl)`` holds the Krylov workspace; ``solve_dense`` runs it on a dense device
matrix and ``stokes_solve`` is SOLVEI (stokesbubb.f:294-415) on a :class:`spectral.MatVec`.
``GMRESBatch`` runs many independent systems in lockstep; ``sweep_solve`` is SOLVEI of every
simulation of a :class:`spectral.Sweep` at once.
Vectors are float64 device tensors; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, _ptr, _stream_ptr, _torch, lib

GMRES_EXPORTS = ("eb_gmres_create", "eb_gmres_solve", "eb_gmres_destroy", "eb_dense_gmres_apply",
                 "eb_matvec_gmres_apply", "eb_matvec_parts", "eb_stokes_solve", "eb_gmres_batch_create",
                 "eb_gmres_batch_solve", "eb_gmres_batch_destroy", "eb_sweep_gmres_apply", "eb_sweep_solve")

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32


class _DenseOp(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("a", ctypes.c_void_p)]


_bound = False


def _bind():
    global _bound
    L = lib()
    if not _bound:
        L.eb_gmres_create.argtypes = [ctypes.POINTER(_vp), _i32, _i32, _vp]
        L.eb_gmres_solve.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_double, _i32, ctypes.POINTER(_i32),
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i32)]
        L.eb_gmres_destroy.argtypes = [_vp]
        L.eb_matvec_parts.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.eb_stokes_solve.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_i32),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i32), _vp]
        L.eb_gmres_batch_create.argtypes = [ctypes.POINTER(_vp), _i32, ctypes.POINTER(_i32), _i32, _vp]
        L.eb_gmres_batch_solve.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp]
        L.eb_gmres_batch_destroy.argtypes = [_vp]
        L.eb_sweep_solve.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        _bound = True
    return L


def _fn(name):
    return ctypes.cast(getattr(lib(), name), _vp)


class GMRES:
    """Workspace of DGMRES for n unknowns and up to maxl Krylov vectors (IGWK(1))."""

    def __init__(self, n, maxl=100, stream=None):
        self._torch = _torch()
        self.n = int(n)
        self.stream = stream
        h = _vp()
        _check(_bind().eb_gmres_create(ctypes.byref(h), self.n, int(maxl), _stream_ptr(stream)))
        self._h = h

    def _solve(self, fn, ctx, b, x, tol, nrmax):
        it, err, ierr = _i32(), ctypes.c_double(), _i32()
        _check(_bind().eb_gmres_solve(self._h, fn, ctx, _ptr(b), _ptr(x), float(tol), int(nrmax), ctypes.byref(it),
                                      ctypes.byref(err), ctypes.byref(ierr)))
        return it.value, err.value, ierr.value

    def solve_dense(self, a, b, x, tol=1e-10, nrmax=0):
        """a: (n, n) float64 device tensor; x (device) holds the initial guess and receives the
        solution.  Returns (iter, err, ierr)."""
        at = a.t().contiguous()  # column-major storage of a
        op = _DenseOp(self.n, at.data_ptr())
        return self._solve(_fn("eb_dense_gmres_apply"), ctypes.cast(ctypes.pointer(op), _vp), b.contiguous(), x, tol,
                           nrmax)

    def solve_matvec(self, mv, b, x, tol=1e-10, nrmax=-1):
        """The Stokes matvec `mv` (spectral.MatVec) as the operator."""
        return self._solve(_fn("eb_matvec_gmres_apply"), mv._h, b.contiguous(), x, tol, nrmax)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().eb_gmres_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stokes_solve(mv, gm, rhs):
    """SOLVEI on the device.  Returns (soln, za, pk, iter, err, ierr)."""
    torch = _torch()
    nb = mv.bodies.nbody
    soln = torch.empty_like(rhs)
    za = torch.empty(nb, dtype=torch.complex128, device=rhs.device)
    pk = torch.empty(nb, dtype=torch.float64, device=rhs.device)
    it, err, ierr = _i32(), ctypes.c_double(), _i32()
    _check(_bind().eb_stokes_solve(mv._h, gm._h, _ptr(rhs.contiguous()), _ptr(soln), _ptr(za), _ptr(pk),
                                   ctypes.byref(it), ctypes.byref(err), ctypes.byref(ierr), _stream_ptr(gm.stream)))
    return soln, za, pk, it.value, err.value, ierr.value


class GMRESBatch:
    """Workspace of the batched GMRES: system s owns entries off[s]..off[s+1]-1."""

    def __init__(self, off, maxl=100, stream=None):
        self._torch = _torch()
        self.off = np.ascontiguousarray(off, np.int32)
        self.nsys = self.off.size - 1
        self.stream = stream
        h = _vp()
        _check(_bind().eb_gmres_batch_create(ctypes.byref(h), self.nsys, self.off.ctypes.data_as(ctypes.POINTER(_i32)),
                                             int(maxl), _stream_ptr(stream)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().eb_gmres_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sweep_solve(sw, gb, rhs):
    """SOLVEI of every simulation of the sweep.  Returns (soln, za, pk, iter, err, ierr), the
    last three as numpy arrays with one entry per simulation."""
    torch = _torch()
    nb = sw.bodies.nbody
    soln = torch.empty_like(rhs)
    za = torch.empty(nb, dtype=torch.complex128, device=rhs.device)
    pk = torch.empty(nb, dtype=torch.float64, device=rhs.device)
    it = np.zeros(sw.nsim, np.int32)
    err = np.zeros(sw.nsim, np.float64)
    ierr = np.zeros(sw.nsim, np.int32)
    _check(_bind().eb_sweep_solve(sw._h, gb._h, _ptr(rhs.contiguous()), _ptr(soln), _ptr(za), _ptr(pk),
                                  it.ctypes.data, err.ctypes.data, ierr.ctypes.data, _stream_ptr(gb.stream)))
    return soln, za, pk, it, err, ierr
