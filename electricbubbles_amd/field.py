"""F, the right-hand side of the drop evolution, on the device (ctypes layer over
include/eb_field.h): the Laplace solve, the Dirichlet-to-Neumann map, SETVELF, SOLVEI and
INTERFACENEW around the spectral tail (electric.f:968-1109), with no gravity and no wall.
Device tensors in and out (float64; complex128 for the Stokes density); nothing here
computes on the CPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, _ptr, _stream_ptr, _torch, lib

FIELD_EXPORTS = ("eb_field_create", "eb_field_destroy", "eb_field_set_mode", "eb_field_rhs", "eb_field_set_geometry",
                 "eb_field_laplace_apply", "eb_field_laplace_solve", "eb_field_dirich_neu", "eb_field_setvelf",
                 "eb_field_stokes_solve", "eb_field_interface", "eb_field_get_geometry", "eb_field_iterations", "eb_sweep_field_create",
                 "eb_sweep_field_rhs", "eb_sweep_field_iterations", "eb_sweep_field_destroy",
                 "eb_sweep_field_rk2_graph", "eb_field_rk2_graph", "eb_field_rk4_graph", "eb_sweep_field_rk4_graph",
                 "eb_sweep_field_set_mode")

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_bound = False


def _bind():
    global _bound
    L = lib()
    if not _bound:
        L.eb_field_create.argtypes = [ctypes.POINTER(_vp), _vp, ctypes.c_double, ctypes.c_double, _vp]
        L.eb_field_destroy.argtypes = [_vp]
        L.eb_field_set_mode.argtypes = [_vp, _i32]
        L.eb_field_rhs.argtypes = [_vp, _vp, _vp]
        L.eb_field_set_geometry.argtypes = [_vp, _vp]
        L.eb_field_laplace_apply.argtypes = [_vp, _vp, _vp]
        L.eb_field_laplace_solve.argtypes = [_vp, _vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]
        L.eb_field_dirich_neu.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.eb_field_setvelf.argtypes = [_vp, _vp, _vp]
        L.eb_field_stokes_solve.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]
        L.eb_field_interface.argtypes = [_vp, _vp, _vp]
        L.eb_field_iterations.argtypes = [_vp] + [ctypes.POINTER(_i32)] * 4
        L.eb_sweep_field_create.argtypes = [ctypes.POINTER(_vp), _vp, _i32, _vp, _vp, _vp, _vp]
        L.eb_sweep_field_rhs.argtypes = [_vp, _vp, _vp]
        L.eb_sweep_field_iterations.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.eb_sweep_field_destroy.argtypes = [_vp]
        L.eb_sweep_field_rk2_graph.argtypes = [_vp, _vp, ctypes.c_double, _i32, ctypes.POINTER(_i32)]
        L.eb_field_rk2_graph.argtypes = [_vp, _vp, ctypes.c_double, _i32, ctypes.POINTER(_i32)]
        L.eb_field_rk4_graph.argtypes = [_vp, _vp, ctypes.c_double, _i32, ctypes.POINTER(_i32)]
        L.eb_sweep_field_rk4_graph.argtypes = [_vp, _vp, ctypes.c_double, _i32, ctypes.POINTER(_i32)]
        L.eb_sweep_field_set_mode.argtypes = [_vp, _i32]
        _bound = True
    return L


class Field:
    """F of one simulation whose bodies are `bodies` (spectral.Bodies); cap and e0 are the
    reference's CAP and E0 (COMMON /flowinfo/)."""

    def __init__(self, bodies, cap=0.0, e0=0.0, stream=None):
        self._torch = _torch()
        self.bodies = bodies
        self.stream = stream
        h = _vp()
        _check(_bind().eb_field_create(ctypes.byref(h), bodies._h, float(cap), float(e0), _stream_ptr(stream)))
        self._h = h

    def set_mode(self, mode):
        """Evaluation mode of every layer potential (MODE_AUTO, MODE_FMM or MODE_DIRECT); takes
        effect at the next set_geometry / rhs."""
        _check(_bind().eb_field_set_mode(self._h, int(mode)))

    def _f64(self, n, like):
        return self._torch.empty(n, dtype=self._torch.float64, device=like.device)

    def rhs(self, rku):
        out = self._f64(rku.numel(), rku)
        _check(_bind().eb_field_rhs(self._h, _ptr(rku.contiguous()), _ptr(out)))
        return out

    def set_geometry(self, rku):
        self._rku = rku.contiguous()
        _check(_bind().eb_field_set_geometry(self._h, _ptr(self._rku)))

    def laplace_apply(self, u):
        w = self._f64(self.bodies.n, u)
        _check(_bind().eb_field_laplace_apply(self._h, _ptr(u.contiguous()), _ptr(w)))
        return w

    def laplace_solve(self, like):
        u = self._f64(self.bodies.n, like)
        it, ie = _i32(), _i32()
        _check(_bind().eb_field_laplace_solve(self._h, _ptr(u), ctypes.byref(it), ctypes.byref(ie)))
        return u, it.value, ie.value

    def dirich_neu(self, ulap):
        n = self.bodies.n
        phi, psi, dphi = self._f64(n, ulap), self._f64(n, ulap), self._f64(n, ulap)
        _check(_bind().eb_field_dirich_neu(self._h, _ptr(ulap.contiguous()), _ptr(phi), _ptr(psi), _ptr(dphi)))
        return phi, psi, dphi

    def setvelf(self, dphi_dn):
        rhs = self._f64(2 * self.bodies.n, dphi_dn)
        _check(_bind().eb_field_setvelf(self._h, _ptr(dphi_dn.contiguous()), _ptr(rhs)))
        return rhs

    def stokes_solve(self, rhs):
        torch = self._torch
        k = self.bodies.nbody
        u = torch.empty(self.bodies.n, dtype=torch.complex128, device=rhs.device)
        za = torch.empty(k, dtype=torch.complex128, device=rhs.device)
        pk = torch.empty(k, dtype=torch.float64, device=rhs.device)
        it, ie = _i32(), _i32()
        _check(_bind().eb_field_stokes_solve(self._h, _ptr(rhs.contiguous()), _ptr(u), _ptr(za), _ptr(pk),
                                             ctypes.byref(it), ctypes.byref(ie)))
        return u, za, pk, it.value, ie.value

    def interface(self, u):
        un = self._f64(self.bodies.n, u)
        _check(_bind().eb_field_interface(self._h, _ptr(u.contiguous()), _ptr(un)))
        return un

    def iterations(self):
        """(Laplace iterations, Laplace IERR, Stokes iterations, Stokes IERR) of the last solves."""
        v = [_i32() for _ in range(4)]
        _check(_bind().eb_field_iterations(self._h, *(ctypes.byref(a) for a in v)))
        return tuple(a.value for a in v)

    def rk2_graph(self, ybar, dt, budget=8):
        """One Heun step in place on ybar (a contiguous device tensor), captured in a hipGraph on
        the first call and replayed after, with device-controlled solves (eb_field_rk2_graph).
        The field must have been created with a non-default stream; returns True when the step
        was redone with the host-controlled solves (a solve needed more than `budget` steps)."""
        assert ybar.is_contiguous() and ybar.dtype == self._torch.float64
        red = _i32()
        _check(_bind().eb_field_rk2_graph(self._h, _ptr(ybar), float(dt), int(budget), ctypes.byref(red)))
        return bool(red.value)

    def rk4_graph(self, ybar, dt, budget=8):
        """One classic RK4 step in place on ybar, captured and replayed as rk2_graph
        (eb_field_rk4_graph); returns True when the step was redone with host-controlled solves."""
        assert ybar.is_contiguous() and ybar.dtype == self._torch.float64
        red = _i32()
        _check(_bind().eb_field_rk4_graph(self._h, _ptr(ybar), float(dt), int(budget), ctypes.byref(red)))
        return bool(red.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().eb_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SweepField:
    """F of many independent simulations at once: simulation s owns bodies
    sim_off[s]..sim_off[s+1]-1 of `bodies` and has its own CAP and E0."""

    def __init__(self, bodies, sim_off, cap, e0, stream=None):
        self._torch = _torch()
        self.bodies = bodies
        self.stream = stream
        self.sim_off = np.ascontiguousarray(sim_off, np.int32)
        self.nsim = self.sim_off.size - 1
        cap = np.ascontiguousarray(np.broadcast_to(np.asarray(cap, np.float64), (self.nsim,)))
        e0 = np.ascontiguousarray(np.broadcast_to(np.asarray(e0, np.float64), (self.nsim,)))
        h = _vp()
        _check(_bind().eb_sweep_field_create(ctypes.byref(h), bodies._h, self.nsim, self.sim_off.ctypes.data,
                                             cap.ctypes.data, e0.ctypes.data, _stream_ptr(stream)))
        self._h = h

    def rhs(self, rku):
        out = self._torch.empty(rku.numel(), dtype=self._torch.float64, device=rku.device)
        _check(_bind().eb_sweep_field_rhs(self._h, _ptr(rku.contiguous()), _ptr(out)))
        return out

    def set_mode(self, mode):
        """0 auto (segmented FMM plans when a simulation has more than 10240 points), 1 segmented
        FMM, 2 each simulation's O(N^2) sums (eb_sweep_field_set_mode)."""
        _check(_bind().eb_sweep_field_set_mode(self._h, int(mode)))

    def iterations(self):
        """(Laplace iterations, Laplace IERR, Stokes iterations, Stokes IERR) of the last call."""
        a = [np.zeros(self.nsim, np.int32) for _ in range(4)]
        _check(_bind().eb_sweep_field_iterations(self._h, *(v.ctypes.data for v in a)))
        return tuple(a)

    def rk2_graph(self, ybar, dt, budget=8):
        """One Heun step of every simulation in place on ybar (a contiguous device tensor),
        captured in a hipGraph on the first call and replayed after (eb_sweep_field_rk2_graph).
        The field must have been created with a non-default stream; returns True when the step
        was redone with the host-controlled solves (a system needed more than `budget` steps)."""
        assert ybar.is_contiguous() and ybar.dtype == self._torch.float64
        red = _i32()
        _check(_bind().eb_sweep_field_rk2_graph(self._h, _ptr(ybar), float(dt), int(budget), ctypes.byref(red)))
        return bool(red.value)

    def rk4_graph(self, ybar, dt, budget=8):
        """One classic RK4 step of every simulation in place on ybar, captured and replayed as
        rk2_graph (eb_sweep_field_rk4_graph)."""
        assert ybar.is_contiguous() and ybar.dtype == self._torch.float64
        red = _i32()
        _check(_bind().eb_sweep_field_rk4_graph(self._h, _ptr(ybar), float(dt), int(budget), ctypes.byref(red)))
        return bool(red.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().eb_sweep_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
