"""The drop evolution on the device (ctypes layer over include/eb_sim.h): MANYDROP's initial
shapes, redistribution to equal arclength, RK2 time loop with the step law and point
doubling, and the bubbin / geometry.data formats (electric.f:1-272, 1360-1423).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, _stream_ptr, _torch, lib

SIM_EXPORTS = ("eb_sim_read_bubbin", "eb_sim_create", "eb_sim_step", "eb_sim_get_state", "eb_sim_copy_state",
               "eb_sim_dump", "eb_sim_restore", "eb_sim_destroy")

KMAX = 150
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_f64 = ctypes.c_double


class Params(ctypes.Structure):
    """READINI's run description (eb_sim_params)."""
    _fields_ = [("k", _i32), ("nd", _i32 * KMAX), ("dt", _f64), ("tfinal", _f64), ("nplot", _i32),
                ("cap", _f64), ("bondx", _f64), ("bondy", _f64), ("e0", _f64), ("ai", _f64 * KMAX),
                ("bi", _f64 * KMAX), ("thetax", _f64 * KMAX), ("cx", _f64 * KMAX), ("cy", _f64 * KMAX),
                ("ncyc", _i32 * KMAX)]

    @classmethod
    def from_dict(cls, d):
        p = cls()
        p.k = int(d["k"])
        for key in ("nd", "ncyc"):
            arr = getattr(p, key)
            for i, v in enumerate(d[key]):
                arr[i] = int(v)
        for key in ("ai", "bi", "thetax", "cx", "cy"):
            arr = getattr(p, key)
            for i, v in enumerate(d[key]):
                arr[i] = float(v)
        p.dt, p.tfinal, p.nplot = float(d["dt"]), float(d["tfinal"]), int(d.get("nplot", 10))
        p.cap, p.e0 = float(d.get("cap", 0.0)), float(d.get("e0", 0.0))
        p.bondx, p.bondy = float(d.get("bondx", 0.0)), float(d.get("bondy", 0.0))
        return p

    def to_dict(self):
        k = self.k
        d = {"k": k, "dt": self.dt, "tfinal": self.tfinal, "nplot": self.nplot, "cap": self.cap,
             "bondx": self.bondx, "bondy": self.bondy, "e0": self.e0}
        for key in ("nd", "ncyc", "ai", "bi", "thetax", "cx", "cy"):
            d[key] = list(getattr(self, key))[:k]
        return d


class _State(ctypes.Structure):
    _fields_ = [("time", _f64), ("dt", _f64), ("area_total", _f64), ("area_total0", _f64), ("area_error", _f64),
                ("k", _i32), ("nbk", _i32), ("steps", _i32), ("doublings", _i32), ("nd", ctypes.POINTER(_i32)),
                ("area", ctypes.POINTER(_f64)), ("arcl", ctypes.POINTER(_f64)), ("arcl0", ctypes.POINTER(_f64)),
                ("ybar", _vp)]


_bound = False


def _bind():
    global _bound
    L = lib()
    if not _bound:
        L.eb_sim_read_bubbin.argtypes = [ctypes.c_char_p, ctypes.POINTER(Params)]
        L.eb_sim_create.argtypes = [ctypes.POINTER(_vp), ctypes.POINTER(Params), _vp]
        L.eb_sim_step.argtypes = [_vp, ctypes.POINTER(_i32)]
        L.eb_sim_get_state.argtypes = [_vp, ctypes.POINTER(_State)]
        L.eb_sim_copy_state.argtypes = [_vp, _vp]
        L.eb_sim_dump.argtypes = [_vp, ctypes.c_char_p]
        L.eb_sim_restore.argtypes = [ctypes.POINTER(_vp), ctypes.POINTER(Params), ctypes.c_char_p, _vp]
        L.eb_sim_destroy.argtypes = [_vp]
        _bound = True
    return L


def read_bubbin(path):
    """READINI: the first run description of a bubbin file (host code only)."""
    p = Params()
    _check(_bind().eb_sim_read_bubbin(str(path).encode(), ctypes.byref(p)))
    return p


class Simulation:
    """MANYDROP on the device from a run description (Params, or a dict of its fields)."""

    def __init__(self, params, stream=None, restore=None):
        self._torch = _torch()
        self.params = params if isinstance(params, Params) else Params.from_dict(params)
        self.stream = stream
        h = _vp()
        if restore is None:
            _check(_bind().eb_sim_create(ctypes.byref(h), ctypes.byref(self.params), _stream_ptr(stream)))
        else:
            _check(_bind().eb_sim_restore(ctypes.byref(h), ctypes.byref(self.params), str(restore).encode(),
                                          _stream_ptr(stream)))
        self._h = h

    def step(self):
        """One pass of the time loop; returns True once TFINAL is reached."""
        done = _i32()
        _check(_bind().eb_sim_step(self._h, ctypes.byref(done)))
        return bool(done.value)

    def state(self):
        """time, dt, areas, nd and a host copy of the interleaved (x, y) state."""
        s = _State()
        _check(_bind().eb_sim_get_state(self._h, ctypes.byref(s)))
        k = s.k
        ybar = np.empty(2 * s.nbk)
        _check(_bind().eb_sim_copy_state(self._h, ybar.ctypes.data_as(_vp)))
        return {"time": s.time, "dt": s.dt, "area_total": s.area_total, "area_total0": s.area_total0,
                "area_error": s.area_error, "steps": s.steps, "doublings": s.doublings,
                "nd": np.array(s.nd[:k], np.int32), "area": np.array(s.area[:k]), "arcl": np.array(s.arcl[:k]),
                "arcl0": np.array(s.arcl0[:k]), "ybar": ybar}

    def dump(self, path):
        _check(_bind().eb_sim_dump(self._h, str(path).encode()))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().eb_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
