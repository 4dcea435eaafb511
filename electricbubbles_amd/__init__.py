"""electricbubbles_amd -- MI355X (gfx950) implementation of ElectricBubbles' Stokes-matvec path.

The native library ``libeb.so`` (HIP kernels + C-ABI, ``include/eb_cadap.h``) does all the
work; this module is a thin ctypes layer over that C-ABI:

* :func:`cadapdr`, :func:`cadapmod`, :func:`cadap` -- the reference's CADAPDR / CADAPMOD /
  CADAP interface (cadapdrop.f:3-5) on host arrays, returning ``(gradw, inform, ier)``;
* :class:`FMMPlan` -- the device-resident plan/apply API on torch tensors (device memory and
  streams only; no torch types cross the C-ABI);
* :func:`tables` / :func:`lists` -- the reference's per-level binning tables and
  interaction lists computed on the GPU (tree parity);
* spectral / RK / FASMVP-epilogue wrappers (see :mod:`electricbubbles_amd.spectral`);
* DAPIF2's adaptive quadtree on the GPU (see :mod:`electricbubbles_amd.atree`);
* the device-resident GMRES solve of the Stokes equation (see :mod:`electricbubbles_amd.gmres`).

There is no CPU fallback: if ``libeb.so`` is missing, or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libeb.so")

EB_OK = 0
EB_ERR_TERMS = 4
EB_ERR_WORKSPACE = 8
EB_ERR_UNSUPPORTED = 16
EB_ERR_ARG = 17
EB_ERR_HIP = 18
EB_ERR_NOMEM = 19

VARIANT_DR, VARIANT_MOD, VARIANT_PLUS, VARIANT_HILBERT = 0, 1, 2, 3
MODE_AUTO, MODE_FMM, MODE_DIRECT = 0, 1, 2

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_ip = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)

# exported symbols of include/eb_cadap.h and include/eb_spectral.h
EXPORTS = (
    "eb_last_error", "eb_version", "eb_cadapdr", "eb_cadapmod", "eb_cadap", "eb_set_mode", "eb_get_mode",
    "eb_fmm_plan_create", "eb_fmm_plan_rebuild", "eb_fmm_apply", "eb_fmm_plan_destroy", "eb_fmm_plan_info",
    "eb_fmm_plan_m2l_split", "eb_afmm_m2l_split",
    "eb_fmm_plan_ntbox", "eb_fmm_enable_timing", "eb_fmm_stage_ms", "eb_fmm_set_p2p_tile", "eb_cadap_tables",
    "eb_cadap_lists", "eb_direct_apply_noalloc", "eb_direct_subset", "eb_dapif2", "eb_afmm_create",
    "eb_afmm_apply", "eb_afmm_info", "eb_afmm_enable_timing", "eb_afmm_stage_ms", "eb_afmm_check",
    "eb_afmm_destroy", "eb_fmm_set_guard", "eb_fmm_apply_transpose", "eb_fmm_plan_create_segmented",
)

_lib = None


class EBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"eb status {code}: {msg}")
        self.code = code


def lib():
    """Load libeb.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIBPATH)
    L.eb_last_error.restype = ctypes.c_char_p
    dropin_args = [_ip, _ip, _ip, _dp, _dp, _dp, _ip, _ip, _ip, _dp, _dp, _ip, _ip, _dp, _ip,
                   _dp, _dp, _ip, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp]
    for name in ("eb_cadapdr", "eb_cadapmod", "eb_cadap"):
        f = getattr(L, name)
        f.argtypes = dropin_args
        f.restype = _i32
    L.eb_dapif2.argtypes = [_ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                            _vp]
    L.eb_dapif2.restype = _i32
    L.eb_set_mode.argtypes = [ctypes.c_int]
    L.eb_get_mode.restype = ctypes.c_int
    L.eb_fmm_plan_create.argtypes = [ctypes.POINTER(_vp), _i32, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _vp]
    L.eb_fmm_apply.argtypes = [_vp, _vp, _vp, _vp]
    L.eb_fmm_apply_transpose.argtypes = [_vp, _vp, _vp, _vp]
    L.eb_fmm_plan_create_segmented.argtypes = [ctypes.POINTER(_vp), _i32, _vp, _vp, _vp, _i32,
                                               ctypes.POINTER(_i32), _i32, _i32, _i32, _vp]
    L.eb_fmm_plan_rebuild.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.eb_fmm_plan_destroy.argtypes = [_vp]
    L.eb_fmm_plan_info.argtypes = [_vp, _dp, _i64p]
    L.eb_fmm_plan_m2l_split.argtypes = [_vp, _i64p]
    L.eb_afmm_m2l_split.argtypes = [_vp, _i64p]
    L.eb_fmm_plan_ntbox.argtypes = [_vp, _ip]
    L.eb_fmm_enable_timing.argtypes = [_vp, _i32]
    L.eb_fmm_set_p2p_tile.argtypes = [_vp, _i32]
    L.eb_fmm_stage_ms.argtypes = [_vp, _dp]
    L.eb_cadap_tables.argtypes = [_i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _ip, _ip, _vp]
    L.eb_cadap_lists.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp]
    L.eb_direct_apply_noalloc.argtypes = [_i32, _vp, _vp, _vp, _vp, _i32, _vp, _vp]
    L.eb_direct_subset.argtypes = [_i32, _vp, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _vp]
    L.eb_afmm_create.argtypes = [ctypes.POINTER(_vp), _i32, _vp, _vp, _vp, _i32, _i32, _i32, _vp]
    L.eb_afmm_apply.argtypes = [_vp, _vp, _vp, _vp]
    L.eb_afmm_info.argtypes = [_vp, _dp, _i64p]
    L.eb_afmm_enable_timing.argtypes = [_vp, _i32]
    L.eb_afmm_stage_ms.argtypes = [_vp, _dp]
    L.eb_afmm_check.argtypes = [_vp, _i64p]
    L.eb_afmm_destroy.argtypes = [_vp]
    L.eb_fmm_set_guard.argtypes = [_vp, _vp]
    _lib = L
    return L


def last_error() -> str:
    return lib().eb_last_error().decode(errors="replace")


def _check(rc):
    if rc != EB_OK:
        raise EBError(rc, last_error())


def nterms_for(tol: float) -> int:
    """NTERMS = -NINT(log(TOL)/log(2)) (cadapdrop.f:126)."""
    v = np.log(tol) / np.log(2.0)
    return -int(np.sign(v) * np.floor(abs(v) + 0.5))


# ----------------------------------------------------------------------------------------
# drop-in interface on host arrays
# ----------------------------------------------------------------------------------------

def _as_c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _call_dropin(fn, iflag1, iflag2, levmax, xat, yat, dzdt, nd, dens, h, tol, leniw, lenw, k0=1):
    nd = _as_c(nd, np.int32)
    kbod = k0 + nd.size - 1
    n = int(nd.sum())
    x = _as_c(xat, np.float64)
    y = _as_c(yat, np.float64)
    dz = _as_c(dzdt, np.complex128).view(np.float64)
    de = _as_c(dens, np.complex128).view(np.float64)
    hh = _as_c(h, np.float64)
    gw = np.zeros(2 * n)
    inform = np.zeros(6, np.int32)
    ier = np.zeros(5, np.int32)
    dummy_d = np.zeros(2)
    dummy_i = np.zeros(1, np.int32)
    I = lambda v: ctypes.byref(ctypes.c_int32(int(v)))  # noqa: E731
    D = lambda v: ctypes.byref(ctypes.c_double(float(v)))  # noqa: E731
    dp = lambda a: a.ctypes.data_as(_dp)  # noqa: E731
    ip = lambda a: a.ctypes.data_as(_ip)  # noqa: E731
    rc = fn(I(iflag1), I(iflag2), I(levmax), dp(x), dp(y), dp(dz), I(k0), I(kbod), ip(nd), dp(de), dp(hh),
            ip(dummy_i), I(leniw), dp(dummy_d), I(lenw), dp(dummy_d), dp(dummy_d), I(0), dp(gw),
            dp(dummy_d), dp(dummy_d), D(tol), ip(inform), ip(ier), D(0.0), D(0.0))
    return rc, gw.view(np.complex128), inform, ier


def _dropin(name, levmax, xat, yat, dzdt, nd, dens, h, tol, iflag1=0, iflag2=0, leniw=None, lenw=None):
    """Shared body of cadapdr / cadapmod / cadap.

    LENIW / LENW default to the workspaces FASMVP passes (stokesbubb.f:503, 564-566).
    Returns (gradw, inform[:2], ier) exactly like the Fortran routine; ier[0] != 0 is an error
    (the reference's callers STOP on it, stokesbubb.f:573-580)."""
    if leniw is None:
        leniw = 1000000
    if lenw is None:
        lenw = 1000000
    rc, gw, inform, ier = _call_dropin(getattr(lib(), name), iflag1, iflag2, levmax, xat, yat, dzdt, nd,
                                       dens, h, tol, leniw, lenw)
    if rc >= EB_ERR_UNSUPPORTED:
        raise EBError(rc, last_error())
    return gw, inform[:2].copy(), ier.copy()


def cadapdr(levmax, xat, yat, dzdt, nd, dens, h, tol, **kw):
    """CADAPDR (cadapdrop.f:3): GW = phi - z conj(phi') - conj(psi) at the sources."""
    return _dropin("eb_cadapdr", levmax, xat, yat, dzdt, nd, dens, h, tol, **kw)


def cadapmod(levmax, xat, yat, dzdt, nd, dens, h, tol, **kw):
    """CADAPMOD (cadapmod.f:3): GW = phi + z conj(phi') + conj(psi) at the sources."""
    return _dropin("eb_cadapmod", levmax, xat, yat, dzdt, nd, dens, h, tol, **kw)


def cadap(levmax, xat, yat, dzdt, nd, dens, h, tol, **kw):
    """CADAP (cadap.f:3): as CADAPMOD with +conj(dens) dz in QB."""
    return _dropin("eb_cadap", levmax, xat, yat, dzdt, nd, dens, h, tol, **kw)


def dapif2(xa, ya, qa, napb=30, tol=1e-11, mex=300, eps7=1e-16, nsp=None, iflag7=3):
    """DAPIF2 (dapif1.f:458) with IFLAG7 = 3 on host arrays: FIELD = the Hilbert-matrix product
    sum_{j != i} QA(j) / (z_i - z_j).  Returns (field, ierr[0:4], inform[0:4]) as the Fortran routine
    leaves them (NSP defaults to 4,000,000 + 400 NAT words; WKSP itself is not used)."""
    x = _as_c(xa, np.float64)
    y = _as_c(ya, np.float64)
    q = _as_c(qa, np.complex128).view(np.float64)
    nat = x.size
    if nsp is None:
        nsp = 4_000_000 + 400 * nat
    field = np.zeros(2 * nat)
    ierr = np.zeros(8, np.int32)
    inform = np.zeros(8, np.int32)
    iout = np.zeros(2, np.int32)
    dummy = np.zeros(2)
    I = lambda v: ctypes.byref(ctypes.c_int32(int(v)))  # noqa: E731
    D = lambda v: ctypes.byref(ctypes.c_double(float(v)))  # noqa: E731
    dp = lambda a: a.ctypes.data_as(_dp)  # noqa: E731
    rc = lib().eb_dapif2(iout.ctypes.data_as(_ip), I(iflag7), I(nat), I(napb), I(2), I(mex), ierr.ctypes.data_as(_ip),
                         inform.ctypes.data_as(_ip), D(tol), D(eps7), dp(x), dp(y), dp(q), dp(dummy), dp(field),
                         dp(dummy), I(nsp), None)
    if rc >= EB_ERR_UNSUPPORTED and ierr[0] < 0:
        raise EBError(rc, last_error())
    return field.view(np.complex128), ierr[:4].copy(), inform[:4].copy()


def set_mode(mode: int) -> None:
    """Evaluation mode of the drop-in layer (CADAP*, FASMVP, DAPIF2 entry points)."""
    _check(lib().eb_set_mode(mode))


def get_mode() -> int:
    return int(lib().eb_get_mode())


# ----------------------------------------------------------------------------------------
# device-resident plan (torch tensors are used only as device memory)
# ----------------------------------------------------------------------------------------

def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("electricbubbles_amd: no GPU visible (torch.cuda.is_available() is False)")
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class FMMPlan:
    """Binned geometry for repeated layer-potential evaluations (one GMRES solve).

    x, y: float64 device tensors (n,); wdz: complex128 device tensor (n,) = DZDT * H.
    apply(dens) returns GW (complex128, n) on the same stream."""

    def __init__(self, x, y, wdz, levmax, nterms, variant=VARIANT_DR, mode=MODE_FMM, stream=None):
        torch = _torch()
        self._torch = torch
        self.n = int(x.numel())
        assert x.dtype == torch.float64 and y.dtype == torch.float64 and wdz.dtype == torch.complex128
        assert x.is_cuda and y.is_cuda and wdz.is_cuda and y.numel() == self.n and wdz.numel() == self.n
        self._keep = (x.contiguous(), y.contiguous(), wdz.contiguous())
        self.levmax, self.nterms, self.variant, self.mode = levmax, nterms, variant, mode
        self.stream = stream
        h = ctypes.c_void_p()
        _check(lib().eb_fmm_plan_create(ctypes.byref(h), self.n, _ptr(self._keep[0]), _ptr(self._keep[1]),
                                        _ptr(self._keep[2]), levmax, nterms, variant, mode,
                                        _stream_ptr(stream)))
        self._h = h

    @classmethod
    def segmented(cls, x, y, wdz, seg_off, levmax, nterms, variant=VARIANT_DR, stream=None):
        """A plan over independent segments (eb_fmm_plan_create_segmented): segment s is points
        seg_off[s] .. seg_off[s+1]-1, binned on its own (own bounding box and scale) with levmax
        levels below its box, and apply() sums within each segment only."""
        torch = _torch()
        self = cls.__new__(cls)
        self._torch = torch
        self.n = int(x.numel())
        assert x.dtype == torch.float64 and y.dtype == torch.float64 and wdz.dtype == torch.complex128
        assert x.is_cuda and y.is_cuda and wdz.is_cuda and y.numel() == self.n and wdz.numel() == self.n
        self._keep = (x.contiguous(), y.contiguous(), wdz.contiguous())
        off = np.ascontiguousarray(seg_off, np.int32)
        self.nseg = off.size - 1
        self.levmax, self.nterms, self.variant, self.mode = levmax, nterms, variant, MODE_FMM
        self.stream = stream
        h = ctypes.c_void_p()
        _check(lib().eb_fmm_plan_create_segmented(ctypes.byref(h), self.n, _ptr(self._keep[0]), _ptr(self._keep[1]),
                                                  _ptr(self._keep[2]), self.nseg,
                                                  off.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), levmax, nterms,
                                                  variant, _stream_ptr(stream)))
        self._h = h
        return self

    def _enter(self):
        """The plan's own stream when it differs from torch's current one (made to wait for the
        current stream's work first), else None."""
        cur = self._torch.cuda.current_stream()
        s = self.stream
        if s is None or s == cur:
            return None
        s.wait_stream(cur)
        return s

    def _leave(self, s, *tensors):
        """Order the current stream after the plan's stream and keep the tensors' memory
        from being reused before the plan's stream is done with it."""
        if s is None:
            return
        self._torch.cuda.current_stream().wait_stream(s)
        for t in tensors:
            t.record_stream(s)

    def rebuild(self, x, y, wdz):
        """Re-bin moved points (same n) in place."""
        self._keep = (x.contiguous(), y.contiguous(), wdz.contiguous())
        s = self._enter()
        _check(lib().eb_fmm_plan_rebuild(self._h, _ptr(self._keep[0]), _ptr(self._keep[1]),
                                         _ptr(self._keep[2]), _stream_ptr(self.stream)))
        self._leave(s, *self._keep)

    def apply(self, dens, out=None):
        torch = self._torch
        assert dens.dtype == torch.complex128 and dens.numel() == self.n and dens.is_cuda
        dens = dens.contiguous()
        if out is None:
            out = torch.empty(self.n, dtype=torch.complex128, device=dens.device)
        s = self._enter()
        _check(lib().eb_fmm_apply(self._h, _ptr(dens), _ptr(out), _stream_ptr(self.stream)))
        self._leave(s, dens, out)
        return out

    def apply_transpose(self, lam, out=None):
        """mu = A^T lam for the real-linear map A: dens -> apply(dens), under the real inner
        product Re sum conj(a) b: the same sums with sources and targets swapped (one set of
        the plan's passes on transposed charges)."""
        torch = self._torch
        assert lam.dtype == torch.complex128 and lam.numel() == self.n and lam.is_cuda
        lam = lam.contiguous()
        if out is None:
            out = torch.empty(self.n, dtype=torch.complex128, device=lam.device)
        s = self._enter()
        _check(lib().eb_fmm_apply_transpose(self._h, _ptr(lam), _ptr(out), _stream_ptr(self.stream)))
        self._leave(s, lam, out)
        return out

    def info(self):
        r = np.zeros(3)
        c = np.zeros(8, np.int64)
        _check(lib().eb_fmm_plan_info(self._h, r.ctypes.data_as(_dp), c.ctypes.data_as(_i64p)))
        keys = ("levmax", "nterms", "kmax", "near_pairs", "m2l_pairs", "boxes", "leaves", "max_leaf")
        d = dict(zip(keys, (int(v) for v in c)))
        d.update(rscal=float(r[0]), rminx=float(r[1]), rminy=float(r[2]))
        return d

    def m2l_split(self):
        """M2L pairs by source group: (near pairs at all p input terms, far pairs, far input terms)."""
        c = np.zeros(3, np.int64)
        _check(lib().eb_fmm_plan_m2l_split(self._h, c.ctypes.data_as(_i64p)))
        return {"m2l_near": int(c[0]), "m2l_far": int(c[1]), "m2l_far_terms": int(c[2])}

    def ntbox(self):
        a = np.zeros(self.levmax + 1, np.int32)
        _check(lib().eb_fmm_plan_ntbox(self._h, a.ctypes.data_as(_ip)))
        return a

    def enable_timing(self, on=True):
        _check(lib().eb_fmm_enable_timing(self._h, 1 if on else 0))

    def set_p2p_tile(self, tmax):
        """Near-field work items of at most tmax targets (16..256; default 42)."""
        _check(lib().eb_fmm_set_p2p_tile(self._h, int(tmax)))
        self.p2p_tile = int(tmax)

    def stage_ms(self):
        a = np.zeros(7)
        _check(lib().eb_fmm_stage_ms(self._h, a.ctypes.data_as(_dp)))
        return dict(zip(("charges", "p2m", "m2m", "m2l", "l2l", "l2p", "p2p"), a.tolist()))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().eb_fmm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AFMMPlan(FMMPlan):
    """The same sums on the adaptive tree of DAPIF2 (eb_afmm_*): childless boxes of at most
    napb points (40 by default) in place of a uniform grid.  x, y, wdz as FMMPlan; apply(dens)
    returns GW.  Creating the plan synchronises (tree build and host-side lists)."""

    def __init__(self, x, y, wdz, nterms, napb=40, variant=VARIANT_DR, stream=None):
        torch = _torch()
        self._torch = torch
        self.n = int(x.numel())
        assert x.dtype == torch.float64 and y.dtype == torch.float64 and wdz.dtype == torch.complex128
        assert x.is_cuda and y.is_cuda and wdz.is_cuda and y.numel() == self.n and wdz.numel() == self.n
        self._keep = (x.contiguous(), y.contiguous(), wdz.contiguous())
        self.nterms, self.napb, self.variant = nterms, napb, variant
        self.stream = stream
        h = ctypes.c_void_p()
        s = self._enter()
        _check(lib().eb_afmm_create(ctypes.byref(h), self.n, _ptr(self._keep[0]), _ptr(self._keep[1]),
                                    _ptr(self._keep[2]), int(napb), int(nterms), int(variant), _stream_ptr(stream)))
        self._leave(s, *self._keep)
        self._h = h

    def rebuild(self, x, y, wdz):
        raise EBError(EB_ERR_UNSUPPORTED, "an adaptive plan is built once per geometry: make a new AFMMPlan")

    def apply(self, dens, out=None):
        torch = self._torch
        assert dens.dtype == torch.complex128 and dens.numel() == self.n and dens.is_cuda
        dens = dens.contiguous()
        if out is None:
            out = torch.empty(self.n, dtype=torch.complex128, device=dens.device)
        s = self._enter()
        _check(lib().eb_afmm_apply(self._h, _ptr(dens), _ptr(out), _stream_ptr(self.stream)))
        self._leave(s, dens, out)
        return out

    def info(self):
        r = np.zeros(3)
        c = np.zeros(8, np.int64)
        _check(lib().eb_afmm_info(self._h, r.ctypes.data_as(_dp), c.ctypes.data_as(_i64p)))
        keys = ("boxes", "leaves", "levels", "near_pairs", "m2l_pairs", "m2p_evals", "p2l_terms", "shifted_boxes")
        d = dict(zip(keys, (int(v) for v in c)))
        d.update(rscal=float(r[0]), xonew=float(r[1]), yonew=float(r[2]), nterms=self.nterms, napb=self.napb)
        return d

    def m2l_split(self):
        """List-V pairs by source group: (near, far, far input terms)."""
        c = np.zeros(3, np.int64)
        _check(lib().eb_afmm_m2l_split(self._h, c.ctypes.data_as(_i64p)))
        return {"m2l_near": int(c[0]), "m2l_far": int(c[1]), "m2l_far_terms": int(c[2])}

    def scaled_points(self):
        """The tree's coordinates (x - XONEW) * RSCAL of the plan's points, as the plan forms
        them (the same two roundings): the sums are RSCAL times the sums there."""
        i = self.info()
        x, y = self._keep[0], self._keep[1]
        return (x - i["xonew"]) * i["rscal"], (y - i["yonew"]) * i["rscal"]

    def ntbox(self):
        raise EBError(EB_ERR_UNSUPPORTED, "no uniform levels in an adaptive plan")

    def apply_transpose(self, lam, out=None):
        raise EBError(EB_ERR_UNSUPPORTED, "the transposed product runs on uniform-grid plans (FMMPlan)")

    def enable_timing(self, on=True):
        _check(lib().eb_afmm_enable_timing(self._h, 1 if on else 0))

    def set_p2p_tile(self, tmax):
        raise EBError(EB_ERR_UNSUPPORTED, "the adaptive near field takes one childless box per work item")

    def stage_ms(self):
        a = np.zeros(8)
        _check(lib().eb_afmm_stage_ms(self._h, a.ctypes.data_as(_dp)))
        return dict(zip(("charges", "p2p", "p2m", "m2m", "m2l", "p2l", "l2l", "far"), a.tolist()))

    def check_lists(self):
        """Ordered pairs of childless boxes whose interaction the passes do not carry exactly
        once (0 for a correct plan); O(leaves^2) on the host."""
        bad = ctypes.c_int64()
        _check(lib().eb_afmm_check(self._h, ctypes.byref(bad)))
        return int(bad.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().eb_afmm_destroy(self._h)
            self._h = None


def direct_apply_noalloc(x, y, wdz, dens, out, variant=VARIANT_DR, stream=None):
    """GW of the O(N^2) sum (as an FMMPlan in MODE_DIRECT) in one call that allocates no device
    memory: every tensor, out included, is the caller's (n <= 10,240 points).  Calls that
    overlap on one GPU share the library's fixed working array and must not run concurrently."""
    n = int(x.numel())
    for t in (x, y, wdz, dens, out):
        assert t.is_cuda and t.is_contiguous() and t.numel() == n
    _check(lib().eb_direct_apply_noalloc(n, _ptr(x), _ptr(y), _ptr(wdz), _ptr(dens), int(variant), _ptr(out),
                                         _stream_ptr(stream)))
    return out


def direct_subset(x, y, wdz, dens, targets, variant=VARIANT_DR, stream=None):
    """GW of the O(N^2) sum over all sources at the listed targets (int32 device tensor of
    0-based indices): the tree-free checker of FMM accuracy.  Returns complex128 (ntgt,)."""
    torch = _torch()
    n = int(x.numel())
    x, y, wdz, dens = (t.contiguous() for t in (x, y, wdz, dens))
    tg = targets.to(device=x.device, dtype=torch.int32).contiguous()
    out = torch.empty(tg.numel(), dtype=torch.complex128, device=x.device)
    _check(lib().eb_direct_subset(n, _ptr(x), _ptr(y), _ptr(wdz), _ptr(dens), int(variant), int(tg.numel()), _ptr(tg),
                                  _ptr(out), _stream_ptr(stream)))
    return out


def tables(ilev, x, y, stream=None):
    """ASSIGN1 tables (cadap2.f:252-409) of level ilev computed on the GPU; 1-based numpy arrays."""
    torch = _torch()
    x = x.contiguous()
    y = y.contiguous()
    n = x.numel()
    nb = 4 ** ilev
    dev = x.device
    ibox = torch.empty(n, dtype=torch.int32, device=dev)
    icnt = torch.empty(nb, dtype=torch.int32, device=dev)
    ioffst = torch.empty(nb + 1, dtype=torch.int32, device=dev)
    iatadr = torch.empty(n, dtype=torch.int32, device=dev)
    iloc = torch.empty(nb, dtype=torch.int32, device=dev)
    ntbox = ctypes.c_int32()
    nsmax = ctypes.c_int32()
    _check(lib().eb_cadap_tables(ilev, n, _ptr(x), _ptr(y), _ptr(ibox), _ptr(icnt), _ptr(ioffst), _ptr(iatadr),
                                 _ptr(iloc), ctypes.byref(ntbox), ctypes.byref(nsmax), _stream_ptr(stream)))
    return dict(ibox=ibox.cpu().numpy(), icnt=icnt.cpu().numpy(), ioffst=ioffst.cpu().numpy(),
                iatadr=iatadr.cpu().numpy(), iloc=iloc.cpu().numpy(), ntbox=ntbox.value, nsmax=nsmax.value)


def lists(ilev, device="cuda", stream=None):
    """MKINTL / MKNBOR lists of every box of level ilev (GPU).  Returns (intl list, nbor list)."""
    torch = _torch()
    nb = 4 ** ilev
    intl = torch.zeros(nb * 27, dtype=torch.int32, device=device)
    nintl = torch.zeros(nb, dtype=torch.int32, device=device)
    nbor = torch.zeros(nb * 9, dtype=torch.int32, device=device)
    nnbor = torch.zeros(nb, dtype=torch.int32, device=device)
    _check(lib().eb_cadap_lists(ilev, _ptr(intl), _ptr(nintl), _ptr(nbor), _ptr(nnbor), _stream_ptr(stream)))
    intl = intl.cpu().numpy().reshape(nb, 27)
    nintl = nintl.cpu().numpy()
    nbor = nbor.cpu().numpy().reshape(nb, 9)
    nnbor = nnbor.cpu().numpy()
    return ([intl[b, : nintl[b]].copy() for b in range(nb)], [nbor[b, : nnbor[b]].copy() for b in range(nb)])


from . import atree, field, gmres, sim, spectral  # noqa: E402

EXPORTS = (EXPORTS + spectral.SPECTRAL_EXPORTS + atree.ATREE_EXPORTS + gmres.GMRES_EXPORTS + field.FIELD_EXPORTS
           + sim.SIM_EXPORTS)
