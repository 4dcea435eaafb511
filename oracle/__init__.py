"""Checkers for the Stokes-matvec path (TEST INFRASTRUCTURE ONLY).

Two CPU checkers live here:

* ``liboracle.so`` -- a plain-C restatement of the reference algorithm (``eb_oracle.c``),
  function by function, with file:line citations of the Fortran it follows;
* ``_ref/libref.so`` -- the reference Fortran itself, compiled in place from its sources by
  ``oracle/Makefile`` (only where the sources are present; the built library travels with
  the tree like any other built ``.so``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import this package, and only as the checker.  The product package
(``electricbubbles_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = os.environ.get("EB_REFERENCE_DIR", "/root/reference")
LIBORACLE = os.path.join(HERE, "liboracle.so")
LIBREF = os.path.join(HERE, "_ref", "libref.so")
# the reference with CADAPDR/CADAPMOD/CADAP bound to the GPU shim instead of its own FMM
LIBREF_ON_EB = os.path.join(HERE, "_ref", "libref_on_eb.so")

DR, MOD, PLUS = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so (always) and _ref/libref.so (when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if ref is None:
        ref = os.path.isdir(REF_SRC)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REF_SRC}"], check=True)
        subprocess.run(["make", "-s", "-C", HERE, "ref_on_eb", f"REF={REF_SRC}"], check=True)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _c_int(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _c_double(v):
    return ctypes.byref(ctypes.c_double(float(v)))


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBORACLE):
            raise RuntimeError(f"{LIBORACLE} is not built; run oracle.build()")
        _lib = ctypes.CDLL(LIBORACLE)
    return _lib


def ref_available() -> bool:
    return os.path.exists(LIBREF)


def ref():
    """The compiled reference (lazy binding: the plot routines of matplot.f stay unresolved)."""
    global _ref
    if _ref is None:
        if not os.path.exists(LIBREF):
            raise RuntimeError(f"{LIBREF} is not built (needs the reference sources)")
        _ref = ctypes.CDLL(LIBREF, mode=os.RTLD_LAZY)
    return _ref


def _cplx(a):
    """complex128 array -> contiguous interleaved float64 view"""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


# ------------------------------------------------------------------------------------
# C restatement
# ------------------------------------------------------------------------------------

def scale(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    n = x.size
    zs = np.empty(2 * n)
    rs = ctypes.c_double()
    rmin = np.empty(2)
    lib().ora_scale(n, _d(x), _d(y), _d(zs), ctypes.byref(rs), _d(rmin))
    return zs.view(np.complex128), rs.value, rmin


def assign1(ilev, zs):
    """ASSIGN1 tables for scaled points zs (complex array). Returns dict of 1-based tables."""
    zsd = _cplx(zs)
    n = zs.size
    nb = 4 ** ilev
    ibox = np.empty(n, np.int32)
    icnt = np.empty(nb, np.int32)
    ioffst = np.empty(nb + 1, np.int32)
    iatadr = np.empty(n, np.int32)
    iloc = np.empty(nb, np.int32)
    ntbox = ctypes.c_int()
    nsmax = ctypes.c_int()
    lib().ora_assign1(ilev, n, _d(zsd), _i(ibox), _i(icnt), _i(ioffst), _i(iatadr), _i(iloc),
                      ctypes.byref(ntbox), ctypes.byref(nsmax))
    return dict(ibox=ibox, icnt=icnt, ioffst=ioffst, iatadr=iatadr, iloc=iloc,
                ntbox=ntbox.value, nsmax=nsmax.value)


def mkintl(ilev, ibox):
    out = np.zeros(30, np.int32)
    n = ctypes.c_int()
    lib().ora_mkintl(ilev, ibox, _i(out), ctypes.byref(n))
    return out[: n.value].copy()


def mknbor(ilev, ibox):
    out = np.zeros(30, np.int32)
    n = ctypes.c_int()
    lib().ora_mknbor(ilev, ibox, _i(out), ctypes.byref(n))
    return out[: n.value].copy()


def cadap(variant, levmax, nd, x, y, dzdt, dens, h, tol):
    """Restated CADAPDR/CADAPMOD/CADAP (IFLAG1=0, IFLAG2=0). Returns (gradw, inform, ier, counts)."""
    nd = np.ascontiguousarray(nd, np.int32)
    h = np.ascontiguousarray(h, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    dz = _cplx(dzdt)
    de = _cplx(dens)
    n = int(nd.sum())
    gw = np.zeros(2 * n)
    inform = np.zeros(2, np.int32)
    ier = np.zeros(5, np.int32)
    counts = np.zeros(4, np.int64)
    lib().ora_cadap(variant, levmax, nd.size, _i(nd), _d(x), _d(y), _d(dz), _d(de), _d(h),
                    ctypes.c_double(tol), _d(gw), _i(inform), _i(ier),
                    counts.ctypes.data_as(_llp))
    return gw.view(np.complex128), inform, ier, counts


def direct(variant, nd, x, y, dzdt, dens, h, targets=None):
    """O(N^2) closed-form sum (long double). targets: 0-based indices (default all)."""
    nd = np.ascontiguousarray(nd, np.int32)
    h = np.ascontiguousarray(h, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    dz = _cplx(dzdt)
    de = _cplx(dens)
    n = int(nd.sum())
    if targets is None:
        targets = np.arange(n, dtype=np.int32)
    targets = np.ascontiguousarray(targets, np.int32)
    gw = np.zeros(2 * targets.size)
    lib().ora_direct(variant, nd.size, _i(nd), _d(x), _d(y), _d(dz), _d(de), _d(h),
                     targets.size, _i(targets), _d(gw))
    return gw.view(np.complex128)


def direct_transpose(variant, x, y, wdz, lam, targets=None):
    """The transpose (real inner product Re sum conj(a) b) of dens -> GW of ``direct`` with
    weights wdz = DZDT * H per point (variant DR / MOD / PLUS), or of the Hilbert product
    GW_i = sum_{j != i} wdz_j dens_j / (z_i - z_j) (variant 3), in long double, at the 0-based
    targets (default all).  Pinned by the dense transposes of ``direct`` and ``cauchy``."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    w = _cplx(wdz)
    l = _cplx(lam)
    n = x.size
    if targets is None:
        targets = np.arange(n, dtype=np.int32)
    targets = np.ascontiguousarray(targets, np.int32)
    out = np.zeros(2 * targets.size)
    lib().ora_direct_transpose(int(variant), n, _d(x), _d(y), _d(w), _d(l), targets.size, _i(targets), _d(out))
    return out.view(np.complex128)


def cauchy(x, y, qa, targets=None):
    """DAPIF2's Hilbert-matrix product (IFLAG7 = 3, dapif1.f:540-558) as an O(N^2) sum in long
    double: sum_{j != i} qa_j / (z_i - z_j) at the 0-based targets (default all)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    q = _cplx(qa)
    n = x.size
    if targets is None:
        targets = np.arange(n, dtype=np.int32)
    targets = np.ascontiguousarray(targets, np.int32)
    out = np.zeros(2 * targets.size)
    lib().ora_cauchy(n, _d(x), _d(y), _d(q), targets.size, _i(targets), _d(out))
    return out.view(np.complex128)


def fasmvp_epilogue(nd, x, y, dz, dsdth, rkappa, zk, u, stress):
    nd = np.ascontiguousarray(nd, np.int32)
    k = nd.size
    nbk = int(nd.sum())
    w = np.zeros(2 * nbk)
    za = np.zeros(2 * k)
    pk = np.zeros(k)
    lib().ora_fasmvp_epilogue(k, _i(nd), _d(np.ascontiguousarray(x, np.float64)),
                              _d(np.ascontiguousarray(y, np.float64)), _d(_cplx(dz)),
                              _d(np.ascontiguousarray(dsdth, np.float64)),
                              _d(np.ascontiguousarray(rkappa, np.float64)), _d(_cplx(zk)),
                              _d(np.ascontiguousarray(u, np.float64)), _d(_cplx(stress)),
                              _d(w), _d(za), _d(pk))
    return w, za.view(np.complex128), pk


def dft(c, sign):
    a = _cplx(c).copy()
    lib().ora_dft(a.size // 2, _d(a), sign)
    return a.view(np.complex128)


def fdifff(f):
    f = _cplx(f)
    out = np.empty_like(f)
    lib().ora_fdifff(f.size // 2, _d(f), _d(out))
    return out.view(np.complex128)


def extractdat(nd, rku):
    nd = np.ascontiguousarray(nd, np.int32)
    k = nd.size
    nbk = int(nd.sum())
    rku = np.ascontiguousarray(rku, np.float64)
    x, y = np.empty(nbk), np.empty(nbk)
    zk = np.empty(2 * k)
    dz = np.empty(2 * nbk)
    tha, dsda, rkappa = np.empty(nbk), np.empty(nbk), np.empty(nbk)
    ckern = np.empty(2 * nbk)
    area = np.empty(k)
    xl, sl = ctypes.c_double(), ctypes.c_double()
    lib().ora_extractdat(k, _i(nd), _d(rku), _d(x), _d(y), _d(zk), _d(dz), _d(tha), _d(dsda),
                         _d(rkappa), _d(ckern), _d(area), ctypes.byref(xl), ctypes.byref(sl))
    return dict(x=x, y=y, zk=zk.view(np.complex128), dz=dz.view(np.complex128), tha=tha,
                dsda=dsda, rkappa=rkappa, ckern=ckern.view(np.complex128), area=area,
                xleft=xl.value, slength=sl.value)


def equit(tha, un):
    tha = np.ascontiguousarray(tha, np.float64)
    un = np.ascontiguousarray(un, np.float64)
    t = np.empty_like(tha)
    lib().ora_equit(tha.size, _d(tha), _d(un), _d(t))
    return t


def filter_(cn):
    a = _cplx(cn).copy()
    lib().ora_filter(a.size // 2, _d(a))
    return a.view(np.complex128)


def ftail(nd, tha, un, dz, dsda):
    nd = np.ascontiguousarray(nd, np.int32)
    nbk = int(nd.sum())
    rkup = np.empty(2 * nbk)
    lib().ora_ftail(nd.size, _i(nd), _d(np.ascontiguousarray(tha, np.float64)),
                    _d(np.ascontiguousarray(un, np.float64)), _d(_cplx(dz)),
                    _d(np.ascontiguousarray(dsda, np.float64)), _d(rkup))
    return rkup


def rk2_stage1(dt, ybar, dy):
    ybar = np.ascontiguousarray(ybar, np.float64)
    dy = np.ascontiguousarray(dy, np.float64)
    phi0, y = np.empty_like(ybar), np.empty_like(ybar)
    lib().ora_rk2_stage1(ybar.size, ctypes.c_double(dt), _d(ybar), _d(dy), _d(phi0), _d(y))
    return phi0, y


def rk2_stage2(dt, phi0, dy, ybar):
    ybar = np.ascontiguousarray(ybar, np.float64).copy()
    lib().ora_rk2_stage2(ybar.size, ctypes.c_double(dt), _d(np.ascontiguousarray(phi0, np.float64)),
                         _d(np.ascontiguousarray(dy, np.float64)), _d(ybar))
    return ybar


# ------------------------------------------------------------------------------------
# The reference library (Fortran calling convention: everything by reference)
# ------------------------------------------------------------------------------------

def ref_cadap(variant, levmax, nd, x, y, dzdt, dens, h, tol, lenw=None, leniw=None):
    """Call the reference CADAPDR / CADAPMOD / CADAP with IFLAG1=0, IFLAG2=0."""
    r = ref()
    fn = {DR: r.cadapdr_, MOD: r.cadapmod_, PLUS: r.cadap_}[variant]
    nd = np.ascontiguousarray(nd, np.int32)
    n = int(nd.sum())
    p = -int(np.round(np.log(tol) / np.log(2.0)))
    if lenw is None:
        lenw = 8 * n + 2 * (p + 1) * min(n, 4 ** levmax) + 16
    if leniw is None:
        leniw = 5 * 4 ** levmax + 2 * n + 16
    iwork = np.zeros(leniw, np.int32)
    cwork = np.zeros(2 * lenw)
    gw = np.zeros(2 * n)
    dummy = np.zeros(4)
    inform = np.zeros(6, np.int32)
    ier = np.zeros(5, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    dz = _cplx(dzdt)
    de = _cplx(dens)
    h = np.ascontiguousarray(h, np.float64)
    fn(_c_int(0), _c_int(0), _c_int(levmax), _d(x), _d(y), _d(dz), _c_int(1), _c_int(nd.size),
       _i(nd), _d(de), _d(h), _i(iwork), _c_int(leniw), _d(cwork), _c_int(lenw),
       _d(dummy), _d(dummy), _c_int(0), _d(gw), _d(dummy), _d(dummy), _c_double(tol),
       _i(inform), _i(ier), _c_double(0.0), _c_double(0.0))
    return gw.view(np.complex128), inform[:2].copy(), ier.copy()


def ref_assign1(ilev, zs):
    r = ref()
    n = zs.size
    nb = 4 ** ilev
    zsd = _cplx(zs)
    ioffst = np.zeros(nb + 1, np.int32)
    itofst = np.zeros(nb + 1, np.int32)
    ibox = np.zeros(n, np.int32)
    icnt = np.zeros(nb, np.int32)
    icnt2 = np.zeros(nb, np.int32)
    iatadr = np.zeros(n, np.int32)
    itradr = np.zeros(1, np.int32)
    iloc = np.zeros(nb, np.int32)
    ntbox = ctypes.c_int()
    nsmax = ctypes.c_int()
    zt = np.zeros(2)
    r.assign1_(_c_int(0), _c_int(ilev), _d(zsd), _i(ioffst), _d(zt), _c_int(0), _i(ibox),
               _c_int(n), _i(icnt), _i(icnt2), _i(iatadr), _i(itradr), _i(itofst), _i(iloc),
               ctypes.byref(ntbox), ctypes.byref(nsmax))
    return dict(ibox=ibox, icnt=icnt, ioffst=ioffst, iatadr=iatadr, iloc=iloc,
                ntbox=ntbox.value, nsmax=nsmax.value)


def ref_mkintl(ilev, ibox):
    out = np.zeros(30, np.int32)
    n = ctypes.c_int()
    ref().mkintl_(_c_int(ilev), _c_int(ibox), _i(out), ctypes.byref(n))
    return out[: n.value].copy()


def ref_mknbor(ilev, ibox):
    out = np.zeros(30, np.int32)
    n = ctypes.c_int()
    ref().mknbor_(_c_int(ilev), _c_int(ibox), _i(out), ctypes.byref(n))
    return out[: n.value].copy()


def _wsave(n):
    w = np.zeros(2 * (4 * n + 15) + 64)
    ref().dcffti_(_c_int(n), _d(w))
    return w


def ref_dcfftf(c):
    a = _cplx(c).copy()
    n = a.size // 2
    w = _wsave(n)
    ref().dcfftf_(_c_int(n), _d(a), _d(w))
    return a.view(np.complex128)


def ref_dcfftb(c):
    a = _cplx(c).copy()
    n = a.size // 2
    w = _wsave(n)
    ref().dcfftb_(_c_int(n), _d(a), _d(w))
    return a.view(np.complex128)


def ref_fdifff(f):
    f = _cplx(f).copy()
    n = f.size // 2
    w = _wsave(n)
    out = np.zeros_like(f)
    ref().fdifff_(_d(f), _d(out), _c_int(n), _d(w))
    return out.view(np.complex128)


def ref_filter(cn):
    a = _cplx(cn).copy()
    ref().filter_(_c_int(a.size // 2), _d(a))
    return a.view(np.complex128)


def ref_equit(tha, un):
    tha = np.ascontiguousarray(tha, np.float64)
    un = np.ascontiguousarray(un, np.float64)
    n = tha.size
    t = np.zeros(n)
    zf = np.zeros(2 * n)
    zfi = np.zeros(2 * n)
    w = _wsave(n)
    ref().equit_(_c_int(n), _d(tha), _d(un), _d(t), _d(zf), _d(zfi), _d(w))
    return t


def ref_extractdat(nd, rku):
    nd = np.ascontiguousarray(nd, np.int32)
    k = nd.size
    nbk = int(nd.sum())
    ifmax = 2 * (4 * int(nd.max()) + 15) // 2 + 64   # complex words per body in WSAVE
    rku = np.ascontiguousarray(rku, np.float64)
    x, y = np.zeros(nbk), np.zeros(nbk)
    z = np.zeros(2 * nbk)
    zk = np.zeros(2 * k)
    dz = np.zeros(2 * nbk)
    tha, dsda, rkappa = np.zeros(nbk), np.zeros(nbk), np.zeros(nbk)
    ckern = np.zeros(2 * nbk)
    za = np.zeros(2 * int(nd.max()))
    zai = np.zeros(2 * int(nd.max()))
    wsave = np.zeros(2 * ifmax * k)
    zc = np.zeros(2 * k)
    xl, sl = ctypes.c_double(), ctypes.c_double()
    zbond = np.zeros(2)
    ref().extractdat_(_c_int(k), _i(nd), _c_int(nbk), _d(x), _d(y), _d(z), _d(zk), _d(dz),
                      _d(tha), _d(dsda), _d(rkappa), _d(ckern), _d(rku), _d(za), _d(zai),
                      _d(wsave), _c_int(ifmax), _d(zc), ctypes.byref(xl), ctypes.byref(sl),
                      _d(zbond), _c_int(0))
    return dict(x=x, y=y, zk=zk.view(np.complex128), dz=dz.view(np.complex128), tha=tha,
                dsda=dsda, rkappa=rkappa, ckern=ckern.view(np.complex128),
                xleft=xl.value, slength=sl.value)


def ref_fasmvp(nd, x, y, dz, dsdth, rkappa, zk, u):
    """Reference FASMVP (stokesbubb.f:496-614): calls CADAPDR (LEVMAX 7, TOL 1e-13) itself."""
    nd = np.ascontiguousarray(nd, np.int32)
    k = nd.size
    nbk = int(nd.sum())
    h = (2 * np.pi / nd).astype(np.float64)
    w = np.zeros(2 * nbk)
    za = np.zeros(2 * k)
    pk = np.zeros(k)
    stress = np.zeros(2 * nbk)
    dens = np.zeros(2 * nbk)
    ref().fasmvp_(_i(nd), _c_int(nbk), _c_int(k), _d(np.ascontiguousarray(x, np.float64)),
                  _d(np.ascontiguousarray(y, np.float64)), _d(_cplx(dz)),
                  _d(np.ascontiguousarray(dsdth, np.float64)),
                  _d(np.ascontiguousarray(rkappa, np.float64)), _d(_cplx(zk)),
                  _d(np.ascontiguousarray(u, np.float64)), _d(w), _d(h), _d(za), _d(pk),
                  _d(stress), _d(dens))
    return w, stress.view(np.complex128), za.view(np.complex128), pk


# ------------------------------------------------------------------------------------
# Adaptive quadtree of DAPIF2 (dapif1.f:740-1680, dapif2.f:3-474)
# ------------------------------------------------------------------------------------

ATREE_TABLES = ("xs", "ys", "nl", "ibox", "xb", "yb", "nbat", "npar", "nchi", "indb", "iat", "jat",
                "l1", "l2", "jl", "l3", "l4", "jl3", "jl4")


def _nint(v):
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))


def dapif_lmx(mex=300, tol=1e-11):
    """Maximum level count DAPIF2 allows (dapif1.f:690-704): N = -NINT(log2 TOL) + 3,
    LMX = INT((MEX log2 10 - 1)/N + 6)."""
    n = -_nint(math.log(tol) / math.log(2.0)) + 3
    rex2 = mex * (math.log(10.0) / math.log(2.0)) - 1
    return int(rex2 / n + 6)


def atree(x, y, napb, lmx=None, eps7=1e-16):
    """The C restatement of DAPIF2's tree.  Returns a dict of the 1-based tables, the
    scaled coordinates xs/ys, geom = (rscal, xonew, yonew) and the counts."""
    L = lib()
    L.ora_atree_build.restype = ctypes.c_void_p
    L.ora_atree_build.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.ora_atree_sizes.argtypes = [ctypes.c_void_p, _ip]
    L.ora_atree_copy.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 20
    L.ora_atree_free.argtypes = [ctypes.c_void_p]
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    lmx = dapif_lmx() if lmx is None else lmx
    h = L.ora_atree_build(x.size, _d(x), _d(y), int(napb), int(lmx), float(eps7))
    try:
        sz = np.zeros(10, np.int32)
        L.ora_atree_sizes(h, _i(sz))
        ierr, nbox, nlev, ln, ln3, ln4, maxp, jatlen, nat = (int(v) for v in sz[:9])
        out = dict(ierr=ierr, nbox=nbox, nlev=nlev, len=ln, len3=ln3, len4=ln4, maxp=maxp)
        if ierr != 0:
            return out
        shapes = dict(xs=(nat, np.float64), ys=(nat, np.float64), geom=(3, np.float64), nl=(nlev + 2, np.int32),
                      ibox=(nat, np.int32), xb=(nbox, np.float64), yb=(nbox, np.float64))
        for k in ("nbat", "npar", "nchi", "indb", "iat"):
            shapes[k] = (nbox, np.int32)
        shapes.update(jat=(jatlen, np.int32), l1=(nbox + 1, np.int32), l2=(nbox + 1, np.int32),
                      jl=(ln, np.int32), l3=(nbox + 1, np.int32), l4=(nbox + 1, np.int32),
                      jl3=(ln3, np.int32), jl4=(ln4, np.int32))
        order = ("xs", "ys", "geom", "nl", "ibox", "xb", "yb", "nbat", "npar", "nchi", "indb", "iat", "jat",
                 "l1", "l2", "jl", "l3", "l4", "jl3", "jl4")
        arrs = {k: np.zeros(shapes[k][0], shapes[k][1]) for k in order}
        L.ora_atree_copy(h, *(arrs[k].ctypes.data for k in order))
        out.update(arrs)
        return out
    finally:
        L.ora_atree_free(h)


_CLOSE_STUB = ctypes.CFUNCTYPE(None)(lambda: None)


def ref_dapif2_scaled(x, y, napb, tol=1e-11, mex=300, eps7=1e-16, seed=0):
    """Run the reference DAPIF2 (IFLAG7 = 3, NINIRE = 2, the parameters of laplace.f:204-215
    apart from NAPB) and return the scaled coordinates it leaves in WKSP(1..2 NAT), with
    INFORM(2) = NBOX and INFORM(3) = number of levels."""
    r = ref()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    nat = x.size
    rng = np.random.default_rng(seed)
    qa = _cplx(rng.normal(size=nat) + 1j * rng.normal(size=nat)).copy()
    poten = np.zeros(nat)
    field = np.zeros(2 * nat)
    nsp = 4_000_000 + 400 * nat
    wksp = np.zeros(nsp)
    iout = np.zeros(2, np.int32)
    ierr = np.zeros(8, np.int32)
    inform = np.zeros(8, np.int32)
    r.dapif2_(_i(iout), _c_int(3), _c_int(nat), _c_int(napb), _c_int(2), _c_int(mex), _i(ierr), _i(inform),
              _c_double(tol), _c_double(eps7), _d(x), _d(y), _d(qa), _d(poten), _d(field), _d(wksp),
              _c_int(nsp), _CLOSE_STUB)
    return dict(ierr=int(ierr[0]), xs=wksp[:nat].copy(), ys=wksp[nat:2 * nat].copy(), nbox=int(inform[1]),
                nlev=int(inform[2]))


def ref_dapif2(x, y, qa, napb=30, tol=1e-11, mex=300, eps7=1e-16, nsp=None):
    """The reference DAPIF2 with IFLAG7 = 3 (the Hilbert-matrix product, as laplace.f:204-215
    and pv_integral.f:171-182 call it).  Returns (field, ierr[0:4], inform[0:4])."""
    r = ref()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    nat = x.size
    q = _cplx(qa).copy()
    poten = np.zeros(nat)
    field = np.zeros(2 * nat)
    if nsp is None:
        nsp = 4_000_000 + 400 * nat
    # the buffer is never smaller than the default: with a short NSP, DALI34 can write its lists
    # past NSP before the next workspace check (dapif1.f:918-929)
    wksp = np.zeros(max(nsp, 4_000_000 + 400 * nat))
    iout = np.zeros(2, np.int32)
    ierr = np.zeros(8, np.int32)
    inform = np.zeros(8, np.int32)
    r.dapif2_(_i(iout), _c_int(3), _c_int(nat), _c_int(napb), _c_int(2), _c_int(mex), _i(ierr), _i(inform),
              _c_double(tol), _c_double(eps7), _d(x), _d(y), _d(q), _d(poten), _d(field), _d(wksp),
              _c_int(nsp), _CLOSE_STUB)
    return field.view(np.complex128), ierr[:4].copy(), inform[:4].copy()


def ref_atree(xs, ys, napb, lmx=None, eps7=1e-16):
    """DAPIF2's table builders called one by one from the reference (DAASSI, DACOMP, DAINDB,
    DAIJAT, DALI12, DALI34, in the order of dapif1.f:794-957) on scaled coordinates, with
    arrays of their own.  SHIFT gets one zero word in front, the value DACOMP reads at
    SHIFT(0) inside DAPIF2's zeroed workspace."""
    r = ref()
    xs = np.ascontiguousarray(xs, np.float64)
    ys = np.ascontiguousarray(ys, np.float64)
    nat = xs.size
    lmx = dapif_lmx() if lmx is None else lmx
    mx = 8 * nat + 64
    qa = np.zeros(2 * nat)
    ierr = np.zeros(8, np.int32)
    nl = np.zeros(128, np.int32)
    ibox = np.zeros(nat + 1, np.int32)
    xb = np.zeros(mx + 2)
    yb = np.zeros(mx + 2)
    nbat = np.zeros(mx + 2, np.int32)
    npar = np.zeros(mx + 2, np.int32)
    nchi = np.zeros(mx + 2, np.int32)
    nbox = ctypes.c_int()
    lev = ctypes.c_int()
    r.daassi_(_c_double(eps7), _c_int(napb), _c_int(nat), ctypes.byref(nbox), _c_int(mx), _i(ierr),
              ctypes.byref(lev), _c_int(lmx), _i(nl), _d(xs), _d(ys), _d(qa), _i(ibox), _d(xb), _d(yb),
              _i(nbat), _i(npar), _i(nchi))
    if ierr[0] != 0:
        return dict(ierr=int(ierr[0]))
    shift = np.zeros(mx + 3, np.int32)
    nbol = ctypes.c_int()
    r.dacomp_(_c_int(nat), ctypes.byref(nbox), ctypes.byref(nbol), ctypes.byref(lev), _i(nl), _i(ibox), _d(xb),
              _d(yb), _i(nbat), _i(npar), _i(nchi), _i(shift[1:]))
    nb = nbox.value
    Lv = lev.value
    indb = np.zeros(nb + 2, np.int32)
    maxp = ctypes.c_int()
    r.daindb_(_c_int(napb), _c_int(nat), ctypes.byref(maxp), ctypes.byref(nbox), ctypes.byref(lev), _i(nl),
              _i(nbat), _i(nchi), _i(indb))
    iat = np.zeros(nb + 2, np.int32)
    jat = np.zeros(nb + nat + 2, np.int32)
    r.daijat_(_c_int(napb), _c_int(nat), ctypes.byref(nbox), ctypes.byref(lev), _i(nl), _i(ibox), _i(nbat),
              _i(indb), _i(iat), _i(jat))
    l1 = np.zeros(nb + 2, np.int32)
    l2 = np.zeros(nb + 2, np.int32)
    jl = np.zeros(36 * (nb + 1) + 2, np.int32)
    ln = ctypes.c_int()
    r.dali12_(ctypes.byref(nbox), ctypes.byref(ln), ctypes.byref(lev), _i(nl), _d(xb), _d(yb), _i(npar),
              _i(nchi), _i(l1), _i(l2), _i(jl))
    l3 = np.zeros(nb + 2, np.int32)
    l4 = np.zeros(nb + 2, np.int32)
    idesc = np.zeros(4 * nb + 64, np.int32)
    big = 256 * nb + 4096
    jl3 = np.zeros(big, np.int32)
    jl4 = np.zeros(big, np.int32)
    ln3 = ctypes.c_int()
    ln4 = ctypes.c_int()
    r.dali34_(_c_int(napb), ctypes.byref(nbox), ctypes.byref(ln3), ctypes.byref(ln4), ctypes.byref(lev), _i(nl),
              _d(xb), _d(yb), _i(nbat), _i(npar), _i(nchi), _i(indb), _i(l1), _i(l2), _i(jl), _i(l3), _i(l4),
              _i(idesc), _i(jl3), _i(jl4))
    lead = int(l1[nb]) - 1
    jatlen = 0
    leaves = np.nonzero(indb[:nb] == 1)[0]
    if leaves.size:
        last = leaves[-1]
        jatlen = int(iat[last]) + int(nbat[last])
    return dict(ierr=0, nbox=nb, nlev=Lv - 1, len=lead, len3=ln3.value, len4=ln4.value, maxp=maxp.value,
                xs=xs.copy(), ys=ys.copy(), nl=nl[:Lv + 1].copy(), ibox=ibox[:nat].copy(), xb=xb[:nb].copy(),
                yb=yb[:nb].copy(), nbat=nbat[:nb].copy(), npar=npar[:nb].copy(), nchi=nchi[:nb].copy(),
                indb=indb[:nb].copy(), iat=iat[:nb].copy(), jat=jat[:jatlen].copy(), l1=l1[:nb + 1].copy(),
                l2=l2[:nb + 1].copy(), jl=jl[:lead].copy(), l3=l3[:nb + 1].copy(), l4=l4[:nb + 1].copy(),
                jl3=jl3[:ln3.value].copy(), jl4=jl4[:ln4.value].copy())


# ------------------------------------------------------------------------------------
# DGMRES (dgmres.f:2-553) of the reference, on a dense operator or on the Stokes MATVEC
# ------------------------------------------------------------------------------------

_MATVEC_T = ctypes.CFUNCTYPE(None, _ip, _dp, _dp, _ip, _ip, _ip, _dp, _ip)


def _dgmres_call(lib_, n, b, x, matvec_ptr, msolve_ptr, maxl, tol, nrmax):
    lrgw = 1 + n * (maxl + 6) + maxl * (maxl + 3) + 16
    rgwk = np.zeros(lrgw)
    igwk = np.zeros(20, np.int32)
    igwk[0] = maxl
    igwk[4] = nrmax
    it = ctypes.c_int()
    err = ctypes.c_double()
    ierr = ctypes.c_int()
    sb = np.ones(n)
    sx = np.ones(n)
    dummy_i = np.zeros(4, np.int32)
    dummy_d = np.zeros(4)
    lib_.dgmres_.argtypes = [_ip, _dp, _dp, _ip, _ip, _ip, _dp, _ip, ctypes.c_void_p, ctypes.c_void_p, _ip, _dp,
                             _ip, _ip, _dp, _ip, _ip, _dp, _dp, _dp, _ip, _ip, _ip, _dp, _ip]
    lib_.dgmres_(_c_int(n), _d(b), _d(x), _c_int(0), _i(dummy_i), _i(dummy_i), _d(dummy_d), _c_int(0),
                 matvec_ptr, msolve_ptr, _c_int(0), _c_double(tol), _c_int(0), ctypes.byref(it), ctypes.byref(err),
                 ctypes.byref(ierr), _c_int(0), _d(sb), _d(sx), _d(rgwk), _c_int(lrgw), _i(igwk), _c_int(20),
                 _d(dummy_d), _i(dummy_i))
    return x, it.value, err.value, ierr.value


def ref_dgmres_dense(a, b, x0, maxl=10, tol=0.0, nrmax=0):
    """The reference DGMRES (ITOL 0, IUNIT 0, no scaling or preconditioner, IGWK(2..4) = 0)
    with a MATVEC that multiplies by the dense matrix a.  Returns (x, iter, err, ierr)."""
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]

    def mv(pn, px, py, *_):
        xv = np.ctypeslib.as_array(px, shape=(n,))
        yv = np.ctypeslib.as_array(py, shape=(n,))
        yv[:] = a @ xv

    cb = _MATVEC_T(mv)
    b = np.ascontiguousarray(b, np.float64).copy()
    x = np.ascontiguousarray(x0, np.float64).copy()
    ptr = ctypes.cast(cb, ctypes.c_void_p)
    return _dgmres_call(ref(), n, b, x, ptr, ptr, maxl, tol, nrmax)


def ref_dgmres_stokes(nd, z, dz, dsdth, rkappa, zk, rhs, maxl=100, tol=1e-10, nrmax=-1):
    """SOLVEI's solve (stokesbubb.f:329-358): the reference DGMRES with the reference MATVEC
    (FASMVP + CADAPDR on the geometry of COMMON /inteqn/ and /problem/), from soln = rhs.
    Returns (soln, iter, err, ierr)."""
    r = ref()
    nd = np.asarray(nd, np.int32)
    k = nd.size
    nbk = int(nd.sum())
    kmax, nmax = 150, 450000
    prob = (ctypes.c_char * (16 * kmax + 4 * kmax + 12)).from_address(
        ctypes.addressof(ctypes.c_char.in_dll(r, "problem_")))
    zkv = np.frombuffer(prob, dtype=np.complex128, count=kmax)
    zkv[:] = 0
    zkv[:k] = zk
    ndv = np.frombuffer(prob, dtype=np.int32, count=kmax, offset=16 * kmax)
    ndv[:] = 0
    ndv[:k] = nd
    np.frombuffer(prob, dtype=np.int32, count=3, offset=20 * kmax)[:] = [k, nbk, 2 * nbk]
    inteqn = (ctypes.c_char * (48 * nmax)).from_address(ctypes.addressof(ctypes.c_char.in_dll(r, "inteqn_")))
    np.frombuffer(inteqn, dtype=np.float64, count=nbk, offset=0)[:] = dsdth
    np.frombuffer(inteqn, dtype=np.float64, count=nbk, offset=8 * nmax)[:] = rkappa
    np.frombuffer(inteqn, dtype=np.complex128, count=nbk, offset=16 * nmax)[:] = z
    np.frombuffer(inteqn, dtype=np.complex128, count=nbk, offset=32 * nmax)[:] = dz
    b = np.ascontiguousarray(rhs, np.float64).copy()
    x = b.copy()
    mv = ctypes.cast(r.matvec_, ctypes.c_void_p)
    ms = ctypes.cast(r.msolve_, ctypes.c_void_p)
    try:
        return _dgmres_call(r, 2 * nbk, b, x, mv, ms, maxl, tol, nrmax)
    finally:
        # FASMVP leaves PRINI(6, 13) set (stokesbubb.f:572): silence the reference's printing again
        r.prini_(_c_int(0), _c_int(0))


# ------------------------------------------------------------------------------------
# The reference's own right-hand side and time step, on a chosen library handle
# ------------------------------------------------------------------------------------

def rk4(rhs, y0, dt, nsteps, record=None):
    """Classic fourth-order Runge-Kutta (the textbook method: y += dt/6 (k1 + 2 k2 + 2 k3 + k4))
    around a given right-hand side, in the order of the library's stage updates.  The reference
    integrates with RK2 only (electric.f:762-786); with RefSim.rhs (the reference's own F) this
    is the checker of the library's RK4 runs.  record(step, y) is called after every step."""
    y = np.ascontiguousarray(y0, np.float64).copy()
    for step in range(nsteps):
        k = rhs(y)
        acc = k.copy()
        ys = y + (0.5 * dt) * k
        k = rhs(ys)
        acc = acc + 2.0 * k
        ys = y + (0.5 * dt) * k
        k = rhs(ys)
        acc = acc + 2.0 * k
        ys = y + dt * k
        k = rhs(ys)
        y = y + (dt / 6.0) * (acc + k)
        if record is not None:
            record(step + 1, y)
    return y


def deformation(nd, rku):
    """D = (L - B) / (L + B) of every body of the interleaved state rku: L, B the semi-axes of
    the ellipse with the body's area and second moments of area, from trapezoid sums of
    Green's-theorem integrands with the spectral tangent (FDIFFF's wavenumbers: the Nyquist
    coefficient dropped).  The reference prints no deformation; pinned by ellipses' closed form."""
    z = np.ascontiguousarray(rku, np.float64).view(np.complex128)
    out = []
    o = 0
    for N in np.asarray(nd, np.int64):
        zb = z[o:o + N]
        o += N
        kw = np.fft.fftfreq(N, 1.0 / N)
        if N % 2 == 0:
            kw[N // 2] = 0.0
        za = np.fft.ifft(1j * kw * np.fft.fft(zb))
        zb = zb - zb.mean()
        x, y, xa, ya = zb.real, zb.imag, za.real, za.imag
        h = 2.0 * np.pi / N
        A = 0.5 * h * np.sum(x * ya - y * xa)
        mx, my = 0.5 * h * np.sum(x * x * ya), -0.5 * h * np.sum(y * y * xa)
        ixx, iyy = h * np.sum(x ** 3 * ya) / 3.0, -h * np.sum(y ** 3 * xa) / 3.0
        ixy = 0.5 * h * np.sum(x * x * y * ya)
        cx, cy = mx / A, my / A
        a, c, b = ixx - A * cx * cx, iyy - A * cy * cy, ixy - A * cx * cy
        rt = np.sqrt(0.25 * (a - c) ** 2 + b * b)
        sp, sm = np.sqrt(0.5 * (a + c) + rt), np.sqrt(0.5 * (a + c) - rt)
        out.append((sp - sm) / (sp + sm))
    return np.array(out)


class RefSim:
    """Drive the reference's F (electric.f:968-1109) and RK2 (electric.f:762-786) in-process.

    `path` is libref.so (the reference as written) or libref_on_eb.so (the same objects with
    CADAPDR / CADAPMOD / CADAP bound to the GPU shim).  COMMON blocks are set the way
    MANYDROP sets them (electric.f:52-56, 67-70): /problem/ nd, k; /flowinfo/ cap, zbond, E0;
    /wallinfo/ iwall.  KMAX = 150 fixes the /problem/ layout (electric.f:10)."""

    KMAX = 150

    def __init__(self, path):
        self.lib = ctypes.CDLL(path, mode=os.RTLD_LAZY)

    def _block(self, name, nbytes):
        addr = ctypes.addressof(ctypes.c_char.in_dll(self.lib, name))
        return (ctypes.c_char * nbytes).from_address(addr)

    def setup(self, nd, cap=0.0, zbond=0j, e0=0.0):
        nd = np.asarray(nd, np.int32)
        k = nd.size
        prob = self._block("problem_", 16 * self.KMAX + 4 * self.KMAX + 12)
        ndv = np.frombuffer(prob, dtype=np.int32, count=self.KMAX, offset=16 * self.KMAX)
        ndv[:] = 0
        ndv[:k] = nd
        ints = np.frombuffer(prob, dtype=np.int32, count=3, offset=20 * self.KMAX)
        ints[:] = [k, int(nd.sum()), 2 * int(nd.sum())]
        flow = np.frombuffer(self._block("flowinfo_", 32), dtype=np.float64, count=4)
        flow[:] = [cap, zbond.real, zbond.imag, e0]
        wall = self._block("wallinfo_", 24)
        np.frombuffer(wall, dtype=np.float64, count=2)[:] = 0.0
        np.frombuffer(wall, dtype=np.int32, count=1, offset=16)[:] = 0
        self.k = k
        self.neq = 2 * int(nd.sum())

    def rhs(self, ybar, t=0.0):
        ybar = np.ascontiguousarray(ybar, np.float64).copy()
        dy = np.zeros_like(ybar)
        self.lib.f_(_c_int(self.k), _c_int(self.neq), _c_double(t), _d(ybar), _d(dy))
        return dy

    def rk2(self, ybar, dt, nsteps, t0=0.0):
        ybar = np.ascontiguousarray(ybar, np.float64).copy()
        y = np.zeros_like(ybar)
        dy = np.zeros_like(ybar)
        phi = np.zeros(4 * ybar.size)
        t = t0
        for _ in range(nsteps):
            self.lib.rk2_(_c_int(self.k), _c_int(self.neq), _c_double(t), _c_double(dt), _d(phi), _d(ybar),
                          _d(y), _d(dy))
            t += dt
        return ybar

    def fasmvp_lap(self, nd, zk, z, dz, rkappa, dsda, u):
        """FASMVP_LAP (laplace.f:159-262) on a real density u: the nbk rows of the Laplace
        operator (the routine also writes k constraint rows past them)."""
        nd = np.ascontiguousarray(nd, np.int32)
        k = nd.size
        nbk = int(nd.sum())
        w = np.zeros(nbk + k + 8)
        ak = np.zeros(k + 8)
        self.lib.fasmvp_lap_(_c_int(k), _i(nd), _c_int(nbk), _d(_cplx(zk)), _d(_cplx(z)), _d(_cplx(dz)),
                             _d(np.ascontiguousarray(rkappa, np.float64)), _d(np.ascontiguousarray(dsda, np.float64)),
                             _d(ak), _d(np.ascontiguousarray(u, np.float64)), _d(w))
        return w[:nbk].copy()

    def fasmvp(self, nd, x, y, dz, dsdth, rkappa, zk, u):
        nd = np.ascontiguousarray(nd, np.int32)
        k = nd.size
        nbk = int(nd.sum())
        h = (2 * np.pi / nd).astype(np.float64)
        w = np.zeros(2 * nbk)
        za = np.zeros(2 * k)
        pk = np.zeros(k)
        stress = np.zeros(2 * nbk)
        dens = np.zeros(2 * nbk)
        self.lib.fasmvp_(_i(nd), _c_int(nbk), _c_int(k), _d(np.ascontiguousarray(x, np.float64)),
                         _d(np.ascontiguousarray(y, np.float64)), _d(_cplx(dz)),
                         _d(np.ascontiguousarray(dsdth, np.float64)),
                         _d(np.ascontiguousarray(rkappa, np.float64)), _d(_cplx(zk)),
                         _d(np.ascontiguousarray(u, np.float64)), _d(w), _d(h), _d(za), _d(pk),
                         _d(stress), _d(dens))
        return w, stress.view(np.complex128)


# ------------------------------------------------------------------------------------
# The reference's F, stage by stage (electric.f:968-1109 with iwall = 0)
# ------------------------------------------------------------------------------------

def ref_field_chain(nd, rku, cap=0.0, e0=0.0, maxl=100):
    """Run the stages of the reference's F on one geometry and return every intermediate:
    EXTRACTDAT, GETBCs, SOLVE_LAPL (laplace.f:335-425), DIRICH_NEU_MAP (laplace.f:266-332),
    SETVELF (stokesbubb.f:4-135), SOLVEI (stokesbubb.f:294-415) and INTERFACENEW
    (electric.f:791-912), in F's order, with COMMON /problem/, /inteqn/, /flowinfo/ and
    /wallinfo/ set as F sees them (zbond = 0, iwall = 0).  Run it from a scratch directory:
    the reference writes fort.13."""
    r = ref()
    nd = np.ascontiguousarray(nd, np.int32)
    k = nd.size
    nbk = int(nd.sum())
    kmax, nmax, ndmax = 150, 450000, 3000
    sim = RefSim(LIBREF)
    sim.lib = r
    sim.setup(nd, cap=cap, zbond=0j, e0=e0)
    inteqn = (ctypes.c_char * (48 * nmax)).from_address(ctypes.addressof(ctypes.c_char.in_dll(r, "inteqn_")))
    base = ctypes.addressof(inteqn)
    c_dsda = ctypes.c_void_p(base)
    c_rkappa = ctypes.c_void_p(base + 8 * nmax)
    c_z = ctypes.c_void_p(base + 16 * nmax)
    c_dz = ctypes.c_void_p(base + 32 * nmax)
    prob = (ctypes.c_char * (16 * kmax)).from_address(ctypes.addressof(ctypes.c_char.in_dll(r, "problem_")))
    c_zk = ctypes.c_void_p(ctypes.addressof(prob))
    ifwork = 10 * ndmax + 50
    x, y = np.zeros(nbk), np.zeros(nbk)
    tha, ckern = np.zeros(nbk), np.zeros(2 * nbk)
    zf, zfi = np.zeros(2 * nbk + 64), np.zeros(2 * nbk + 64)
    wsave = np.zeros(2 * ifwork * k)
    zc = np.zeros(2 * k)
    xl, sl = ctypes.c_double(), ctypes.c_double()
    zbond = np.zeros(2)
    rku = np.ascontiguousarray(rku, np.float64)
    r.extractdat_(_c_int(k), _i(nd), _c_int(nbk), _d(x), _d(y), c_z, c_zk, c_dz, _d(tha), c_dsda, c_rkappa,
                  _d(ckern), _d(rku), _d(zf), _d(zfi), _d(wsave), _c_int(ifwork), _d(zc), ctypes.byref(xl),
                  ctypes.byref(sl), _d(zbond), _c_int(0))
    view = lambda off, cnt, dt: np.frombuffer(inteqn, dtype=dt, count=cnt, offset=off).copy()  # noqa: E731
    dsda = view(0, nbk, np.float64)
    rkappa = view(8 * nmax, nbk, np.float64)
    z = view(16 * nmax, nbk, np.complex128)
    dz = view(32 * nmax, nbk, np.complex128)
    zk = np.frombuffer(prob, dtype=np.complex128, count=k).copy()
    # Laplace: GETBCs and SOLVE_LAPL (N = nbk, MAXL 100, TOL 1e-10, no restart)
    rhs = np.zeros(2 * nbk + kmax)
    r.getbcs_(_c_int(k), _i(nd), _c_int(nbk), c_zk, c_z, _d(rhs))
    rhs_lap = rhs[:nbk].copy()
    n_l = nbk
    lrw = 10 + (2 * nbk + k) * (maxl + 6) + maxl * (maxl + 3) + 4096  # FASMVP_LAP writes k rows past N
    rwork = np.zeros(lrw)
    iwork = np.zeros(64, np.int32)
    iwork[0] = maxl
    soln = np.zeros(2 * nbk + kmax)
    ulap = np.zeros(nbk + kmax)
    ak = np.zeros(kmax)
    r.solve_lapl_(_i(nd), _c_int(k), _c_int(nbk), _d(rhs), _d(soln), _d(ulap), _d(ak), _d(rwork), _c_int(lrw),
                  _i(iwork), _c_int(30), c_dsda)
    ulap = ulap[:nbk].copy()
    # DIRICH_NEU_MAP
    phi, psi, dphi = np.zeros(nbk), np.zeros(nbk), np.zeros(nbk)
    r.dirich_neu_map_(_c_int(k), _i(nd), _c_int(nbk), c_z, _d(ulap), c_dz, c_zk, _d(ak), _d(phi), _d(psi),
                      _d(dphi), c_dsda)
    # SETVELF
    arg = np.zeros(nbk + 64)
    zb = np.zeros(2 * k)
    rhs_s = np.zeros(2 * nbk + kmax)
    r.setvelf_(_c_int(k), _i(nd), _c_int(nbk), c_z, c_zk, c_dz, _d(arg), _d(zb), _d(zc), _d(zf), _d(zfi),
               _d(wsave), _c_int(ifwork), _d(rhs_s), _c_int(0), _d(zbond), _c_double(cap), c_dsda, _d(dphi),
               _c_double(e0))
    rhs_stokes = rhs_s[:2 * nbk].copy()
    # SOLVEI (N = 2 nbk)
    rwork[:] = 0.0
    iwork[:] = 0
    iwork[0] = maxl
    soln[:] = 0.0
    pk = np.zeros(k)
    pinf = ctypes.c_double()
    za = np.zeros(2 * k)
    u = np.zeros(2 * nbk)
    r.solvei_(_i(nd), _c_int(k), _c_int(nbk), _d(rhs_s), _d(soln), _d(pk), ctypes.byref(pinf), _d(za), c_zk, c_z,
              c_dz, _d(u), _d(rwork), _c_int(lrw), _i(iwork), _c_int(30))
    # INTERFACENEW
    h = (2 * np.pi / nd).astype(np.float64)
    un = np.zeros(nbk)
    r.interfacenew_(_i(nd), _c_int(k), _c_int(nbk), _d(h), _d(x), _d(y), c_dz, _d(u), _d(un), c_rkappa,
                    _d(ckern), _d(wsave), c_zk, _c_int(ifwork), _d(zc), _c_int(0), ctypes.byref(xl),
                    ctypes.byref(sl), _c_double(cap), c_dsda, _d(zb))
    r.prini_(_c_int(0), _c_int(0))
    return dict(x=x, y=y, z=z, zk=zk, dz=dz, dsda=dsda, rkappa=rkappa, tha=tha, ckern=ckern.view(np.complex128),
                rhs_lap=rhs_lap, ulap=ulap, phi=phi, psi=psi, dphi_dn=dphi, rhs_stokes=rhs_stokes,
                u=u.view(np.complex128), pk=pk, za=za.view(np.complex128), un=un)


def ref_manydrop(p, nsteps):
    """MANYDROP's run (electric.f:62-245) with the reference's own routines, for a run
    description p (dict: k, nd, dt, tfinal, cap, e0, ai, bi, thetax, cx, cy, ncyc; no gravity):
    GEOINIT, CLUSTPARAM, GEOSHIFT, INITCOND, then nsteps passes of the time loop (RK2,
    EXTRACTSOL, GETFOURIER, DOUBLE, the step law).  Plotting is left out.  Run it from a
    scratch directory: the reference writes fort.13.  Returns the state after the steps and
    the per-step (time, dt, total area)."""
    r = ref()
    kmax, ndmax = 150, 3000
    nmax = ndmax * kmax
    k = int(p["k"])
    sim = RefSim(LIBREF)
    sim.lib = r
    nd0 = np.asarray(p["nd"], np.int32)
    sim.setup(nd0, cap=float(p["cap"]), zbond=0j, e0=float(p["e0"]))
    prob = (ctypes.c_char * (16 * kmax + 4 * kmax + 12)).from_address(
        ctypes.addressof(ctypes.c_char.in_dll(r, "problem_")))
    c_nd = ctypes.c_void_p(ctypes.addressof(prob) + 16 * kmax)  # COMMON /problem/ nd, as MANYDROP's nd
    ndv = np.frombuffer(prob, dtype=np.int32, count=kmax, offset=16 * kmax)
    ints = np.frombuffer(prob, dtype=np.int32, count=3, offset=20 * kmax)
    f64 = lambda key: np.ascontiguousarray(np.asarray(p[key], np.float64)[:k])  # noqa: E731
    ai, bi, th, cx, cy = f64("ai"), f64("bi"), f64("thetax"), f64("cx"), f64("cy")
    ncyc = np.ascontiguousarray(np.asarray(p["ncyc"], np.int32)[:k])
    zk = (cx + 1j * cy).astype(np.complex128)
    nbk = int(nd0.sum())
    z = np.zeros((k, ndmax), np.complex128)  # Fortran z(ndmax, k)
    dsdth, rkappa = np.zeros(nmax), np.zeros(nmax)
    area0, arcl0 = np.zeros(k), np.zeros(k)
    r.geoinit_(_c_int(k), c_nd, _c_int(nbk), _d(ai), _d(bi), _d(th), _i(ncyc), _d(_cplx(zk)), _d(z), _d(dsdth),
               _d(rkappa), _c_int(10), _d(area0), _c_int(ndmax))
    ifwork = 10 * ndmax + 50
    wsave = np.zeros(2 * ifwork * kmax)
    za, zai = np.zeros(2 * nmax), np.zeros(2 * nmax)
    beta = np.zeros(nmax)
    r.clustparam_(_c_int(k), c_nd, _c_int(nbk), _d(dsdth), _d(arcl0), _d(beta), _d(za), _d(zai), _d(wsave),
                  _c_int(ifwork))
    r.geoshift_(_c_int(k), c_nd, _c_int(nbk), _d(ai), _d(bi), _d(th), _i(ncyc), _d(_cplx(zk)), _d(beta), _d(z),
                _d(rkappa), _c_int(10), _c_int(ndmax))
    dt = float(p["dt"])
    cmin = max(dt * ndv[b] / arcl0[b] for b in range(k))
    neqmax = 2 * nmax
    ybar, y, dy = np.zeros(neqmax), np.zeros(neqmax), np.zeros(neqmax)
    phi = np.zeros(4 * neqmax)
    nbk_c, neq_c = ctypes.c_int(0), ctypes.c_int(0)
    r.initcond_(_c_int(k), _c_int(ndmax), c_nd, ctypes.byref(nbk_c), ctypes.byref(neq_c), _d(z), _d(ybar))
    ints[:] = [k, nbk_c.value, neq_c.value]
    areatot0 = float(area0.sum())
    time, tfinal = 0.0, float(p["tfinal"])
    dz = np.zeros(2 * nmax)
    amode, fmode = np.zeros(nmax), np.zeros(nmax)
    z2 = np.zeros(2 * nmax)
    nzwork = 32 * nmax + 100
    zwork = np.zeros(2 * nzwork)
    area, arcl = np.zeros(k), np.zeros(k)
    hist = []
    for _ in range(nsteps):
        time += dt
        if time > tfinal:
            dt = tfinal - time + dt
            time = tfinal
        neq = int(ints[2])
        r.rk2_(_c_int(k), _c_int(neq), _c_double(time), _c_double(dt), _d(phi), _d(ybar), _d(y), _d(dy))
        icut = 0
        ist, ist2 = 0, 0
        for b in range(k):
            a_c, l_c, e_c, km_c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            zb = z[b]
            nb = ctypes.c_int(int(ndv[b]))
            r.extractsol_(ctypes.byref(nb), _c_int(neq), ctypes.byref(a_c), ctypes.byref(l_c), _d(zb),
                          ctypes.c_void_p(dz.ctypes.data + 16 * ist), ctypes.c_void_p(ybar.ctypes.data + 8 * ist2),
                          _d(za), _d(zai), ctypes.c_void_p(wsave.ctypes.data + 16 * ifwork * b),
                          ctypes.byref(e_c), ctypes.byref(km_c), _d(rkappa))
            area[b], arcl[b] = a_c.value, l_c.value
            idoub = ctypes.c_int(0)
            r.getfourier_(ctypes.byref(nb), _d(amode), _d(fmode), _d(zb), _d(za),
                          ctypes.c_void_p(wsave.ctypes.data + 16 * ifwork * b), ctypes.byref(idoub))
            if idoub.value == 1:
                icut = 1
                istop = ctypes.c_int(0)
                ndp = ctypes.c_void_p(ctypes.addressof(prob) + 16 * kmax + 4 * b)
                r.double_(ndp, _d(zb), _d(z2), _d(za), ctypes.c_void_p(wsave.ctypes.data + 16 * ifwork * b),
                          _d(zwork), _c_int(nzwork), _c_int(ndmax), ctypes.byref(istop))
                ist += int(ndv[b]) // 2
                ist2 += int(ndv[b])
            else:
                ist += int(ndv[b])
                ist2 += 2 * int(ndv[b])
        dt = min(cmin * arcl[b] / ndv[b] for b in range(k))
        if icut:
            r.initcond_(_c_int(k), _c_int(ndmax), c_nd, ctypes.byref(nbk_c), ctypes.byref(neq_c), _d(z), _d(ybar))
            ints[:] = [k, nbk_c.value, neq_c.value]
        hist.append((time, dt, float(area.sum())))
    r.prini_(_c_int(0), _c_int(0))
    neq = int(ints[2])
    return dict(ybar=ybar[:neq].copy(), nd=ndv[:k].copy(), time=time, dt=dt, hist=np.array(hist),
                area0=area0, arcl0=arcl0, areatot0=areatot0, cmin=cmin)
