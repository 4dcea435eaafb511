"""Write the golden fixtures of tests/golden/ from the compiled reference (oracle/_ref/libref.so).

Runs only where the reference sources exist (they are compiled in place by oracle/Makefile);
the fixtures it writes are data -- inputs and the reference's outputs -- and are committed so
the tests run without the reference.  Usage:  python tests/golden/make_golden.py [atree|gmres|sweep|field|sim|adjoint|adjoint16|dapif2]
("atree" / "gmres" / "sweep" / "field" / "sim" / "adjoint" / "adjoint16" / "dapif2" rewrite only those
fixtures; "adjoint16" runs the reference's F some 680 times, in parallel worker processes).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from electricbubbles_amd.workloads import curves_cfg, uniform_cfg, perturbed_circle  # noqa: E402


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", os.path.relpath(path, ROOT), os.path.getsize(path), "bytes")


def fmm_cases():
    rng = np.random.default_rng(20260101)
    cases = []
    # cfg1-like single drop r = 1 + 0.05 cos 2theta, 256 points, LEVMAX 7, TOL 1e-13
    z, dz, nd, h = perturbed_circle(256, [(2, 0.05)])
    cases.append(("drop256", z, dz, nd, h, 7, 1e-13))
    # many small curves (the cfg3 shape at small scale)
    z, dz, nd, h = curves_cfg(16, 64, rng)
    cases.append(("curves1024", z, dz, nd, h, 5, 1e-13))
    z, dz, nd, h = curves_cfg(16, 256, rng)
    cases.append(("curves4096", z, dz, nd, h, 6, 1e-13))
    z, dz, nd, h = uniform_cfg(4096, rng)
    cases.append(("uniform4096", z, dz, nd, h, 6, 1e-10))
    return cases, rng


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (needs the reference sources)")
    cases, rng = fmm_cases()
    for name, z, dz, nd, h, lev, tol in cases:
        n = z.size
        dens = (rng.normal(size=n) + 1j * rng.normal(size=n)) / np.sqrt(2.0)
        out = {}
        for var, tag in ((oracle.DR, "dr"), (oracle.MOD, "mod"), (oracle.PLUS, "plus")):
            gw, inform, ier = oracle.ref_cadap(var, lev, nd, z.real, z.imag, dz, dens, h, tol)
            assert ier[0] == 0
            out["gw_" + tag] = gw
            out["inform"] = inform
        save("cadap_" + name, x=z.real, y=z.imag, dzdt=dz, dens=dens, nd=nd, h=h,
             levmax=np.int32(lev), tol=np.float64(tol), **out)

    # ASSIGN1 tables and lists
    z, dz, nd, h = curves_cfg(20, 100, rng)
    zs, rs, rmin = oracle.scale(z.real, z.imag)
    tab = {"x": z.real, "y": z.imag, "zs": zs, "rscal": np.float64(rs), "rmin": rmin}
    for lev in range(1, 8):
        t = oracle.ref_assign1(lev, zs)
        for k in ("ibox", "ioffst", "iatadr", "iloc"):
            tab[f"{k}_{lev}"] = t[k]
        tab[f"ntbox_{lev}"] = np.int32(t["ntbox"])
        tab[f"nsmax_{lev}"] = np.int32(t["nsmax"])
    save("assign1_curves2000", **tab)
    lst = {}
    for lev in range(1, 6):
        intl, nbor = [], []
        for b in range(1, 4 ** lev + 1):
            a = oracle.ref_mkintl(lev, b)
            intl.append(np.pad(a, (0, 27 - a.size), constant_values=0))
            c = oracle.ref_mknbor(lev, b)
            nbor.append(np.pad(c, (0, 9 - c.size), constant_values=0))
        lst[f"intl_{lev}"] = np.array(intl, np.int32)
        lst[f"nbor_{lev}"] = np.array(nbor, np.int32)
    save("lists_l1to5", **lst)

    # workspace / precision failures (cadapdrop.f:126-133, 191-197, 283-290, 312-319)
    z, dz, nd, h = curves_cfg(4, 64, rng)
    n = z.size
    dens = (rng.normal(size=n) + 1j * rng.normal(size=n)) / np.sqrt(2.0)
    err = {"x": z.real, "y": z.imag, "dzdt": dz, "dens": dens, "nd": nd, "h": h}
    for tag, lev, tol, lenw, leniw in (("tol", 5, 1e-19, 100000, 100000),
                                       ("lenw_p", 5, 1e-13, 8 * n, 100000),
                                       ("leniw", 5, 1e-13, 100000, 5 * 4 ** 3 + 2 * n + 3),
                                       ("lenw_c", 5, 1e-13, 8 * n + 1 + 200, 100000)):
        gw, inform, ier = oracle.ref_cadap(oracle.DR, lev, nd, z.real, z.imag, dz, dens, h, tol, lenw=lenw,
                                           leniw=leniw)
        err[f"ier_{tag}"] = ier
        err[f"inform_{tag}"] = inform
        err[f"args_{tag}"] = np.array([lev, lenw, leniw], np.int64)
        err[f"tol_{tag}"] = np.float64(tol)
    save("cadap_errors", **err)

    # spectral pieces: DCFFTF/DCFFTB/FDIFFF on several lengths, FILTER, EQUIT, EXTRACTDAT
    sp = {}
    for n in (8, 12, 30, 64, 100, 256, 1234):
        c = rng.normal(size=n) + 1j * rng.normal(size=n)
        sp[f"c_{n}"] = c
        sp[f"fwd_{n}"] = oracle.ref_dcfftf(c)
        sp[f"bwd_{n}"] = oracle.ref_dcfftb(c)
        sp[f"dif_{n}"] = oracle.ref_fdifff(c)
    for n in (64, 256):
        cn = (rng.normal(size=n) + 1j * rng.normal(size=n)) * np.exp(-0.05 * np.minimum(np.arange(n), n - np.arange(n)))
        cn[np.abs(cn) < 1e-12] += 3e-14
        sp[f"filt_in_{n}"] = cn
        sp[f"filt_out_{n}"] = oracle.ref_filter(cn)
        tha = rng.normal(size=n)
        un = rng.normal(size=n)
        sp[f"equit_tha_{n}"] = tha
        sp[f"equit_un_{n}"] = un
        sp[f"equit_t_{n}"] = oracle.ref_equit(tha, un)
    save("spectral", **sp)

    z1, _, _, _ = perturbed_circle(128, [(2, 0.05), (3, 0.02)], center=0.3 - 0.1j)
    z2, _, _, _ = perturbed_circle(96, [(2, -0.03)], center=3.0 + 1.0j, radius=0.7)
    nd = np.array([128, 96], np.int32)
    rku = np.concatenate([z1, z2]).view(np.float64).copy()
    ex = oracle.ref_extractdat(nd, rku)
    u = rng.normal(size=2 * int(nd.sum()))
    w, stress, za, pk = oracle.ref_fasmvp(nd, ex["x"], ex["y"], ex["dz"], ex["dsda"], ex["rkappa"], ex["zk"], u)
    save("geometry_two_drops", nd=nd, rku=rku, x=ex["x"], y=ex["y"], zk=ex["zk"], dz=ex["dz"], tha=ex["tha"],
         dsda=ex["dsda"], rkappa=ex["rkappa"], ckern=ex["ckern"], xleft=np.float64(ex["xleft"]),
         slength=np.float64(ex["slength"]), u=u, w=w, stress=stress, za=za, pk=pk)


def atree_cases():
    """Point sets for the adaptive tree of DAPIF2: uniform (no list 3/4), curves and a tight
    cluster in a sparse background (deep, uneven tree: long lists 3 and 4)."""
    rng = np.random.default_rng(20260202)
    cases = []
    z, _, _, _ = curves_cfg(16, 128, rng)
    cases.append(("curves2048_napb30", z.real, z.imag, 30))
    z, _, _, _ = uniform_cfg(5000, rng)
    cases.append(("uniform5000_napb40", z.real, z.imag, 40))
    c = np.concatenate([rng.normal(0.3, 0.01, (1500, 2)), rng.uniform(0, 1, (1500, 2))])
    cases.append(("cluster3000_napb40", c[:, 0].copy(), c[:, 1].copy(), 40))
    z, _, _, _ = curves_cfg(64, 256, rng)
    cases.append(("curves16384_napb40", z.real, z.imag, 40))
    return cases


def main_atree():
    """DAPIF2's scaled coordinates (from its workspace) and the tables of DAASSI, DACOMP,
    DAINDB, DAIJAT, DALI12 and DALI34 run on them."""
    for name, x, y, napb in atree_cases():
        w = oracle.ref_dapif2_scaled(x, y, napb)
        assert w["ierr"] == 0
        t = oracle.ref_atree(w["xs"], w["ys"], napb)
        assert t["ierr"] == 0 and t["nbox"] == w["nbox"] and t["nlev"] == w["nlev"]
        counts = np.array([t[k] for k in ("nbox", "nlev", "len", "len3", "len4", "maxp")], np.int32)
        save("atree_" + name, x=x, y=y, napb=np.int32(napb), lmx=np.int32(oracle.dapif_lmx()),
             eps7=np.float64(1e-16), counts=counts, **{k: t[k] for k in oracle.ATREE_TABLES})


def gmres_dense_cases():
    """Dense systems for DGMRES: well conditioned with restarts, non-normal, one solve
    without restarts, and a hard one that stops at the restart limit (IERR = 1)."""
    rng = np.random.default_rng(20260404)
    out = []
    for name, n, kind, maxl, nrmax in (("well50_maxl10", 50, "well", 10, 0), ("nonnormal80_maxl20", 80, "nn", 20, 0),
                                        ("well60_maxl100_norestart", 60, "well", 100, -1),
                                        ("hard200_maxl15_nrmax3", 200, "hard", 15, 3)):
        if kind == "well":
            a = np.eye(n) + 0.3 * rng.normal(size=(n, n)) / np.sqrt(n)
        elif kind == "nn":
            a = np.eye(n) + 0.2 * np.triu(rng.normal(size=(n, n)), 1) + 0.1 * rng.normal(size=(n, n)) / np.sqrt(n)
        else:
            a = np.diag(np.linspace(0.01, 1, n)) + 0.05 * rng.normal(size=(n, n)) / np.sqrt(n)
        b = rng.normal(size=n)
        x0 = 0.1 * rng.normal(size=n)
        out.append((name, a, b, x0, maxl, nrmax))
    return out


def stokes_16drops():
    """16 perturbed drops on a 4x4 grid (spacing 3), N per drop 128..512: above the direct-sum
    size, so the FMM path is the operator."""
    rng = np.random.default_rng(20260405)
    zs, nds = [], []
    for k in range(16):
        n = int(rng.choice([128, 192, 256, 384, 512]))
        c = 3.0 * (k % 4) + 3.0j * (k // 4)
        modes = [(m, rng.uniform(-0.04, 0.04)) for m in (2, 3, 5)]
        z, _, _, _ = perturbed_circle(n, modes, center=c, radius=rng.uniform(0.8, 1.2))
        zs.append(z)
        nds.append(n)
    return np.array(nds, np.int32), np.concatenate(zs).view(np.float64).copy()


def main_gmres():
    """DGMRES of the reference on dense matrices, and SOLVEI's solve (DGMRES with the
    reference MATVEC: FASMVP + CADAPDR) on two geometries."""
    cases = gmres_dense_cases()
    d = {}
    for name, a, b, x0, maxl, nrmax in cases:
        x, it, err, ierr = oracle.ref_dgmres_dense(a, b, x0, maxl, 1e-10, nrmax)
        d.update({f"{name}_a": a, f"{name}_b": b, f"{name}_x0": x0, f"{name}_x": x,
                  f"{name}_meta": np.array([maxl, nrmax, it, ierr], np.int32), f"{name}_err": np.float64(err)})
    d["names"] = np.array([c[0] for c in cases])
    save("gmres_dense", **d)
    import tempfile
    cwd = os.getcwd()
    os.chdir(tempfile.mkdtemp())  # the reference writes fort.13
    try:
        sim = oracle.RefSim(oracle.LIBREF)
        for tag in ("two_drops", "16drops"):
            if tag == "two_drops":
                t = np.load(os.path.join(HERE, "geometry_two_drops.npz"))
                nd, rku = t["nd"], t["rku"]
            else:
                nd, rku = stokes_16drops()
            ex = oracle.ref_extractdat(nd, rku)
            th = np.concatenate([2 * np.pi * np.arange(m) / m for m in nd])
            u_true = np.stack([np.cos(th) + 0.3 * np.sin(2 * th), 0.5 * np.sin(th) - 0.2 * np.cos(3 * th)], 1).ravel()
            sim.setup(nd)
            rhs = oracle.ref_fasmvp(nd, ex["x"], ex["y"], ex["dz"], ex["dsda"], ex["rkappa"], ex["zk"], u_true)[0]
            z = ex["x"] + 1j * ex["y"]
            x, it, err, ierr = oracle.ref_dgmres_stokes(nd, z, ex["dz"], ex["dsda"], ex["rkappa"], ex["zk"], rhs)
            _, _, za, pk = oracle.ref_fasmvp(nd, ex["x"], ex["y"], ex["dz"], ex["dsda"], ex["rkappa"], ex["zk"], x)
            print(tag, "reference GMRES iterations", it, "err", err, "ierr", ierr)
            save("gmres_stokes_" + tag, nd=nd, rku=rku, rhs=rhs, u_true=u_true, x=x,
                 meta=np.array([it, ierr], np.int32), err=np.float64(err), za=za, pk=pk)
    finally:
        os.chdir(cwd)


def main_sweep():
    """SOLVEI's solve by the reference for several single-drop simulations, each solved on its
    own, as the reference runs one simulation per process: the batched sweep's parity data."""
    import tempfile
    from electricbubbles_amd.workloads import perturbed_circle
    cwd = os.getcwd()
    os.chdir(tempfile.mkdtemp())  # the reference writes fort.13
    try:
        sim = oracle.RefSim(oracle.LIBREF)
        nds = [128, 130, 256, 322, 512, 578]
        extra = [(3, 0.02), (4, -0.03), (5, 0.01), (3, -0.04), (6, 0.02), (2, 0.03)]
        out = {k: [] for k in ("rku", "rhs", "u_true", "x", "za", "pk")}
        meta, errs = [], []
        for n, mode in zip(nds, extra):
            nd = np.array([n], np.int32)
            rku = perturbed_circle(n, [(2, 0.05), mode])[0].view(np.float64).copy()
            ex = oracle.ref_extractdat(nd, rku)
            th = 2 * np.pi * np.arange(n) / n
            u_true = np.stack([np.cos(th) + 0.3 * np.sin(2 * th), 0.5 * np.sin(th) - 0.2 * np.cos(3 * th)], 1).ravel()
            sim.setup(nd)
            rhs = oracle.ref_fasmvp(nd, ex["x"], ex["y"], ex["dz"], ex["dsda"], ex["rkappa"], ex["zk"], u_true)[0]
            z = ex["x"] + 1j * ex["y"]
            x, it, err, ierr = oracle.ref_dgmres_stokes(nd, z, ex["dz"], ex["dsda"], ex["rkappa"], ex["zk"], rhs)
            _, _, za, pk = oracle.ref_fasmvp(nd, ex["x"], ex["y"], ex["dz"], ex["dsda"], ex["rkappa"], ex["zk"], x)
            print("sweep drop", n, "reference GMRES iterations", it, "err", err, "ierr", ierr)
            for k, v in (("rku", rku), ("rhs", rhs), ("u_true", u_true), ("x", x), ("za", za), ("pk", pk)):
                out[k].append(np.asarray(v))
            meta.append([it, ierr])
            errs.append(err)
        save("gmres_sweep", nd=np.array(nds, np.int32), meta=np.array(meta, np.int32), err=np.array(errs),
             **{k: np.concatenate(v) for k, v in out.items()})
    finally:
        os.chdir(cwd)


def main_field():
    """The reference's F stage by stage (oracle.ref_field_chain), its whole F (RefSim.rhs) and a
    few RK2 steps, on three geometries: the single drop of the plain method (r = 1 + 0.05 cos
    2 theta, N = 256), two drops of different N, and the four unit circles of bubbin."""
    import tempfile
    from electricbubbles_amd.workloads import perturbed_circle
    cwd = os.getcwd()
    os.chdir(tempfile.mkdtemp())  # the reference writes fort.13
    try:
        cases = []
        cases.append(("single_drop", [256], [0j], [(2, 0.05)], 0.1, 1.0))
        cases.append(("two_drops", [128, 160], [0j, 3.0 + 0.5j], [(2, 0.05), (3, 0.03)], 0.0, 0.64))
        cases.append(("four_circles", [128] * 4, [1.5 + 1.5j, -1.5 + 1.5j, -1.5 - 1.5j, 1.5 - 1.5j], [(2, 0.0)], 0.0,
                      0.64))
        # under-resolved bodies (content at mode N/2 - 1, where PVINTEV's FINTER differs from
        # FINTERC, pv_integral.f:236-237), off the origin, with the straining flow on
        cases.append(("underresolved", [32, 48], [0.3 - 0.2j, 2.8 + 0.4j], [(5, 0.1), (14, 0.02)], 0.05, 0.5))
        only = os.environ.get("EB_GOLDEN_FIELD")
        if only:
            cases = [c for c in cases if c[0] == only]
        for tag, nds, centres, modes, cap, e0 in cases:
            nd = np.array(nds, np.int32)
            rku = np.concatenate([perturbed_circle(int(n), modes, center=c)[0] for n, c in zip(nds, centres)])
            rku = rku.view(np.float64).copy()
            ch = oracle.ref_field_chain(nd, rku, cap=cap, e0=e0)
            sim = oracle.RefSim(oracle.LIBREF)
            sim.setup(nd, cap=cap, e0=e0)
            rkup = sim.rhs(rku)
            # FASMVP_LAP on a smooth density
            nbk = int(nd.sum())
            th = np.concatenate([2 * np.pi * np.arange(m) / m for m in nd])
            ut = np.cos(th) + 0.3 * np.sin(3 * th)
            w = np.zeros(nbk + 150)
            ak = np.zeros(150)
            r = oracle.ref()
            r.fasmvp_lap_(oracle._c_int(nd.size), oracle._i(nd), oracle._c_int(nbk), oracle._d(oracle._cplx(ch["zk"])),
                          oracle._d(oracle._cplx(ch["z"])), oracle._d(oracle._cplx(ch["dz"])),
                          oracle._d(ch["rkappa"]), oracle._d(ch["dsda"]), oracle._d(ak), oracle._d(ut), oracle._d(w))
            r.prini_(oracle._c_int(0), oracle._c_int(0))
            dt, nsteps = 1e-3, 3
            sim.setup(nd, cap=cap, e0=e0)
            traj = sim.rk2(rku, dt, nsteps)
            print(tag, "F norm", np.linalg.norm(rkup), "Un max", np.abs(ch["un"]).max())
            save("field_" + tag, nd=nd, rku=rku, cap=np.float64(cap), e0=np.float64(e0), rkup=rkup, lap_u=ut,
                 lap_w=w[:nbk].copy(), rk2=traj, rk2_dt=np.float64(dt), rk2_steps=np.int32(nsteps),
                 **{k: v for k, v in ch.items() if k in ("rhs_lap", "ulap", "phi", "psi", "dphi_dn", "rhs_stokes",
                                                         "u", "za", "pk", "un")})
        if only:
            return
        # a many-drop configuration (8 drops of 256..1024 points, E = (1.5, 0)): F and two RK2
        # steps at the configuration's dt
        from electricbubbles_amd.workloads import many_drops_cfg
        nd, rku = many_drops_cfg(8, np.random.default_rng(20260508))
        sim = oracle.RefSim(oracle.LIBREF)
        sim.setup(nd, cap=0.0, e0=2.25)
        rkup = sim.rhs(rku)
        sim.setup(nd, cap=0.0, e0=2.25)
        traj = sim.rk2(rku, 5e-4, 2)
        print("many_drops_8 points", int(nd.sum()), "F norm", np.linalg.norm(rkup))
        save("field_many_drops_8", nd=nd, rku=rku, cap=np.float64(0.0), e0=np.float64(2.25), rkup=rkup, rk2=traj,
             rk2_dt=np.float64(5e-4), rk2_steps=np.int32(2))
    finally:
        os.chdir(cwd)


def main_sim():
    """MANYDROP's run with the reference's routines (oracle.ref_manydrop): the initial shapes
    after CLUSTPARAM and GEOSHIFT, and the state, step sizes and areas after a few passes of
    the time loop, for the four circles of bubbin, an ellipse and a starfish."""
    import tempfile
    cwd = os.getcwd()
    os.chdir(tempfile.mkdtemp())  # the reference writes fort.13
    try:
        cases = {
            "four_circles": (dict(k=4, nd=[128] * 4, dt=0.01, tfinal=0.5, nplot=10, cap=0.0, e0=0.64, ai=[1.0] * 4,
                                  bi=[1.0] * 4, thetax=[0.0] * 4, cx=[-1.5, -1.5, 1.5, 1.5],
                                  cy=[1.5, -1.5, 1.5, -1.5], ncyc=[0] * 4), 5),
            "ellipse": (dict(k=1, nd=[256], dt=1e-3, tfinal=1.0, nplot=10, cap=0.0, e0=1.0, ai=[1.1],
                             bi=[1.0 / 1.1], thetax=[0.3], cx=[0.2], cy=[-0.1], ncyc=[0]), 10),
            "starfish": (dict(k=1, nd=[256], dt=5e-4, tfinal=1.0, nplot=10, cap=0.0, e0=0.5, ai=[1.0], bi=[0.1],
                              thetax=[0.0], cx=[0.0], cy=[0.0], ncyc=[3]), 4),
            # a five-lobed star on 32 points beside an ellipse: GETFOURIER's test (modes n-10..n-1
            # above 1e-11, electric.f:560-566) fires on the star in each of the first three steps,
            # and DOUBLE (electric.f:721-757) takes it to 64, 128 and 256 points
            "doubling": (dict(k=2, nd=[32, 64], dt=1e-3, tfinal=1.0, nplot=10, cap=0.05, e0=0.5, ai=[1.0, 0.8],
                              bi=[0.1, 0.8], thetax=[0.0, 0.2], cx=[0.0, 3.0], cy=[0.0, 0.5], ncyc=[5, 0]), 4),
        }
        only = os.environ.get("EB_GOLDEN_SIM")
        if only:
            cases = {only: cases[only]}
        for tag, (p, nsteps) in cases.items():
            r0 = oracle.ref_manydrop(p, 0)
            r1 = oracle.ref_manydrop(p, nsteps)
            # body lengths after each step (the doubling history)
            nds = np.array([oracle.ref_manydrop(p, s)["nd"] for s in range(1, nsteps + 1)], np.int32)
            print(tag, "steps", nsteps, "dt", r1["dt"], "area", r1["hist"][-1], "nd", nds.tolist())
            save("sim_" + tag, nsteps=np.int32(nsteps), ybar0=r0["ybar"], ybar=r1["ybar"], nd_after=r1["nd"],
                 nd_steps=nds, hist=r1["hist"], arcl0=r1["arcl0"], area0=r1["area0"], cmin=np.float64(r1["cmin"]),
                 **{k: np.asarray(v) for k, v in p.items()})
    finally:
        os.chdir(cwd)


def _np_fdifff(z):
    """FDIFFF (dcfft.f:1371-1427) per row: i k FFT / N with k = 0 and Nyquist zeroed."""
    N = z.shape[1]
    k = np.fft.fftfreq(N, 1.0 / N)
    k[N // 2] = 0.0
    return np.fft.ifft(1j * k * np.fft.fft(z, axis=1), axis=1)


def np_deformation(z):
    """D = (L - B) / (L + B) per body from the area and second moments of area (the definition
    the adjoint's loss uses)."""
    N = z.shape[1]
    h = 2 * np.pi / N
    za = _np_fdifff(z)
    x, y, xa, ya = z.real, z.imag, za.real, za.imag
    A = 0.5 * h * (x * ya - y * xa).sum(1)
    cx = 0.5 * h * (x * x * ya).sum(1) / A
    cy = -0.5 * h * (y * y * xa).sum(1) / A
    a = h * (x ** 3 * ya).sum(1) / 3 - A * cx * cx
    c = -h * (y ** 3 * xa).sum(1) / 3 - A * cy * cy
    b = 0.5 * h * (x * x * y * ya).sum(1) - A * cx * cy
    rt = np.sqrt(0.25 * (a - c) ** 2 + b * b)
    sp, sm = np.sqrt(0.5 * (a + c) + rt), np.sqrt(0.5 * (a + c) - rt)
    return (sp - sm) / (sp + sm)


def np_shapes(centres, modes, N, decay=0.0):
    """z = c + (1 + sum_m m^-decay (a_m cos m th + b_m sin m th)) e^{i th} (the adjoint's shapes)."""
    M = modes.shape[1] // 2
    th = 2 * np.pi * np.arange(N) / N
    m = np.arange(1, M + 1)
    w = m.astype(np.float64) ** (-float(decay))
    r = 1.0 + modes[:, :M] @ (w[:, None] * np.cos(np.outer(m, th))) + modes[:, M:] @ (w[:, None] * np.sin(np.outer(m, th)))
    return centres[:, None] + r * np.exp(1j * th)[None, :]


def _ref_rk4_loss(args):
    """Summed deformation after `steps` RK4 steps whose F is the reference's own (RefSim.rhs), run
    in a scratch directory of its own (the reference writes fort.13): one worker process."""
    import tempfile
    centres, modes, E_, Ca_, N, steps, dt, decay = args
    os.chdir(tempfile.mkdtemp())
    nd = np.full(centres.size, N, np.int32)
    sim = oracle.RefSim(oracle.LIBREF)
    sim.setup(nd, cap=Ca_, e0=E_ * E_)
    z = np_shapes(centres, modes, N, decay)

    def F(zz):
        y = np.stack([zz.real.ravel(), zz.imag.ravel()], 1).ravel()
        d = sim.rhs(y)
        return (d[0::2] + 1j * d[1::2]).reshape(zz.shape)
    for _ in range(steps):
        k1 = F(z)
        k2 = F(z + 0.5 * dt * k1)
        k3 = F(z + 0.5 * dt * k2)
        k4 = F(z + dt * k3)
        z = z + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return float(np_deformation(z).sum())


def main_adjoint16():
    """The adjoint configuration at its size: 16 drops of 512 points on a 4 x 4 grid (spacing 4),
    64 shape modes per drop with amplitudes U[0, 0.02] weighted 1/m^2, |E| = 1, Ca = 0.1, 10 RK4
    steps at dt = 1e-3 with the reference's own F.  Richardson-extrapolated central differences
    (h = 1e-4) of the summed deformation in |E|, Ca and two random unit directions of all 1026
    parameters; the forwards run in parallel worker processes."""
    import multiprocessing as mp
    # adjoint.grid_problem(4, 512, 4.0, 64, 0.02, seed=5): the same draws, on the host
    rng = np.random.default_rng(5)
    g = (np.arange(4) - 1.5) * 4.0
    cx, cy = np.meshgrid(g, g, indexing="xy")
    centres = (cx + 1j * cy).ravel()
    modes = rng.uniform(0.0, 0.02, (16, 64))
    E, Ca, dt, steps, N, decay, h = 1.0, 0.1, 1e-3, 10, 512, 2.0, 1e-4
    rng = np.random.default_rng(16)
    dirs = rng.standard_normal((2, 2 + modes.size))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    pert = [(1.0, 0.0, np.zeros_like(modes)), (0.0, 1.0, np.zeros_like(modes))]
    pert += [(v[0], v[1], v[2:].reshape(modes.shape)) for v in dirs]
    jobs = [(centres, modes, E, Ca, N, steps, dt, decay)]
    for dE, dCa, dm in pert:
        for e in (h, -h, h / 2, -h / 2):
            jobs.append((centres, modes + e * dm, E + e * dE, Ca + e * dCa, N, steps, dt, decay))
    with mp.get_context("spawn").Pool(min(8, os.cpu_count() or 1)) as pool:
        vals = pool.map(_ref_rk4_loss, jobs)
    L0, rest = vals[0], np.array(vals[1:]).reshape(len(pert), 4)
    fd = (4 * (rest[:, 2] - rest[:, 3]) / h - (rest[:, 0] - rest[:, 1]) / (2 * h)) / 3
    print("adjoint16 fixture: loss", L0, "fd", fd)
    save("adjoint_fd_16drops", centres=centres, modes=modes, E=np.float64(E), Ca=np.float64(Ca), dt=np.float64(dt),
         steps=np.int32(steps), N=np.int32(N), decay=np.float64(decay), h=np.float64(h), loss=np.float64(L0),
         fd_E=np.float64(fd[0]), fd_Ca=np.float64(fd[1]), dirs=dirs, fd_dirs=fd[2:])


def main_adjoint():
    """Central differences (Richardson-extrapolated) of the summed deformation after a few RK4
    steps whose F is the reference's own (RefSim.rhs): four drops of 64 points on a 2 x 2 grid, eight shape modes
    each; E0 = |E|^2, CAP = Ca.  The adjoint's gradient is checked against these."""
    import tempfile
    cwd = os.getcwd()
    os.chdir(tempfile.mkdtemp())  # the reference writes fort.13
    try:
        k_side, N, nm, spacing, amp = 2, 64, 8, 3.0, 0.03
        rng = np.random.default_rng(2)
        g = (np.arange(k_side) - 0.5 * (k_side - 1)) * spacing
        cx, cy = np.meshgrid(g, g, indexing="xy")
        centres = (cx + 1j * cy).ravel()
        modes = rng.uniform(0.0, amp, (k_side * k_side, nm))
        E, Ca, dt, steps = 1.0, 0.1, 1e-3, 4
        nd = np.full(k_side * k_side, N, np.int32)
        sim = oracle.RefSim(oracle.LIBREF)

        def loss(E_, Ca_, modes_):
            sim.setup(nd, cap=Ca_, e0=E_ * E_)
            z = np_shapes(centres, modes_, N)

            def F(zz):
                y = np.stack([zz.real.ravel(), zz.imag.ravel()], 1).ravel()
                d = sim.rhs(y)
                return (d[0::2] + 1j * d[1::2]).reshape(zz.shape)
            for _ in range(steps):
                k1 = F(z)
                k2 = F(z + 0.5 * dt * k1)
                k3 = F(z + 0.5 * dt * k2)
                k4 = F(z + dt * k3)
                z = z + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
            return float(np_deformation(z).sum())

        h = 1e-4

        def rich(fn):
            # central differences at h and h/2, Richardson-extrapolated (error O(h^4))
            d1 = (fn(h) - fn(-h)) / (2 * h)
            d2 = (fn(h / 2) - fn(-h / 2)) / h
            return (4 * d2 - d1) / 3

        L0 = loss(E, Ca, modes)
        fd_E = rich(lambda e: loss(E + e, Ca, modes))
        fd_Ca = rich(lambda e: loss(E, Ca + e, modes))
        dirs = rng.standard_normal((3,) + modes.shape)
        dirs /= np.linalg.norm(dirs.reshape(3, -1), axis=1)[:, None, None]
        fd_dirs = np.array([rich(lambda e: loss(E, Ca, modes + e * v)) for v in dirs])
        print("adjoint fixture: loss", L0, "dE", fd_E, "dCa", fd_Ca, "dirs", fd_dirs)
        save("adjoint_fd_four_drops", centres=centres, modes=modes, E=np.float64(E), Ca=np.float64(Ca),
             dt=np.float64(dt), steps=np.int32(steps), N=np.int32(N), h=np.float64(h), loss=np.float64(L0),
             fd_E=np.float64(fd_E), fd_Ca=np.float64(fd_Ca), dirs=dirs, fd_dirs=fd_dirs)
    finally:
        os.chdir(cwd)


def main_dapif2():
    """DAPIF2's Hilbert-matrix product (IFLAG7 = 3) as laplace.f:204-215 and pv_integral.f:171-182
    call it (NAPB 30, NINIRE 2, MEX 300, EPS7 1e-16): FIELD, IERR and INFORM of the reference on
    three point sets -- closed curves carrying u dz h, as FASMVP_LAP forms them (TOL 1e-11), the
    alternating-weight points of PVINTEV (TOL 1e-10), and uniform points with complex charges --
    plus its error returns (TOL too small, NAT <= NAPB, workspace too short)."""
    rng = np.random.default_rng(20261204)
    cwd = os.getcwd()
    os.chdir(os.path.join(ROOT, "gpurun_out") if os.path.isdir(os.path.join(ROOT, "gpurun_out")) else "/tmp")
    try:
        out = {}
        z, dz, nd, h = curves_cfg(24, 128, rng)
        u = rng.normal(size=z.size)
        cases = [("lap", z, u * dz * np.repeat(h, nd), 1e-11)]
        z2, dz2, nd2, h2 = curves_cfg(12, 256, rng)
        q2 = np.where(np.arange(z2.size) % 2 == 1, rng.normal(size=z2.size) * np.repeat(h2, nd2), 0.0)
        cases.append(("pv", z2, q2.astype(complex), 1e-10))
        z3 = rng.uniform(-1, 1, 3000) + 1j * rng.uniform(-1, 1, 3000)
        q3 = (rng.normal(size=3000) + 1j * rng.normal(size=3000)) / np.sqrt(2)
        cases.append(("uniform", z3, q3, 1e-11))
        for tag, zz, qq, tol in cases:
            field, ierr, inform = oracle.ref_dapif2(zz.real, zz.imag, qq, napb=30, tol=tol)
            assert ierr[0] == 0, (tag, ierr)
            out.update({f"x_{tag}": zz.real, f"y_{tag}": zz.imag, f"qa_{tag}": qq, f"tol_{tag}": np.float64(tol),
                        f"field_{tag}": field, f"ierr_{tag}": ierr, f"inform_{tag}": inform})
        # error returns, on the "lap" points
        z, q = cases[0][1], cases[0][2]
        need = int(out["inform_lap"][3])  # INFORM(4): the workspace the call uses
        errs = {"tol": dict(tol=1e-18), "few": dict(napb=5000), "nsp": dict(nsp=2 * z.size - 1),
                "nsp_last": dict(nsp=need - 2), "nsp_half": dict(nsp=need // 2),
                "nsp_tables": dict(nsp=need // 8)}
        for tag, kw in errs.items():
            args = dict(napb=30, tol=1e-11)
            args.update(kw)
            sel = slice(None)
            _, ierr, inform = oracle.ref_dapif2(z[sel].real, z[sel].imag, q[sel], **args)
            assert ierr[0] != 0, tag
            out[f"err_{tag}_ierr"] = ierr
            out[f"err_{tag}_inform"] = inform
            out[f"err_{tag}_args"] = np.array([args["napb"], args.get("nsp", 0)], np.int64)
            out[f"err_{tag}_tol"] = np.float64(args["tol"])
            print("dapif2 error case", tag, ierr, inform)
        save("dapif2_hilbert", **out)
    finally:
        os.chdir(cwd)


if __name__ == "__main__":
    if sys.argv[1:] == ["adjoint"]:
        main_adjoint()
    elif sys.argv[1:] == ["dapif2"]:
        main_dapif2()
    elif sys.argv[1:] == ["adjoint16"]:
        main_adjoint16()
    elif sys.argv[1:] == ["atree"]:
        main_atree()
    elif sys.argv[1:] == ["gmres"]:
        main_gmres()
    elif sys.argv[1:] == ["sweep"]:
        main_sweep()
    elif sys.argv[1:] == ["field"]:
        main_field()
    elif sys.argv[1:] == ["sim"]:
        main_sim()
    else:
        main()
        main_atree()
        main_gmres()
        main_sweep()
        main_field()
        main_sim()
        main_dapif2()
