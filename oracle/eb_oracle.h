/*
 * eb_oracle.h -- CPU restatement of the reference algorithm on the Stokes-matvec path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (electricbubbles_amd/, include/) may
 * link, load or call this library; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Conventions follow the reference Fortran: complex arrays are interleaved (re, im)
 * doubles (COMPLEX*16), index tables are 1-based, bodies are described by nd[k] points
 * and a per-body quadrature spacing h[k].
 */
#ifndef EB_ORACLE_H
#define EB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel variants of the cadap family.
 *   ORA_DR  : CADAPDR  (cadapdrop.f)  qb = (-conj(s) dz + s conj(dz)) h/(2 pi i),  GW = phi - z conj(phi') - conj(psi)
 *   ORA_MOD : CADAPMOD (cadapmod.f)   same qb,                                    GW = phi + z conj(phi') + conj(psi)
 *   ORA_PLUS: CADAP    (cadap.f)      qb = (+conj(s) dz + s conj(dz)) h/(2 pi i),  GW = phi + z conj(phi') + conj(psi) */
enum { ORA_DR = 0, ORA_MOD = 1, ORA_PLUS = 2 };

/* cadapdrop.f:137-171,199-205: bounding box, RSCAL = 64/max(range), scaled points. */
void ora_scale(int n, const double *x, const double *y, double *zs, double *rscal, double *rmin);

/* cadap2.f:252-409 (sources only, IFLAG1=0).  icnt/iloc hold 4^ilev entries, ioffst 4^ilev+1. */
void ora_assign1(int ilev, int n, const double *zs, int *ibox, int *icnt, int *ioffst,
                 int *iatadr, int *iloc, int *ntbox, int *nsmax);
/* cadap2.f:780-828 */
void ora_mkintl(int ilev, int ibox, int *intact, int *nintac);
/* cadap2.f:705-742 */
void ora_mknbor(int ilev, int ibox, int *nbors, int *nnbors);

/* cadapdrop.f:3-387 / cadapmod.f / cadap.f with IFLAG1=0, IFLAG2=0 (free space, evaluation at
 * the sources).  counts (may be NULL) receives {#M2L shifts, #M2P evaluations, #L2P point
 * evaluations, #P2P ordered pairs}.  Returns IER(1). */
int ora_cadap(int variant, int levmax, int kbod, const int *nd, const double *x, const double *y,
              const double *dzdt, const double *dens, const double *h, double tol,
              double *gradw, int *inform, int *ier, long long *counts);

/* O(N^2) direct sum of the same kernel in the original coordinates, long double accumulation.
 * Evaluates at the ntgt target indices tgt[] (0-based); sources are all points except the target.
 * Targets are split over ORA_THREADS host threads (default: the online cores); each target's
 * sum runs in source order, so the result does not depend on the thread count. */
void ora_direct(int variant, int kbod, const int *nd, const double *x, const double *y,
                const double *dzdt, const double *dens, const double *h,
                int ntgt, const int *tgt, double *gradw);

/* The transpose of the real-linear map dens -> GW of ora_direct (variant ORA_DR / ORA_MOD /
 * ORA_PLUS) or of the Hilbert product GW_i = sum_{j != i} wdz_j dens_j / (z_i - z_j) (variant 3),
 * under the real inner product Re sum conj(a) b: with GW_i = sum_j [c_ij dens_j + d_ij conj(dens_j)],
 * mu[t] = sum_{i != j} [conj(c_ij) lam_i + d_ij conj(lam_i)] at j = tgt[t], long double.  wdz =
 * DZDT * H per point.  The reference has no transposes: this is the checker of the build's
 * adjoint products, pinned by the dense transpose of ora_direct / ora_cauchy. */
void ora_direct_transpose(int variant, int n, const double *x, const double *y, const double *wdz,
                          const double *lam, int ntgt, const int *tgt, double *mu);

/* The Hilbert-matrix product of DAPIF2 with IFLAG7 = 3 (dapif1.f:540-558):
 * field[t] = sum_{j != tgt[t]} qa_j / (z_tgt[t] - z_j), long double, at the ntgt targets.
 * qa and field are interleaved complex. */
void ora_cauchy(int n, const double *x, const double *y, const double *qa, int ntgt, const int *tgt, double *field);

/* stokesbubb.f:527-549,585-611: the FASMVP epilogue around a given GW (``stress``).
 * u: 2*nbk interleaved density, w: 2*nbk output; za (complex, k) and pk (k) are outputs. */
void ora_fasmvp_epilogue(int k, const int *nd, const double *x, const double *y, const double *dz,
                         const double *dsdth, const double *rkappa, const double *zk,
                         const double *u, const double *stress, double *w, double *za, double *pk);

/* Unnormalised DFT, sign -1 = DCFFTF (dcfft.f:1-78), +1 = DCFFTB (dcfft.f:715-791). In place. */
void ora_dft(int n, double *c, int sign);
/* dcfft.f:1371-1427 */
void ora_fdifff(int n, const double *f, double *df);
/* electric.f:636-716 (per body geometry). rku interleaved 2*nbk. Outputs are per point except
 * zk (complex per body), area (per body), xleft/slength (scalars). */
void ora_extractdat(int k, const int *nd, const double *rku, double *x, double *y, double *zk,
                    double *dz, double *tha, double *dsda, double *rkappa, double *ckern,
                    double *area, double *xleft, double *slength);
/* electric.f:1270-1302 (+ FOURINT electric.f:1307-1331) */
void ora_equit(int nd, const double *tha, const double *un, double *t);
/* electric.f:1152-1187, in place on nd complex Fourier coefficients */
void ora_filter(int nd, double *cn);
/* electric.f:1076-1103: tangential frame, FFT, filter, inverse FFT -> rkup (2*nbk interleaved) */
void ora_ftail(int k, const int *nd, const double *tha, const double *un, const double *dz,
               const double *dsda, double *rkup);
/* electric.f:773-783: the two Heun stage updates */
void ora_rk2_stage1(int neq, double dt, const double *ybar, const double *dy, double *phi0, double *y);
void ora_rk2_stage2(int neq, double dt, const double *phi0, const double *dy, double *ybar);

/* Adaptive quadtree of DAPIF2: scaling (dapif1.f:740-770), DAASSI..DAIJAT (dapif1.f:1165-1680),
 * DALI12/DALI34 (dapif2.f:3-474).  Built on the host; ora_atree_sizes gives
 * {ierr, nbox, nlev, len, len3, len4, maxp, jatlen, nat, napb}; ora_atree_copy fills caller arrays
 * of those sizes (nl: nlev+2, ibox/xs/ys: nat, per-box arrays: nbox, l1..l4: nbox+1, jat: jatlen,
 * jl: len, jl3: len3, jl4: len4, geom: {rscal, xonew, yonew}).  Any pointer may be NULL. */
typedef struct ora_atree_s ora_atree;
ora_atree *ora_atree_build(int nat, const double *xa, const double *ya, int napb, int lmx, double eps7);
void ora_atree_sizes(const ora_atree *t, int *out);
void ora_atree_copy(const ora_atree *t, double *xs, double *ys, double *geom, int *nl, int *ibox, double *xb,
                    double *yb, int *nbat, int *npar, int *nchi, int *indb, int *iat, int *jat, int *l1, int *l2,
                    int *jl, int *l3, int *l4, int *jl3, int *jl4);
void ora_atree_free(ora_atree *t);

#ifdef __cplusplus
}
#endif
#endif
