/*
 * eb_spectral.h -- C-ABI of the per-body spectral geometry, the Runge-Kutta stage update and
 * the Stokes matvec (FASMVP epilogue + GPU FMM), batched over all bodies (drops) of one or
 * many simulations on one MI355X.
 *
 * Bodies are stored back to back: body k owns points off[k] .. off[k]+nd[k]-1 of every
 * per-point array.  Complex arrays are interleaved (re, im) doubles, as COMPLEX*16 in the
 * reference.  All data pointers are device pointers; `stream` is a hipStream_t (NULL = default).
 * Transforms keep the reference's conventions: DCFFTF is the unnormalised forward DFT
 * C_j = sum_k c_k exp(-2 pi i jk/N), DCFFTB the unnormalised backward one (dcfft.f:1-78,
 * 715-791).  Any body length 2 <= N <= EB_SPECTRAL_NMAX is accepted (mixed radix with a
 * direct stage for large prime factors).
 */
#ifndef EB_SPECTRAL_H
#define EB_SPECTRAL_H

#include <stdint.h>

#include "eb_cadap.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EB_SPECTRAL_NMAX 3200  /* three LDS sequences of N complex doubles per body: 150 KB */

typedef struct eb_bodies_s *eb_bodies;

/* Layout of nbody bodies with nd[k] points each (host array); builds the per-length twiddle
 * tables and factorisations on the device. */
int eb_bodies_create(eb_bodies *bodies, int32_t nbody, const int32_t *nd, void *stream);
int eb_bodies_destroy(eb_bodies bodies);
int eb_bodies_npoints(eb_bodies bodies);

/* DCFFTF (sign = -1, dcfft.f:1-78) / DCFFTB (sign = +1, dcfft.f:715-791) on every body, in place. */
int eb_dcfft(eb_bodies bodies, double *c, int32_t sign, void *stream);
/* FDIFFF (dcfft.f:1371-1427): spectral derivative of every body. f and df may alias. */
int eb_fdifff(eb_bodies bodies, const double *f, double *df, void *stream);
/* FILTER (electric.f:1152-1187) on the Fourier coefficients of every body, in place. */
int eb_filter(eb_bodies bodies, double *cn, void *stream);
/* EQUIT + FOURINT (electric.f:1270-1331): tangential velocity T from tha and Un (real arrays). */
int eb_equit(eb_bodies bodies, const double *tha, const double *un, double *t, void *stream);

/* EXTRACTDAT (electric.f:636-716) for all bodies from the RK state rku (2*nbk interleaved):
 * x, y, dz, tha, dsda, rkappa, ckern per point; zk (complex) and area per body;
 * xleft_slength[2] = {xleft, slength}.  Any output pointer may be NULL. */
int eb_extractdat(eb_bodies bodies, const double *rku, double *x, double *y, double *zk, double *dz,
                  double *tha, double *dsda, double *rkappa, double *ckern, double *area,
                  double *xleft_slength, void *stream);

/* Tail of F (electric.f:1076-1103): EQUIT, the frame velocity (-Un i dz - T dz)/dsda, FFT,
 * FILTER, inverse FFT -> rkup (2*nbk interleaved). */
int eb_ftail(eb_bodies bodies, const double *tha, const double *un, const double *dz,
             const double *dsda, double *rkup, void *stream);

/* RK2 (Heun) stage updates (electric.f:773-783) over neq doubles:
 * stage 1: phi0 = dt*dy, y = ybar + phi0;  stage 2: ybar += 0.5*phi0 + 0.5*dt*dy. */
int eb_rk2_stage1(int32_t neq, double dt, const double *ybar, const double *dy, double *phi0, double *y,
                  void *stream);
int eb_rk2_stage2(int32_t neq, double dt, const double *phi0, const double *dy, double *ybar,
                  void *stream);
/* Classic fourth-order Runge-Kutta, one stage update after each of its four evaluations
 * dy = F(stage state) (stage 0: F(y), 1: F(ys after stage 0), ...): acc accumulates
 * k1 + 2 k2 + 2 k3, ys receives the next stage's state, and stage 3 sets
 * y += dt/6 (k1 + 2 k2 + 2 k3 + k4).  The reference integrates with RK2 only (electric.f:762-786);
 * no counterpart there: RK4 serves the runs that ask for a fourth-order integrator. */
int eb_rk4_stage(int32_t neq, int32_t stage, double dt, double *y, const double *dy, double *acc, double *ys,
                 void *stream);
/* D[k] = (L - B) / (L + B) of every body of the state rku (2*nbk interleaved): L, B the semi-axes
 * of the ellipse with the body's area and second moments of area (trapezoid sums with the
 * spectral tangent).  The reference prints no deformation: the measure of drop deformation
 * used by the single-drop runs and the adjoint's loss. */
int eb_deformation(eb_bodies bodies, const double *rku, double *D, void *stream);

/* ---------------------------------------------------------------------------------------
 * The Stokes matvec of the GMRES solve: MATVEC (stokesbubb.f:419-465) + FASMVP
 * (stokesbubb.f:496-614) with the layer potential from the GPU FMM (eb_fmm_plan).
 * Geometry (x, y, dz = DZ, dsdth = |z_alpha|, rkappa, zk per body) is bound at creation;
 * each apply maps the interleaved density xx (2*nbk) to yy (2*nbk).  h_k = 2 pi / nd_k.
 * mode: EB_MODE_FMM, EB_MODE_DIRECT, or EB_MODE_AUTO (direct sum up to EB_DIRECT_MAX points).
 * --------------------------------------------------------------------------------------- */
typedef struct eb_matvec_s *eb_matvec;

int eb_matvec_create(eb_matvec *mv, eb_bodies bodies, const double *x, const double *y, const double *dz,
                     const double *dsdth, const double *rkappa, const double *zk, int32_t levmax,
                     int32_t nterms, int32_t mode, void *stream);
/* yy = A xx.  When non-NULL, stress (nbk complex) receives GW + self + pressure terms, za / pk
 * (per body) the reference's za and pk (after the pinf shift). */
int eb_matvec_apply(eb_matvec mv, const double *xx, double *yy, double *stress, double *za, double *pk,
                    void *stream);
/* New geometry for the same bodies (the RK stages of a time step): copies the arrays and
 * rebuilds the FMM tree. */
int eb_matvec_rebuild(eb_matvec mv, const double *x, const double *y, const double *dz, const double *dsdth,
                      const double *rkappa, const double *zk, void *stream);
/* The layer potential's guard flag (eb_fmm_set_guard) for later applies; NULL clears it. */
int eb_matvec_set_guard(eb_matvec mv, const int32_t *flag);
int eb_matvec_destroy(eb_matvec mv);

/* ---------------------------------------------------------------------------------------
 * Many independent simulations (a parameter sweep) in one operator: simulation s owns the
 * bodies [sim_off[s], sim_off[s+1]) of `bodies` (sim_off: nsim + 1 host ints, sim_off[0] = 0,
 * sim_off[nsim] = the body count).  eb_sweep_apply is MATVEC + FASMVP (stokesbubb.f:419-614)
 * of every simulation at once: each simulation's layer potential sums over its own points
 * only (the O(N^2) sum that EB_MODE_AUTO also picks for small simulations) and its pressure
 * term uses its own first body's pk (pinf = -pk(1), stokesbubb.f:546).  Geometry as for
 * eb_matvec_create, all simulations back to back.  Replaces the reference's one process per
 * simulation; it has no Fortran counterpart.
 * --------------------------------------------------------------------------------------- */
typedef struct eb_sweep_s *eb_sweep;

int eb_sweep_create(eb_sweep *sw, eb_bodies bodies, int32_t nsim, const int32_t *sim_off, const double *x,
                    const double *y, const double *dz, const double *dsdth, const double *rkappa,
                    const double *zk, void *stream);
int eb_sweep_apply(eb_sweep sw, const double *xx, double *yy, void *stream);
/* za (nbody complex) and pk (nbody, after each simulation's pinf shift) of a density;
 * either output may be NULL. */
int eb_sweep_parts(eb_sweep sw, const double *xx, double *za, double *pk, void *stream);
/* The layer potential alone, GW of CADAPDR (EB_VARIANT_DR) or CADAPMOD (EB_VARIANT_MOD), of every
 * simulation over its own points: dens and gw are nbk complex. */
int eb_sweep_layer(eb_sweep sw, const double *dens, double *gw, int32_t variant, void *stream);
/* New geometry for the same bodies (the RK stages of a time step). */
int eb_sweep_rebuild(eb_sweep sw, const double *x, const double *y, const double *dz, const double *dsdth,
                     const double *rkappa, const double *zk, void *stream);
/* unknowns of each simulation: offsets into the 2*nbk vector (nsim + 1 ints, host memory) */
const int32_t *eb_sweep_unknown_offsets(eb_sweep sw);
/* Per-simulation flags (device, nsim ints, may be NULL): while set, eb_sweep_apply skips the
 * layer potential of simulations whose flag is 0 (their output is left unspecified).  The
 * batched solves set it to the solver's active flags around their operator calls. */
int eb_sweep_set_mask(eb_sweep sw, const int32_t *mask);
/* How the sweep's layer potentials run: EB_MODE_DIRECT (each simulation's O(N^2) sum, the tiles
 * above), EB_MODE_FMM (segmented FMM plans, one segment per simulation, eb_fmm_plan_create_segmented:
 * each simulation binned with its own bounding box and scale, 43 terms), EB_MODE_AUTO (default:
 * FMM when a simulation has more than EB_DIRECT_MAX points).  In FMM mode every simulation's GW is
 * formed on each apply: the mask above is not consulted.  No counterpart in the reference, which
 * runs one simulation per process. */
int eb_sweep_set_mode(eb_sweep sw, int32_t mode);
int eb_sweep_destroy(eb_sweep sw);

/* Drop-in for FASMVP (stokesbubb.f:496-498) on host arrays, Fortran convention (all by
 * reference): (ND, NBK, K, X, Y, DZ, DSDTH, RKAPPA, ZK, U, W, H, ZA, PK, STRESS, DENS).
 * The FMM runs with the reference's LEVMAX 7 and TOL 1e-13 (stokesbubb.f:564-567). */
int32_t eb_fasmvp(const int32_t *nd, const int32_t *nbk, const int32_t *k, const double *x, const double *y,
                  const double *dz, const double *dsdth, const double *rkappa, const double *zk,
                  const double *u, double *w, const double *h, double *za, double *pk, double *stress,
                  double *dens);

#ifdef __cplusplus
}
#endif
#endif
