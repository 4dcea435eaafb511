/*
 * eb_gmres.h -- C-ABI of the device-resident GMRES solve of the Stokes integral equation.
 *
 * DGMRES (dgmres.f:2-553, with DPIGMR dpigmr.f:295-437, DORTH dorth.f:90-121, DHEQR
 * dheqr.f:82-176, DHELS dhels.f, DRLCAL drlcal.f and the ITOL = 0 test of ISDGMR
 * isdgmr.f:289-394) for the settings its caller uses: ITOL = 0, no scaling (JSCAL = 0), no
 * preconditioner (JPRE = 0) and full modified Gram-Schmidt (KMP = MAXL).  The Krylov basis,
 * the iterate and every n-vector stay in device memory; only the O(MAXL) Hessenberg column,
 * the norms and the Givens rotations are handled on the host, in the reference's operation
 * order.  The operator is any device function y = A x.
 *
 * All vector pointers are device pointers; `stream` is a hipStream_t (NULL = default).
 */
#ifndef EB_GMRES_H
#define EB_GMRES_H

#include <stdint.h>

#include "eb_spectral.h"

#ifdef __cplusplus
extern "C" {
#endif

/* y = A x on the device, stream-ordered; returns EB_OK or an EB_ERR_* code. */
typedef int (*eb_apply_fn)(void *ctx, const double *x, double *y, void *stream);

typedef struct eb_gmres_s *eb_gmres;

/* Workspace for n unknowns and a Krylov space of at most maxl vectors (IGWK(1) = MAXL:
 * 0 means 10, and values above n are cut to n, dgmres.f). */
int eb_gmres_create(eb_gmres *g, int32_t n, int32_t maxl, void *stream);

/* Solves A x = b.  x holds the initial guess on entry and the solution on return.
 * tol: ITOL = 0 tolerance on |b - A x| / |b| (0 means 500 machine epsilons); nrmax: IGWK(5)
 * (0 means 10 restarts, -1 none).  Outputs: *iter (ITER), *err (ERR) and *ierr (IERR: 0
 * converged, 1 not converged after the restarts, 2 stagnation).  Returns EB_OK or an
 * EB_ERR_* code (a failing operator call). */
int eb_gmres_solve(eb_gmres g, eb_apply_fn apply, void *ctx, const double *b, double *x, double tol,
                   int32_t nrmax, int32_t *iter, double *err, int32_t *ierr);
int eb_gmres_destroy(eb_gmres g);

/* eb_apply_fn for a dense n x n matrix in device memory, column-major (Fortran order), as
 * the operator of small systems and of the solver's own tests (ctx = an eb_dense_op). */
typedef struct {
  int32_t n;
  const double *a;
} eb_dense_op;
int eb_dense_gmres_apply(void *op, const double *x, double *y, void *stream);

/* eb_apply_fn for the Stokes matvec (ctx = an eb_matvec): MATVEC, stokesbubb.f:419-465. */
int eb_matvec_gmres_apply(void *mv, const double *x, double *y, void *stream);

/* za and pk (after the pinf shift) of a density, without the layer potential: the
 * post-processing of SOLVEI (stokesbubb.f:379-409).  Either output may be NULL. */
int eb_matvec_parts(eb_matvec mv, const double *xx, double *za, double *pk, void *stream);

/* SOLVEI (stokesbubb.f:294-415) on the device: soln = rhs as the initial guess, DGMRES with
 * the workspace's MAXL (SOLVEI's caller passes 100, electric.f:1011,1060), TOL = 1e-10 and no
 * restart, then za and pk of the solution density.  The workspace must use `stream`.
 * rhs, soln: 2*nbk doubles; za: nbody complex; pk: nbody doubles (may be NULL). */
int eb_stokes_solve(eb_matvec mv, eb_gmres g, const double *rhs, double *soln, double *za, double *pk,
                    int32_t *iter, double *err, int32_t *ierr, void *stream);

/* ---------------------------------------------------------------------------------------
 * Batched GMRES: nsys independent systems in lockstep, each with DGMRES's arithmetic for
 * IGWK(5) = -1 (one cycle, no restart: SOLVEI's setting, stokesbubb.f:294-415).  System s
 * owns entries [off[s], off[s+1]) of every vector (off: nsys + 1 host ints, off[0] = 0); the
 * operator applies all systems at once.  MAXL is cut to each system's size.  No
 * counterpart in the reference, which solves one simulation per process.
 * --------------------------------------------------------------------------------------- */
typedef struct eb_gmres_batch_s *eb_gmres_batch;

int eb_gmres_batch_create(eb_gmres_batch *g, int32_t nsys, const int32_t *off, int32_t maxl, void *stream);
/* x: initial guesses in, solutions out.  iter, err, ierr: nsys entries each (may be NULL);
 * ierr as eb_gmres_solve (1: not converged within MAXL). */
int eb_gmres_batch_solve(eb_gmres_batch g, eb_apply_fn apply, void *ctx, const double *b, double *x, double tol,
                         int32_t *iter, double *err, int32_t *ierr);
int eb_gmres_batch_destroy(eb_gmres_batch g);

/* The same solve with the per-system scalar work (DORTH's test, DHEQR, ISDGMR's test, DHELS)
 * in device kernels, one thread per system, in the same order and arithmetic: nothing is
 * synchronised or copied to the host, so the launches can be captured in a hipGraph.  It
 * always runs min(budget, MAXL) Arnoldi steps (systems that are done drop out of the
 * kernels); a system still iterating after `budget` steps stops there with the update of
 * those steps, IERR = 3, and is counted in the workspace's failure counter.
 * eb_gmres_batch_results fetches ITER, ERR, IERR of the last such solve and the counter
 * (synchronises); eb_gmres_batch_clear_fail zeroes the counter with a kernel. */
int eb_gmres_batch_solve_async(eb_gmres_batch g, eb_apply_fn apply, void *ctx, const double *b, double *x,
                               double tol, int32_t budget);
int eb_gmres_batch_clear_fail(eb_gmres_batch g);
/* The workspace's per-system active flags (device, nsys ints): 1 while a system iterates (all
 * 1 for the initial residual).  An operator may skip the systems whose flag is 0. */
const int32_t *eb_gmres_batch_active(eb_gmres_batch g);
int eb_gmres_batch_results(eb_gmres_batch g, int32_t *iter, double *err, int32_t *ierr, int32_t *nfail);

/* eb_apply_fn for the sweep operator (ctx = an eb_sweep) */
int eb_sweep_gmres_apply(void *sw, const double *x, double *y, void *stream);
/* eb_sweep_solve's solve through eb_gmres_batch_solve_async (no za / pk). */
int eb_sweep_solve_async(eb_sweep sw, eb_gmres_batch g, const double *rhs, double *soln, int32_t budget, void *stream);

/* SOLVEI of every simulation of the sweep at once: soln = rhs as the initial guesses,
 * TOL = 1e-10, the workspace's MAXL, then za and pk of each solution.  The workspace must
 * have been created with the sweep's unknown offsets (eb_sweep_unknown_offsets) and `stream`.
 * iter, err, ierr: one entry per simulation. */
int eb_sweep_solve(eb_sweep sw, eb_gmres_batch g, const double *rhs, double *soln, double *za, double *pk,
                   int32_t *iter, double *err, int32_t *ierr, void *stream);

#ifdef __cplusplus
}
#endif
#endif
