/*
 * eb_field.h -- C-ABI of F, the right-hand side of the drop evolution, on the device.
 *
 * F (electric.f:968-1109) for one simulation of k bodies, with no gravity (zbond = 0) and no
 * wall (iwall = 0), the settings of the reference's drivers for these runs:
 *
 *   EXTRACTDAT (electric.f:636-716)
 *   GETBCs + SOLVE_LAPL (laplace.f:5-25, 335-425): DGMRES on the Laplace operator FASMVP_LAP
 *     (laplace.f:159-262), whose Hilbert product (DAPIF2 with IFLAG7 = 3) runs on the GPU FMM
 *   DIRICH_NEU_MAP (laplace.f:266-332) with PVINTEV (pv_integral.f:53-212): spectral
 *     interpolation to twice the points, alternating-point Hilbert product, FDIFFF
 *   SETVELF (stokesbubb.f:4-135), SOLVEI (stokesbubb.f:294-415),
 *   INTERFACENEW (electric.f:791-912) with CADAPMOD on the GPU FMM
 *   the F tail: EQUIT, frame velocity, FFT, FILTER, inverse FFT (electric.f:1068-1103)
 *
 * All arrays are device pointers; rku / rkup are interleaved (x, y) per point, bodies back to
 * back.  cap and e0 are COMMON /flowinfo/'s CAP and E0.  Every call is stream-ordered except
 * the GMRES solves inside, which synchronise their stream once per iteration.
 */
#ifndef EB_FIELD_H
#define EB_FIELD_H

#include <stdint.h>

#include "eb_spectral.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eb_field_s *eb_field;

/* Bodies up to 2048 points each (PVINTEV works on twice the points). */
int eb_field_create(eb_field *f, eb_bodies bodies, double cap, double e0, void *stream);
int eb_field_destroy(eb_field f);
/* Evaluation mode of every layer potential of F (the Laplace and PVINTEV Hilbert products, the
 * Stokes matvec, CADAPMOD): EB_MODE_AUTO (default: the O(N^2) sum up to EB_DIRECT_MAX points,
 * the FMM above), EB_MODE_FMM or EB_MODE_DIRECT.  Drops the current plans; the next
 * eb_field_set_geometry (or eb_field_rhs) makes them in the new mode.  No counterpart in the
 * reference (its products are always FMMs); a knob for parity runs of either path. */
int eb_field_set_mode(eb_field f, int32_t mode);

/* rkup = F(rku): the whole chain above, on the field's stream. */
int eb_field_rhs(eb_field f, const double *rku, double *rkup);

/* One Heun RK2 step (electric.f:773-783), ybar += (dt F(ybar) + dt F(y)) / 2, captured once into
 * a hipGraph and replayed on later calls: the Laplace and Stokes solves of both F run the
 * device-controlled DGMRES of one system (eb_gmres_batch_solve_async, nsys = 1) with at most
 * `budget` Arnoldi steps each, so the step has no host synchronisation; in direct mode the
 * pair sums of the products after convergence are skipped (eb_fmm_set_guard).  After the
 * replay the solves' failure counters are read; if a solve needed more than `budget` steps,
 * ybar is restored and the step is run again with the host-controlled solves (*redone = 1).
 * ybar: 2 nbk doubles on the device, updated in place.  The graph is rebuilt when ybar, dt or
 * budget change, or after eb_field_set_mode.  The field must have been created on a
 * non-default stream.  No counterpart in the reference (its solver stops on the host). */
int eb_field_rk2_graph(eb_field f, double *ybar, double dt, int32_t budget, int32_t *redone);
/* The same for one classic RK4 step (four F, eb_rk4_stage after each), captured and replayed
 * the same way; the graph is rebuilt when the call switches between the two integrators.  The
 * reference integrates with RK2 only: no counterpart there. */
int eb_field_rk4_graph(eb_field f, double *ybar, double dt, int32_t budget, int32_t *redone);

/* The stages one at a time, for tests and for callers that replace part of the chain.
 * eb_field_set_geometry runs EXTRACTDAT on rku and rebuilds every FMM tree; the others act on
 * that geometry. */
int eb_field_set_geometry(eb_field f, const double *rku);
/* w = FASMVP_LAP(u) for the nbk unknowns of the Laplace system (the reference also writes k
 * constraint rows past the system's end, laplace.f:251-259; they are not part of it). */
int eb_field_laplace_apply(eb_field f, const double *u, double *w);
/* GETBCs + SOLVE_LAPL: ulap (nbk); *iter, *ierr as DGMRES reports them (may be NULL). */
int eb_field_laplace_solve(eb_field f, double *ulap, int32_t *iter, int32_t *ierr);
/* DIRICH_NEU_MAP: phi, psi, dphi_dn (nbk each; phi and psi may be NULL). */
int eb_field_dirich_neu(eb_field f, const double *ulap, double *phi, double *psi, double *dphi_dn);
/* SETVELF: rhs (2 nbk) from dphi_dn. */
int eb_field_setvelf(eb_field f, const double *dphi_dn, double *rhs);
/* SOLVEI: u (nbk complex = 2 nbk doubles), za (k complex), pk (k); *iter, *ierr may be NULL. */
int eb_field_stokes_solve(eb_field f, const double *rhs, double *u, double *za, double *pk, int32_t *iter,
                          int32_t *ierr);
/* INTERFACENEW: the normal velocity un (nbk) from the density u. */
int eb_field_interface(eb_field f, const double *u, double *un);

/* GMRES iteration counts and IERR of the last Laplace and Stokes solves (any may be NULL). */
int eb_field_iterations(eb_field f, int32_t *lap_iter, int32_t *lap_ierr, int32_t *stokes_iter, int32_t *stokes_ierr);

/* Geometry of the current call (device pointers owned by the field). */
typedef struct {
  const double *x, *y, *dz, *tha, *dsda, *rkappa, *ckern, *zk;
} eb_field_geometry;
int eb_field_get_geometry(eb_field f, eb_field_geometry *g);

/* ---------------------------------------------------------------------------------------
 * F of many independent simulations at once (a parameter sweep).  Simulation s owns bodies
 * [sim_off[s], sim_off[s+1]) of `bodies` (nsim + 1 host ints) and has its own CAP and E0
 * (cap, e0: nsim host doubles).  Every stage is one launch for all simulations; the Laplace
 * and Stokes solves run in the batched GMRES, one system per simulation; every layer
 * potential sums over the simulation's own points.  The reference runs one simulation per
 * process: there is no counterpart in it.
 * --------------------------------------------------------------------------------------- */
typedef struct eb_sweep_field_s *eb_sweep_field;

int eb_sweep_field_create(eb_sweep_field *f, eb_bodies bodies, int32_t nsim, const int32_t *sim_off,
                          const double *cap, const double *e0, void *stream);
/* rkup = F(rku) of every simulation (all bodies back to back, interleaved (x, y)). */
int eb_sweep_field_rhs(eb_sweep_field f, const double *rku, double *rkup);
/* Evaluation mode of every layer potential of the sweep's F: EB_MODE_DIRECT (each simulation's
 * O(N^2) sums), EB_MODE_FMM (segmented FMM plans, one segment per simulation: the Laplace and
 * PVINTEV Hilbert products at 37 and 33 terms as their TOL 1e-11 and 1e-10, laplace.f:212,
 * pv_integral.f:175; the Stokes operator and CADAPMOD at 43, eb_sweep_set_mode), EB_MODE_AUTO
 * (default: FMM when a simulation has more than EB_DIRECT_MAX points).  Drops a captured step. */
int eb_sweep_field_set_mode(eb_sweep_field f, int32_t mode);
/* GMRES iteration counts and IERR of the last call's Laplace and Stokes solves (nsim each;
 * any may be NULL). */
int eb_sweep_field_iterations(eb_sweep_field f, int32_t *lap_iter, int32_t *lap_ierr, int32_t *stokes_iter,
                              int32_t *stokes_ierr);
int eb_sweep_field_destroy(eb_sweep_field f);

/* One Heun RK2 step (electric.f:773-783) of every simulation, ybar += (dt F(ybar) + dt F(y)) / 2,
 * captured once into a hipGraph and replayed on later calls: the two F of the step with their
 * four batched solves run the device-controlled DGMRES (eb_gmres_batch_solve_async) with at
 * most `budget` Arnoldi steps each, so the step has no host synchronisation.  After the replay
 * the solves' failure counters are read; if any system needed more than `budget` steps, ybar
 * is restored and the step is run again with the host-controlled solves (*redone = 1).
 * ybar: 2 nbk doubles on the device, updated in place.  The graph is rebuilt when ybar, dt or
 * budget change.  The field must have been created on a non-default stream. */
int eb_sweep_field_rk2_graph(eb_sweep_field f, double *ybar, double dt, int32_t budget, int32_t *redone);
/* One classic RK4 step of every simulation (four F with their eight batched solves), captured
 * and replayed as eb_sweep_field_rk2_graph.  No counterpart in the reference (RK2 only). */
int eb_sweep_field_rk4_graph(eb_sweep_field f, double *ybar, double dt, int32_t budget, int32_t *redone);

#ifdef __cplusplus
}
#endif
#endif
