/*
 * eb_sim.h -- C-ABI of the whole time evolution on the device: the driver MANYDROP
 * (electric.f:1-272) around F (include/eb_field.h), for no gravity and no wall.
 *
 *   READINI (electric.f:275-337): the `bubbin` input (eb_sim_read_bubbin)
 *   GEOINIT + INITPARAM (electric.f:342-500): ellipses (NCYC = 0), starfish (NCYC > 0) and
 *     the NCYC < 0 family; CLUSTPARAM (electric.f:1192-1265): redistribution to equal
 *     arclength by Newton's method on the Fourier series of the arclength; GEOSHIFT
 *   the time loop (electric.f:144-255): RK2 steps, EXTRACTSOL (electric.f:575-631),
 *     GETFOURIER's doubling test and DOUBLE (electric.f:531-570, 721-757), the step law
 *     dt = min_k cmin arcl_k / nd_k with cmin = max_k dt0 nd_k / arcl0_k
 *   DUMPDATA / READDATA (electric.f:1360-1423): `geometry.data`
 *
 * Device memory for the state; the parameters and the per-body scalars in host memory.
 */
#ifndef EB_SIM_H
#define EB_SIM_H

#include <stdint.h>

#include "eb_field.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EB_SIM_KMAX 150 /* bodies (the reference's KMAX, electric.f:10) */
#define EB_SIM_NDMAX 2048 /* points per body (F's PVINTEV interpolation; the reference's NDMAX is 3000) */

/* The run description that READINI reads from `bubbin`. */
typedef struct {
  int32_t k;
  int32_t nd[EB_SIM_KMAX];
  double dt, tfinal;
  int32_t nplot;
  double cap, bondx, bondy, e0;
  double ai[EB_SIM_KMAX], bi[EB_SIM_KMAX], thetax[EB_SIM_KMAX], cx[EB_SIM_KMAX], cy[EB_SIM_KMAX];
  int32_t ncyc[EB_SIM_KMAX];
} eb_sim_params;

/* READINI: the first configuration of a `bubbin` file (list-directed: K; ND(1..K); DT,
 * TFINAL, NPLOT; CAP, BONDX, BONDY, E0; then AI, BI, THETAX, CX, CY, NCYC per body).
 * Returns EB_OK or EB_ERR_ARG (unreadable or inconsistent file). */
int eb_sim_read_bubbin(const char *path, eb_sim_params *p);

typedef struct eb_sim_s *eb_sim;

/* GEOINIT + CLUSTPARAM + GEOSHIFT + INITCOND on the device.  Gravity (BONDX, BONDY not 0,
 * which the reference runs with a wall) is rejected. */
int eb_sim_create(eb_sim *s, const eb_sim_params *p, void *stream);
/* One pass of the time loop: the RK2 step, EXTRACTSOL and GETFOURIER of every body, DOUBLE
 * where the spectrum asks for it, the new dt.  *done = 1 once time reaches TFINAL.  A body
 * whose spectrum asks for doubling past EB_SIM_NDMAX points ends the run with
 * EB_ERR_UNSUPPORTED after the RK2 step (the reference STOPs there, electric.f:190-199). */
int eb_sim_step(eb_sim s, int32_t *done);

typedef struct {
  double time, dt, area_total, area_total0, area_error;
  int32_t k, nbk, steps, doublings;
  const int32_t *nd;     /* host, k entries */
  const double *area;    /* host, k entries (EXTRACTSOL of the last step) */
  const double *arcl;    /* host, k entries */
  const double *arcl0;   /* host, k entries (CLUSTPARAM) */
  const double *ybar;    /* device, 2 nbk: the interleaved (x, y) state */
} eb_sim_state;
int eb_sim_get_state(eb_sim s, eb_sim_state *st);
/* Copies the state (2 nbk doubles) to dst, host or device memory; synchronous. */
int eb_sim_copy_state(eb_sim s, double *dst);

/* DUMPDATA: `geometry.data` in the reference's format. */
int eb_sim_dump(eb_sim s, const char *path);
/* READDATA (electric.f:85-104 with IDUMP = 1): the state of a `geometry.data` (K, ND, DT,
 * TSTART, AREA0, ARCL0, the points); p supplies CAP, E0 and TFINAL, which runs to
 * TSTART + TFINAL as in the reference. */
int eb_sim_restore(eb_sim *s, const eb_sim_params *p, const char *path, void *stream);
int eb_sim_destroy(eb_sim s);

#ifdef __cplusplus
}
#endif
#endif
