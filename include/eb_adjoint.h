/*
 * eb_adjoint.h -- C-ABI of the pieces a hand-written adjoint of F needs: the layer potentials
 * of F as O(N^2) pair sums with their vector-Jacobian products, and the Laplace and Stokes
 * operators of F with their transposes, each solvable by the device GMRES.
 *
 * The reference has no derivatives: there is no counterpart in it.  The operators are the ones
 * F (electric.f:968-1109) solves, for one simulation of k bodies with no gravity and no wall:
 *
 *   Laplace, FASMVP_LAP (laplace.f:159-262) on the nbk unknowns:
 *     (A u)_i = u_i (1 + h kappa_i s_i / (2 pi) - 2 h s_i) - Im(CF_i) / pi,
 *     CF_i = sum_{j != i} u_j h dz_j / (z_i - z_j)
 *   Stokes, FASMVP (stokesbubb.f:496-614) on the complex density sigma (2 nbk doubles):
 *     (A sigma)_i = sigma_i + za_b + GW_i + c0_i (sigma_i - conj(sigma_i) zn_i^2) + (i/2) pk_b z_i,
 *     GW = CADAPDR's kernel (cadapdrop.f:218-222, cadapdrop2.f:564-620) summed directly,
 *     za_b = h sum_b sigma, pk_b = (2h/pi) sum_b Re(sigma dz / (z - zk_b)^2),
 *     c0 = h s kappa / (2 pi), zn = -i dz / s
 *
 * Complex arrays are interleaved (re, im) doubles; every pointer is a device pointer.  The
 * adjoint convention: for a real loss L and complex v, vbar = dL/dRe v + i dL/dIm v.
 *
 * Pair sums:
 *   eb_adj_cauchy:      g_i = sum_j c_j / (t_i - s_j)      (self = 1: t = s, j = i left out)
 *   eb_adj_stokes_pairs: gw_i = sum_{j != i} [qa_j / r + s2 (conj(qb_j) / conj(r)
 *                                                    - conj(qa_j) r / conj(r)^2)],  r = z_j - z_i
 *     (s2 = -1: CADAPDR's signs; s2 = +1: CADAPMOD's, cadap2.f:518,617)
 * and their vector-Jacobian products given gbar / gwbar.
 */
#ifndef EB_ADJOINT_H
#define EB_ADJOINT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eb_adj_s *eb_adj;

/* nbody bodies of nd[b] points each (host array), back to back; GMRES with MAXL maxl and no
 * restart (SOLVEI's and SOLVE_LAPL's setting, stokesbubb.f:294-415, laplace.f:335-425). */
int eb_adj_create(eb_adj *a, int32_t nbody, const int32_t *nd, int32_t maxl, void *stream);
int eb_adj_destroy(eb_adj a);

/* The geometry the operators act on: z, dz (n complex), dsda, kappa (n doubles), zk (nbody
 * complex) as EXTRACTDAT gives them (electric.f:636-716).  Copied. */
int eb_adj_set_geometry(eb_adj a, const double *z, const double *dz, const double *dsda, const double *kappa,
                        const double *zk);

/* How the operators below evaluate their layer potentials: EB_MODE_DIRECT (the O(N^2) pair
 * sums), EB_MODE_FMM (FMM plans over the same points: CADAPDR's for GW, the Hilbert product for
 * CF; the transposed operators through eb_fmm_apply_transpose, the same sums with sources and
 * targets swapped), EB_MODE_AUTO (default: direct up to EB_DIRECT_MAX points, FMM above).  The
 * plans have LEVMAX with 4^LEVMAX >= n / 48 (3..10) and 43 terms (TOL 1e-13).  The pair sums
 * and their reverses further down stay O(N^2) in every mode. */
int eb_adj_set_mode(eb_adj a, int32_t mode);

/* y = A x or y = A^T x (transpose = 1) for the Laplace (n doubles) and Stokes (2n doubles)
 * operators above. */
int eb_adj_laplace_apply(eb_adj a, int32_t transpose, const double *x, double *y);
int eb_adj_stokes_apply(eb_adj a, int32_t transpose, const double *x, double *y);
/* Solves A x = b or A^T x = b by DGMRES (ITOL 0, relative residual tol); x holds the initial
 * guess on entry.  iter, err, ierr as eb_gmres_solve (may be NULL). */
int eb_adj_laplace_solve(eb_adj a, int32_t transpose, const double *b, double *x, double tol, int32_t *iter,
                         double *err, int32_t *ierr);
int eb_adj_stokes_solve(eb_adj a, int32_t transpose, const double *b, double *x, double tol, int32_t *iter,
                        double *err, int32_t *ierr);

/* Pair sums on any point sets (they use the handle's stream and scratch only). */
int eb_adj_cauchy(eb_adj a, int32_t nt, const double *t, int32_t ns, const double *s, const double *c, int32_t self,
                  double *g);
/* tbar (nt), sbar (ns), cbar (ns) complex; any may be NULL. */
int eb_adj_cauchy_vjp(eb_adj a, int32_t nt, const double *t, int32_t ns, const double *s, const double *c,
                      int32_t self, const double *gbar, double *tbar, double *sbar, double *cbar);
int eb_adj_stokes_pairs(eb_adj a, int32_t n, const double *z, const double *qa, const double *qb, double s2,
                        double *gw);
/* zbar, qabar, qbbar (n complex each); zbar may be NULL. */
int eb_adj_stokes_pairs_vjp(eb_adj a, int32_t n, const double *z, const double *qa, const double *qb, double s2,
                            const double *gwbar, double *zbar, double *qabar, double *qbbar);

#ifdef __cplusplus
}
#endif
#endif
