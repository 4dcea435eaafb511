/*
 * eb_atree.h -- C-ABI of the adaptive quadtree of DAPIF2 built on the GPU.
 *
 * The tree that the reference's adaptive FMM builds before its passes (DAPIF2,
 * dapif1.f:458-459): scaling of the points into [-64, 64]^2 (dapif1.f:745-770, with the
 * REAL*4 factor 0.99), DAASSI / DAINIA / DAIBNB / DAFAMI (dapif1.f:1165-1453: a box with
 * more than NAPB points gets four children NW, NE, SW, SE, numbered level by level in the
 * order of their parents), DACOMP (dapif1.f:1456-1539: empty boxes removed), DAINDB
 * (dapif1.f:1542-1606), DAIJAT (dapif1.f:1609-1680), DALI12 (dapif2.f:3-210: interaction
 * list #1 and colleagues #2) and DALI34 (dapif2.f:212-474: lists #3 and #4 of childless
 * boxes).  Every table equals the reference's bit for bit, with its 1-based values.
 *
 * All data pointers are device pointers; `stream` is a hipStream_t (NULL = default).  The
 * build synchronises the stream once per tree level and once per list depth (it sizes the
 * next level from the counts of the last).
 */
#ifndef EB_ATREE_H
#define EB_ATREE_H

#include <stdint.h>

#include "eb_cadap.h"

#ifdef __cplusplus
extern "C" {
#endif

/* IERR(1) values of DAPIF2 that the tree build can produce (dapif1.f:711-717, 1222-1236). */
#define EB_ATREE_EPS_STOP   1   /* smallest box below EPS7: the reference continues with a
                                   partly formed level; this build stops and reports it */
#define EB_ATREE_FEW_POINTS 16  /* NAT <= NAPB */
#define EB_ATREE_TOO_DEEP   32  /* more than LMX levels needed */
#define EB_ATREE_WORKSPACE  8   /* more boxes than NBOXMX before compression (DAFAMI,
                                   dapif1.f:1436-1441; eb_atree_build_bounded only) */

typedef struct eb_atree_s *eb_atree;

/* Device views of the tables (valid until eb_atree_destroy).  Sizes from eb_atree_counts:
 *   xs, ys, ibox: nat;  geom: {RSCAL, XONEW, YONEW};  nl: nlev + 2 (NL(1..L+1));
 *   xb, yb, nbat, npar, nchi, indb, iat: nbox;  l1, l2, l3, l4: nbox + 1;
 *   jat: jatlen;  jl: len;  jl3: len3;  jl4: len4. */
typedef struct {
  const double *xs, *ys, *geom, *xb, *yb;
  const int32_t *nl, *ibox, *nbat, *npar, *nchi, *indb, *iat, *jat, *l1, *l2, *jl, *l3, *l4, *jl3, *jl4;
} eb_atree_view;

/* Builds the tree of nat points (x, y device arrays) with at most napb points per childless
 * box above the finest level and at most lmx levels (DAPIF2's LMX; 30 for MEX = 300,
 * TOL = 1e-11).  Returns EB_OK, one of the EB_ATREE_* codes above (ierr[0..2] then hold
 * IERR(1..3) as DAPIF2 sets them and *t stays NULL), or an EB_ERR_* code. */
int eb_atree_build(eb_atree *t, int32_t nat, const double *x, const double *y, int32_t napb, int32_t lmx,
                   double eps7, int32_t *ierr, void *stream);
/* The same with DAPIF2's bound on the boxes before compression (NBOXMX, dapif1.f:784): the
 * build stops with EB_ATREE_WORKSPACE, IERR(1) = 8 and IERR(2) as DAFAMI sets them, where the
 * reference's tree would outgrow its part of the workspace.  nboxmx < 0: no bound. */
int eb_atree_build_bounded(eb_atree *t, int32_t nat, const double *x, const double *y, int32_t napb, int32_t lmx,
                           double eps7, int64_t nboxmx, int32_t *ierr, void *stream);
/* counts[8] = {nbox, nlev, len, len3, len4, maxp, jatlen, nat} (host array). */
int eb_atree_counts(eb_atree t, int32_t *counts);
int eb_atree_tables(eb_atree t, eb_atree_view *view);
int eb_atree_destroy(eb_atree t);

#ifdef __cplusplus
}
#endif
#endif
