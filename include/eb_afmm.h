/*
 * eb_afmm.h -- C-ABI of the FMM on the adaptive tree of DAPIF2.
 *
 * The layer-potential sums of eb_fmm_apply (CADAPDR / CADAPMOD / CADAP, or the Hilbert
 * product of DAPIF2 with IFLAG7 = 3) evaluated on the adaptive quadtree that DAPIF2 builds
 * (dapif1.f:1165-1680, eb_atree): a box is split while it holds more than NAPB points
 * (NAPB = 40 at 2^20 points), so the leaves follow the points instead of a uniform grid of
 * LEVMAX levels.  The reference runs its uniform-grid CADAPDR on this path; the adaptive form
 * is this build's throughput mode at large N and has no counterpart there.  Results agree
 * with the O(N^2) sum to the expansion length's accuracy (about 2^-NTERMS relative).
 *
 * Data pointers are device pointers; `stream` is a hipStream_t (NULL = default).  The plan
 * holds the tree, the interaction lists and the expansion storage; its creation synchronises
 * the stream (the tree build sizes each level from the last, and the lists are formed on the
 * host).
 */
#ifndef EB_AFMM_H
#define EB_AFMM_H

#include <stdint.h>

#include "eb_cadap.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eb_afmm_s *eb_afmm;

/* n points (x, y), weights wdz (n complex: dzdt * h per point, as eb_fmm_plan_create), at most
 * napb points per childless box above the deepest level, nterms terms, variant EB_VARIANT_*. */
int eb_afmm_create(eb_afmm *plan, int32_t n, const double *x, const double *y, const double *wdz, int32_t napb,
                   int32_t nterms, int32_t variant, void *stream);
/* gradw (n complex) = the sum over j != i for the density dens (n complex), as eb_fmm_apply. */
int eb_afmm_apply(eb_afmm plan, const double *dens, double *gradw, void *stream);
/* reals3 = {RSCAL, XONEW, YONEW}: the plan computes in the tree's coordinates
 * (x - XONEW) * RSCAL and returns RSCAL * (the sums there) (may be NULL); counts8 = {boxes,
 * childless boxes, levels, ordered near pairs, M2L translations, M2P evaluations (point x
 * box), P2L terms (point x box), boxes of level >= 3 (the M2M and the L2L translations
 * each)}. */
int eb_afmm_info(eb_afmm plan, double *reals3, int64_t *counts8);
/* The list-V (M2L) pairs by source group, as eb_fmm_plan_m2l_split: out3[0] near (all p
 * input terms), out3[1] far, out3[2] the far group's input terms. */
int eb_afmm_m2l_split(eb_afmm plan, int64_t *out3);
/* Milliseconds of the last apply's stages [charges, P2P, P2M, M2M, M2L, P2L, L2L, L2P + M2P]
 * from HIP events, while timing is enabled (the apply then synchronises). */
int eb_afmm_enable_timing(eb_afmm plan, int32_t on);
int eb_afmm_stage_ms(eb_afmm plan, double *ms8);
/* Host check of the lists: *bad = the ordered pairs of childless boxes whose interaction the
 * passes do not carry exactly once (0 for a correct plan).  O(leaves^2 x depth^2). */
int eb_afmm_check(eb_afmm plan, int64_t *bad);
int eb_afmm_destroy(eb_afmm plan);

#ifdef __cplusplus
}
#endif
#endif
