/*
 * eb_cadap.h -- C-ABI of the MI355X (gfx950) Stokes-matvec path of ElectricBubbles.
 *
 * Two layers:
 *
 * 1. Drop-in entry points with the exact 26-argument Fortran interface of the reference's
 *    complex-variable FMM (all arguments by reference, COMPLEX*16 = interleaved double pairs,
 *    ND(K0:*) and H(K0:*) point at element K0).  A thin ISO_C_BINDING shim
 *    (electricbubbles_amd/fortran/eb_cadap_shim.f90) exports the reference names
 *    CADAPDR / CADAPMOD / CADAP and forwards here, so the reference drivers link unchanged.
 *    Host buffers in, host buffers out; the call blocks until GRADW is back in host memory.
 *
 * 2. A device-resident, stream-ordered plan/apply API used by the drop-in layer itself, by
 *    the benchmark and by batched callers: geometry is binned once per plan (the tree is a
 *    pure function of the point coordinates), and each apply evaluates the layer potential
 *    for a new density.
 *
 * Error behaviour mirrors the reference: the int return value equals IER(1) (0 ok, 4 too many
 * terms, 8 workspace too small); build-specific failures use EB_ERR_* (>= 16) and leave a
 * message for eb_last_error().
 */
#ifndef EB_CADAP_H
#define EB_CADAP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (0, 4, 8 are the reference's IER(1) values, cadapdrop.f:89-103) */
#define EB_OK             0
#define EB_ERR_TERMS      4    /* NTERMS > 60 (cadapdrop.f:129-133) */
#define EB_ERR_WORKSPACE  8    /* LENW / LENIW too small (cadapdrop.f:191-197, 283-290, 312-319) */
#define EB_ERR_UNSUPPORTED 16  /* IFLAG1=1 (targets) or IFLAG2=1 (wall): not built */
#define EB_ERR_ARG        17   /* invalid argument (n < 1, LEVMAX out of range, ...) */
#define EB_ERR_HIP        18   /* HIP runtime error */
#define EB_ERR_NOMEM      19

/* Kernel variants: sign conventions of the three reference drivers. */
#define EB_VARIANT_DR   0  /* CADAPDR  (cadapdrop.f): GW = phi - z conj(phi') - conj(psi) */
#define EB_VARIANT_MOD  1  /* CADAPMOD (cadapmod.f):  GW = phi + z conj(phi') + conj(psi) */
#define EB_VARIANT_PLUS 2  /* CADAP    (cadap.f):     as MOD, with +conj(dens) dz in QB (cadap.f:215) */
/* DAPIF2 with IFLAG7 = 3 (dapif1.f:458-560): the Hilbert-matrix product
 * GW_i = sum_{j != i} DENS_j WDZ_j / (z_i - z_j), WDZ = the plan's complex weights.  Plans only
 * (eb_fmm_plan_create); used by the Laplace side. */
#define EB_VARIANT_HILBERT 3

/* Evaluation modes of the drop-in layer. */
#define EB_MODE_AUTO   0   /* direct O(N^2) sum up to EB_DIRECT_MAX points, FMM above */
#define EB_MODE_FMM    1
#define EB_MODE_DIRECT 2
/* Below about 10^4 points one FMM apply costs its launch latency (0.17-0.21 ms on one MI355X
 * at LEVMAX 5..7, p = 43) while the chunked O(N^2) sum costs 0.024 ms at 2048 points and
 * 0.093 ms at 8192; they cross near 12,000. */
#define EB_DIRECT_MAX  10240

const char *eb_last_error(void);
int eb_version(void);

/* ---------------------------------------------------------------------------------------
 * Layer 1: drop-in replacements.  Interface of CADAPDR, cadapdrop.f:3-5:
 *   (IFLAG1, IFLAG2, LEVMAX, XAT, YAT, DZDT, K0, KBOD, ND, DENS, H, IWORK, LENIW, CWORK,
 *    LENW, XTARG, YTARG, NTARG, GRADW, PHIP, PHIPP, TOL, INFORM, IER, XLEFT, SLENGTH)
 * Supported: IFLAG1 = 0 (evaluate at the sources, self excluded), IFLAG2 = 0 (free space),
 * which is what every caller on the path passes (stokesbubb.f:559-571, electric.f:835-837).
 * With IFLAG1 = 0 the target arrays and NTARG are not read (FASMVP leaves NTARG unset).
 * IWORK/CWORK are not used, but LENIW/LENW are checked exactly as the reference does, so
 * the same inputs fail the same way (IER(2..5) = need/have).
 * --------------------------------------------------------------------------------------- */
int32_t eb_cadapdr(const int32_t *iflag1, const int32_t *iflag2, const int32_t *levmax,
                   const double *xat, const double *yat, const double *dzdt,
                   const int32_t *k0, const int32_t *kbod, const int32_t *nd,
                   const double *dens, const double *h, int32_t *iwork, const int32_t *leniw,
                   double *cwork, const int32_t *lenw, const double *xtarg, const double *ytarg,
                   const int32_t *ntarg, double *gradw, double *phip, double *phipp,
                   const double *tol, int32_t *inform, int32_t *ier, const double *xleft,
                   const double *slength);
/* CADAPMOD, cadapmod.f:3-5 (same arguments) */
int32_t eb_cadapmod(const int32_t *iflag1, const int32_t *iflag2, const int32_t *levmax,
                    const double *xat, const double *yat, const double *dzdt,
                    const int32_t *k0, const int32_t *kbod, const int32_t *nd,
                    const double *dens, const double *h, int32_t *iwork, const int32_t *leniw,
                    double *cwork, const int32_t *lenw, const double *xtarg, const double *ytarg,
                    const int32_t *ntarg, double *gradw, double *phip, double *phipp,
                    const double *tol, int32_t *inform, int32_t *ier, const double *xleft,
                    const double *slength);
/* CADAP, cadap.f:3-5 (same arguments) */
int32_t eb_cadap(const int32_t *iflag1, const int32_t *iflag2, const int32_t *levmax,
                 const double *xat, const double *yat, const double *dzdt,
                 const int32_t *k0, const int32_t *kbod, const int32_t *nd,
                 const double *dens, const double *h, int32_t *iwork, const int32_t *leniw,
                 double *cwork, const int32_t *lenw, const double *xtarg, const double *ytarg,
                 const int32_t *ntarg, double *gradw, double *phip, double *phipp,
                 const double *tol, int32_t *inform, int32_t *ier, const double *xleft,
                 const double *slength);

/* DAPIF2 (dapif1.f:458-459), the adaptive FMM of the Laplace side, with its 18 arguments by
 * reference: (IOUT, IFLAG7, NAT, NAPB, NINIRE, MEX, IERR, INFORM, TOL, EPS7, XA, YA, QA, POTEN,
 * FIELD, WKSP, NSP, CLOSE).  IFLAG7 = 3 only, the Hilbert-matrix product that FASMVP_LAP
 * (laplace.f:214-215) and PVINTEV (pv_integral.f:181-182) call:
 * FIELD(i) = sum_{j != i} QA(j) / (z_i - z_j), QA and FIELD COMPLEX*16.  IERR(1..3) and INFORM(1..4)
 * are set as the reference sets them (IERR 4, 16, 1, 32 and the workspace checks with IERR 8),
 * from the adaptive tree of DAPIF2's rules built on the GPU; WKSP is not used and CLOSE is never
 * called.  Other IFLAG7 values return EB_ERR_UNSUPPORTED with IERR(1) = -1.  Returns IERR(1)
 * or an EB_ERR_* status. */
int32_t eb_dapif2(int32_t *iout, const int32_t *iflag7, const int32_t *nat, const int32_t *napb,
                  const int32_t *ninire, const int32_t *mex, int32_t *ierr, int32_t *inform, const double *tol,
                  const double *eps7, const double *xa, const double *ya, const double *qa, double *poten,
                  double *field, double *wksp, const int32_t *nsp, const void *close);

/* Evaluation mode used by the drop-in layer (process-wide; default EB_MODE_AUTO): the CADAP
 * entry points, eb_fasmvp and eb_dapif2 take the O(N^2) sum or the FMM by it.  No counterpart
 * in the reference (its CADAPDR is always the FMM); a knob of this build for parity runs. */
int eb_set_mode(int mode);
int eb_get_mode(void);

/* ---------------------------------------------------------------------------------------
 * Layer 2: device-resident plan / apply.  All pointers are device pointers; `stream` is a
 * hipStream_t (NULL = default stream).  Complex arrays are interleaved (re, im) doubles.
 * --------------------------------------------------------------------------------------- */
typedef struct eb_fmm_plan_s *eb_fmm_plan;

/* Bin n points (x, y) into the reference's uniform grids of levels 1..levmax
 * (cadap2.f:252-409) and prepare the expansions.  wdz[i] = DZDT(i) * H(body of i): the
 * quadrature-weighted tangent.  nterms = p (the reference uses -NINT(log2 TOL),
 * cadapdrop.f:126).  mode: EB_MODE_FMM or EB_MODE_DIRECT. */
int eb_fmm_plan_create(eb_fmm_plan *plan, int32_t n, const double *x, const double *y,
                       const double *wdz, int32_t levmax, int32_t nterms, int32_t variant,
                       int32_t mode, void *stream);
/* nseg independent point sets in one plan (the segments of a batch of simulations): segment s
 * holds points seg_off[s] .. seg_off[s+1]-1 (host array of nseg + 1 offsets, seg_off[0] = 0,
 * seg_off[nseg] = n, empty segments allowed).  Each segment is binned as CADAPDR bins one point
 * set (its own bounding box and scale, cadapdrop.f:145-171) into its own box of level
 * k = ceil(log4 nseg) of one grid, with levmax levels below that box (total depth k + levmax
 * <= 10), and the sums run within each segment only: gradw[i] is the GW of point i's segment
 * alone.  One launch per FMM stage serves all segments.  FMM mode; eb_fmm_apply,
 * eb_fmm_apply_transpose and eb_fmm_plan_rebuild (same segments) work as for other plans.  No
 * counterpart in the reference, which evaluates one point set per call: a batch of simulations
 * each too large for the O(N^2) sum. */
int eb_fmm_plan_create_segmented(eb_fmm_plan *plan, int32_t n, const double *x, const double *y, const double *wdz,
                                 int32_t nseg, const int32_t *seg_off, int32_t levmax, int32_t nterms,
                                 int32_t variant, void *stream);
/* Re-bin new coordinates / weights of the same n into an existing plan (no allocation);
 * used when the geometry moves (every Runge-Kutta stage). */
int eb_fmm_plan_rebuild(eb_fmm_plan plan, const double *x, const double *y, const double *wdz,
                        void *stream);
/* gradw[i] = GW at point i for density dens (the CADAPDR result with IFLAG1 = 0). */
int eb_fmm_apply(eb_fmm_plan plan, const double *dens, double *gradw, void *stream);
/* mu = A^T lam: the transpose of the real-linear map dens -> gradw of eb_fmm_apply under the real
 * inner product Re sum_i conj(a_i) b_i, so that Re <lam, A dens> = Re <A^T lam, dens>.  It is
 * the same layer potential with sources and targets swapped (for the CADAP family
 * mu_j = sum_{i != j} [conj(c_ij) lam_i + d_ij conj(lam_i)] with GW_i = sum_j [c_ij dens_j +
 * d_ij conj(dens_j)]), evaluated by the plan's own passes (FMM or O(N^2) by its mode) on
 * transposed charges.  No counterpart in the reference, which has no derivatives: it transposes
 * CADAPDR / CADAPMOD / CADAP (cadapdrop.f:3-387) and the Hilbert product of DAPIF2
 * (dapif1.f:540-558), for adjoint solves. */
int eb_fmm_apply_transpose(eb_fmm_plan plan, const double *lam, double *mu, void *stream);
int eb_fmm_plan_destroy(eb_fmm_plan plan);

/* The O(N^2) evaluation of eb_fmm_apply in EB_MODE_DIRECT, in one call that allocates nothing:
 * gradw[i] = GW at point i (caller's order) of the kernel `variant` for n points (x, y),
 * weighted tangents wdz and density dens; every pointer is a device pointer, stream-ordered on
 * `stream`.  No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  Its working memory is a fixed
 * device array of about 8.5 MiB inside the library, held by each GPU it runs on until the
 * process ends, so n is limited to EB_DIRECT_MAX points (larger n: EB_ERR_ARG).  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams, other host threads)
 * race and must be serialised by the caller. */
int eb_direct_apply_noalloc(int32_t n, const double *x, const double *y, const double *wdz, const double *dens,
                            int32_t variant, double *gradw, void *stream);

/* The closed-form O(N^2) sum of the kernel `variant` over all n sources at the ntgt target
 * indices tgt[] (0-based, device int32): gradw[t] = GW at point tgt[t], the point itself
 * excluded.  Device pointers; synchronises `stream` (it allocates and frees its chunk
 * partials).  Independent of the tree, lists and expansions: the checker of accuracy
 * measurements (the autotuner and the benchmark).  No counterpart in the reference, whose
 * direct routines (SLDIRDR / SLSELFDR, cadapdrop2.f:564-769) sum only near boxes. */
int eb_direct_subset(int32_t n, const double *x, const double *y, const double *wdz, const double *dens,
                     int32_t variant, int32_t ntgt, const int32_t *tgt, double *gradw, void *stream);

/* Plan facts: [0] rscal, [1] rminx, [2] rminy; counts: [0] levmax, [1] nterms, [2] kmax
 * (padded p), [3] #ordered near-field pairs, [4] #M2L pairs (both boxes non-empty),
 * [5] #non-empty boxes over levels 2..levmax, [6] #non-empty leaf boxes, [7] max leaf count.
 * Synchronises the plan's stream. */
int eb_fmm_plan_info(eb_fmm_plan plan, double *reals3, int64_t *counts8);
/* The M2L pairs of [4] above by source group: out3[0] pairs whose source is one of the 12
 * ring-2 offsets at centre distance 2 or sqrt 5 box widths (all p input terms), out3[1] pairs
 * whose source is one of the 15 farther ones, out3[2] the input terms those take (about two
 * thirds of p).  For the flop count of the translations.  Synchronises the plan's stream.
 * No counterpart in the reference, whose SLSHIFT (cadap2.f:1149-1215) takes NTERMS terms of
 * every interaction-list source. */
int eb_fmm_plan_m2l_split(eb_fmm_plan plan, int64_t *out3);
/* Non-empty boxes per level, ntbox[0..levmax] (entry l = the reference's NTBOX of level l,
 * cadap2.f:331-336; entry 0 unused).  Synchronises the plan's stream. */
int eb_fmm_plan_ntbox(eb_fmm_plan plan, int32_t *ntbox);

/* Per-stage timing of the last apply on the plan's stream: milliseconds of
 * [charges, P2M, M2M, M2L, L2L, L2P, P2P] measured with HIP events when enabled. */
int eb_fmm_enable_timing(eb_fmm_plan plan, int32_t on);
/* Near-field tile: a leaf of T points becomes ceil(T / tmax) work items of the P2P kernel
 * (tmax in 16..256, 42 by default); the items are rebuilt for the current tree.  A tuning
 * knob of this build (no counterpart in the reference): results change only by rounding. */
int eb_fmm_set_p2p_tile(eb_fmm_plan plan, int32_t tmax);
/* Direct mode: a device int read by the pair-sum kernel of every later apply; where it holds 0
 * the kernel returns at once and the apply's output is left undefined (NULL: always sum).  For
 * operators inside a device-controlled solve, whose products after convergence are not used
 * (the solver's active flag, eb_gmres_batch_active).  No counterpart in the reference, whose
 * solver stops on the host.  FMM-mode plans ignore it. */
int eb_fmm_set_guard(eb_fmm_plan plan, const int32_t *flag);
int eb_fmm_stage_ms(eb_fmm_plan plan, double *ms7);

/* ---------------------------------------------------------------------------------------
 * Tree parity: the reference's per-level binning tables, computed on the GPU.
 * For level ilev (1..15), n points (device x, y), with the bounding box and RSCAL of
 * cadapdrop.f:145-171 recomputed on the device.  Outputs (device, 1-based like the
 * reference): ibox[n] (ASSIGN1 IBOX), icnt[4^ilev] (source histogram), ioffst[4^ilev+1],
 * iatadr[n] (stable by source index), iloc[4^ilev] (first-touch numbering); host scalars
 * ntbox, nsmax.  Synchronises `stream`.
 * --------------------------------------------------------------------------------------- */
int eb_cadap_tables(int32_t ilev, int32_t n, const double *x, const double *y,
                    int32_t *ibox, int32_t *icnt, int32_t *ioffst, int32_t *iatadr,
                    int32_t *iloc, int32_t *ntbox, int32_t *nsmax, void *stream);
/* MKINTL (cadap2.f:780-828) and MKNBOR (cadap2.f:705-742) for every box of level ilev,
 * generated by the same device routine the M2L and near-field kernels use.
 * intl[4^ilev * 27] / nintl[4^ilev], nbor[4^ilev * 9] / nnbor[4^ilev] (device, 1-based ids). */
int eb_cadap_lists(int32_t ilev, int32_t *intl, int32_t *nintl, int32_t *nbor, int32_t *nnbor,
                   void *stream);

#ifdef __cplusplus
}
#endif
#endif
