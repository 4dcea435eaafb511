/*
 * eb_oracle.c -- CPU restatement of the reference algorithm on the Stokes-matvec path.
 *
 * TEST INFRASTRUCTURE ONLY (see eb_oracle.h).  Every function follows the operation order
 * of the reference routine it cites, so that the tree tables come out bit-identical and
 * the floating-point results agree with the reference to rounding.  Pinning:
 *   - tree / lists: bit-exact against the reference library (oracle/_ref) and against
 *     brute-force integer rules in the tests;
 *   - FMM: against the reference CADAPDR and against ora_direct (closed form);
 *   - spectral: ora_dft against numpy.fft, ora_fdifff on trigonometric polynomials,
 *     ora_extractdat on ellipses, ora_filter/ora_equit/ora_ftail against the reference.
 */
#include "eb_oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef double complex cplx;

static inline cplx ld(const double *a, long i) { return CMPLX(a[2 * i], a[2 * i + 1]); }
static inline void st(double *a, long i, cplx v) { a[2 * i] = creal(v); a[2 * i + 1] = cimag(v); }

/* ------------------------------------------------------------------------------------- */
/* Scaling: cadapdrop.f:145-171 (IFLAG2=0 branch) and 199-205.                            */
/* ------------------------------------------------------------------------------------- */
void ora_scale(int n, const double *x, const double *y, double *zs, double *rscal, double *rmin)
{
    double rminx = x[0], rmaxx = x[0], rminy = y[0];
    /* cadapdrop.f:150 initialises RMAXY from YAT(2); the loop covers every point anyway. */
    double rmaxy = (n > 1) ? y[1] : y[0];
    for (int i = 0; i < n; ++i) {
        if (x[i] > rmaxx) rmaxx = x[i];
        if (y[i] > rmaxy) rmaxy = y[i];
        if (x[i] < rminx) rminx = x[i];
        if (y[i] < rminy) rminy = y[i];
    }
    double rsx = rmaxx - rminx, rsy = rmaxy - rminy;
    double r = (rsx > rsy) ? rsx : rsy;
    r = 64.0 / r;
    for (int i = 0; i < n; ++i) {
        zs[2 * i] = (x[i] - rminx) * r;
        zs[2 * i + 1] = (y[i] - rminy) * r;
    }
    *rscal = r;
    rmin[0] = rminx;
    rmin[1] = rminy;
}

/* ------------------------------------------------------------------------------------- */
/* ASSIGN1 (cadap2.f:252-409), sources only.                                              */
/* ------------------------------------------------------------------------------------- */
void ora_assign1(int ilev, int n, const double *zs, int *ibox, int *icnt, int *ioffst,
                 int *iatadr, int *iloc, int *ntbox, int *nsmax)
{
    const int nside = 1 << ilev;
    const int nboxes = nside * nside;
    const double hb = 64.0 / nside;
    int *icnt2 = (int *)malloc(sizeof(int) * nboxes);
    for (int j = 0; j < nboxes; ++j) { icnt2[j] = 1; icnt[j] = 0; iloc[j] = 0; }
    int nt = 0, ns = 0;
    for (int j = 0; j < n; ++j) {
        double X = zs[2 * j], Y = zs[2 * j + 1];
        int ixh = (int)(X / hb), iyh = (int)(Y / hb);
        if (ixh >= nside) ixh = nside - 1;
        if (iyh >= nside) iyh = nside - 1;
        if (ixh < 0) ixh = 0;
        if (iyh < 0) iyh = 0;
        int iadr = iyh * nside + ixh + 1;
        icnt[iadr - 1] += 1;
        ibox[j] = iadr;
        if (icnt[iadr - 1] > ns) ns = icnt[iadr - 1];
        if (iloc[iadr - 1] == 0) iloc[iadr - 1] = ++nt;
    }
    ioffst[0] = 0;
    for (int j = 1; j <= nboxes; ++j) ioffst[j] = ioffst[j - 1] + icnt[j - 1];
    for (int j = 0; j < n; ++j) {
        int iadr = ibox[j];
        int indx = ioffst[iadr - 1] + icnt2[iadr - 1];
        iatadr[indx - 1] = j + 1;
        icnt2[iadr - 1] += 1;
    }
    free(icnt2);
    *ntbox = nt;
    *nsmax = ns;
}

/* MKINTL (cadap2.f:780-828): column-outer, row-inner over the parent's neighbourhood. */
void ora_mkintl(int ilev, int ibox, int *intact, int *nintac)
{
    int nside = 1 << ilev, n = 0;
    int irowb = (ibox - 1) / nside + 1;
    int icolb = ibox - (irowb - 1) * nside;
    int imin, imax, jmin, jmax;
    if (icolb % 2 == 0) { imin = icolb - 3; imax = icolb + 2; }
    else                { imin = icolb - 2; imax = icolb + 3; }
    if (irowb % 2 == 0) { jmin = irowb - 3; jmax = irowb + 2; }
    else                { jmin = irowb - 2; jmax = irowb + 3; }
    if (imin < 1) imin = 1;
    if (imax > nside) imax = nside;
    if (jmin < 1) jmin = 1;
    if (jmax > nside) jmax = nside;
    for (int i = imin; i <= imax; ++i)
        for (int j = jmin; j <= jmax; ++j) {
            if (abs(icolb - i) <= 1 && abs(irowb - j) <= 1) continue;
            intact[n++] = (j - 1) * nside + i;
        }
    *nintac = n;
}

/* MKNBOR (cadap2.f:705-742): clipped 3x3 including the box itself. */
void ora_mknbor(int ilev, int ibox, int *nbors, int *nnbors)
{
    int nside = 1 << ilev, n = 0;
    int irowb = (ibox - 1) / nside + 1;
    int icolb = ibox - (irowb - 1) * nside;
    int imin = icolb - 1 < 1 ? 1 : icolb - 1;
    int imax = icolb + 1 > nside ? nside : icolb + 1;
    int jmin = irowb - 1 < 1 ? 1 : irowb - 1;
    int jmax = irowb + 1 > nside ? nside : irowb + 1;
    for (int i = imin; i <= imax; ++i)
        for (int j = jmin; j <= jmax; ++j) nbors[n++] = (j - 1) * nside + i;
    *nnbors = n;
}

/* ------------------------------------------------------------------------------------- */
/* Expansion kernels                                                                      */
/* ------------------------------------------------------------------------------------- */

/* SLMKXP (cadap2.f:412-468) */
static void slmkxp(cplx zc, const cplx *zs, const cplx *qa, const cplx *qb, const cplx *qc,
                   int nin, cplx *m1, cplx *m2, int p)
{
    for (int i = 0; i < p; ++i) { m1[i] = 0; m2[i] = 0; }
    for (int i = 0; i < nin; ++i) {
        cplx zsh = zs[i] - zc;
        cplx z0phi = -qa[i], z0psi = -qb[i];
        for (int j = 0; j < p; ++j) {
            m1[j] += z0phi;
            m2[j] += z0psi;
            z0phi *= zsh;
            z0psi *= zsh;
        }
        z0psi = qc[i];
        for (int j = 2; j <= p; ++j) {
            m2[j - 1] += z0psi * (double)(j - 1);
            z0psi *= zsh;
        }
    }
}

/* SLMVALDR (cadapdrop2.f:251-301) / SLMVAL (cadap2.f:470-520): sgn=-1 / +1 */
static void slmval(double sgn, const cplx *m1, const cplx *m2, cplx zc, int p, cplx zt, cplx *gw)
{
    cplx zsinv = 1.0 / (zt - zc);
    cplx phi = m1[p - 1] * zsinv;
    cplx phip = (double)p * m1[p - 1] * zsinv;
    cplx psi = m2[p - 1] * zsinv;
    for (int j = p - 1; j >= 1; --j) {
        phi = (phi + m1[j - 1]) * zsinv;
        psi = (psi + m2[j - 1]) * zsinv;
        phip = (phip + (double)j * m1[j - 1]) * zsinv;
    }
    phip = -phip * zsinv;
    *gw += phi + sgn * (zt * conj(phip)) + sgn * conj(psi);
}

/* SLLVALDR (cadapdrop2.f:304-350) / SLLVAL (cadap2.f:575-621) */
static void sllval(double sgn, const cplx *l1, const cplx *l2, int p, cplx zc, cplx zt, cplx *gw)
{
    cplx zsh = zt - zc, phi = 0, psi = 0, phip = 0;
    for (int j = p; j >= 0; --j) {
        phi = l1[j] + zsh * phi;
        psi = l2[j] + zsh * psi;
    }
    for (int j = p; j >= 1; --j) phip = (double)j * l1[j] + zsh * phip;
    *gw += phi + sgn * (zt * conj(phip)) + sgn * conj(psi);
}

/* SLSHIFT (cadap2.f:1149-1215) with the Pascal table C(i,j)=binom(i-1,j-1) (cadapdrop2.f:53-63) */
static void slshift(const cplx *m1, const cplx *m2, cplx zc, int n, cplx zc2, cplx *b1, cplx *b2,
                    const double *C /* 120x120 column-major, 1-based (i,j) -> C[(j-1)*120+(i-1)] */)
{
    cplx z0 = zc2 - zc;
    cplx cd = 1.0 / z0, cdd = cd;
    cplx z0p[61], mm1[60], mm2[60];
    z0p[0] = 1.0;
    for (int i = 1; i <= n; ++i) { z0p[i] = cdd; cdd *= cd; }
    for (int i = 1; i <= n; ++i) {
        mm1[i - 1] = m1[i - 1] * z0p[i];
        mm2[i - 1] = m2[i - 1] * z0p[i];
        b1[i] = 0; b2[i] = 0;
    }
    b1[0] = 0; b2[0] = 0;
    for (int m = 0; m <= n - 1; ++m)
        for (int k = 1; k <= n; ++k) {
            double c = C[(k - 1) * 120 + (m + k - 1)];
            b1[m] += mm1[k - 1] * c;
            b2[m] += mm2[k - 1] * c;
        }
    for (int m = 0; m <= n - 1; m += 2) { b1[m] *= z0p[m]; b2[m] *= z0p[m]; }
    for (int m = 1; m <= n - 1; m += 2) { b1[m] = -b1[m] * z0p[m]; b2[m] = -b2[m] * z0p[m]; }
}

/* SLDIRDR (cadapdrop2.f:564-620) / SLDIR (cadap2.f:1090-1146), I1=0 branch */
static void sldir(double sgn, cplx zt, const cplx *zs, int ns, const cplx *qa, const cplx *qb,
                  const cplx *qc, cplx *gw)
{
    for (int j = 0; j < ns; ++j) {
        cplx zinv = 1.0 / (zs[j] - zt);
        cplx zinv2c = conj(zinv * zinv);
        *gw += qa[j] * zinv + sgn * conj(qb[j] * zinv) + sgn * (zt * conj(qa[j]) + conj(qc[j])) * zinv2c;
    }
}

/* SLSELFDR (cadapdrop2.f:724-769) / SLSELF (cadap2.f:1319-1364) */
static void slself(double sgn, const cplx *zs, int ioff, const int *iatadr, int ns, const cplx *qa,
                   const cplx *qb, const cplx *qc, cplx *gw)
{
    for (int k = 0; k < ns; ++k) {
        cplx phi = 0, phip = 0, psi = 0;
        int kt = iatadr[ioff + k] - 1;
        for (int j = 0; j < ns; ++j) {
            if (j == k) continue;
            cplx zd = zs[j] - zs[k];
            cplx zd2 = zd * zd;
            phi += qa[j] / zd;
            phip += qa[j] / zd2;
            psi += qb[j] / zd;
            psi += qc[j] / zd2;
        }
        gw[kt] += phi + sgn * (zs[k] * conj(phip)) + sgn * conj(psi);
    }
}

/* ------------------------------------------------------------------------------------- */
/* CADAPDR / CADAPMOD / CADAP (cadapdrop.f:3-387), IFLAG1=0, IFLAG2=0                      */
/* ------------------------------------------------------------------------------------- */
int ora_cadap(int variant, int levmax, int kbod, const int *nd, const double *x, const double *y,
              const double *dzdt, const double *dens, const double *h, double tol,
              double *gradw, int *inform, int *ier, long long *counts)
{
    const double pi = 4.0 * atan(1.0);
    const cplx z2pi = 1.0 / (2.0 * pi * I);
    const double sgn = (variant == ORA_DR) ? -1.0 : 1.0;   /* GW sign of z conj(phi') + conj(psi) */
    const double sqb = (variant == ORA_PLUS) ? 1.0 : -1.0;  /* sign of conj(dens) dz in QB */
    long long nshift = 0, nslmval = 0, nl2p = 0, npair = 0;

    for (int i = 0; i < 5; ++i) ier[i] = 0;
    int nterms = -(int)lround(log(tol) / log(2.0));
    inform[1] = nterms;
    if (nterms > 60) { ier[0] = 4; ier[1] = nterms; return 4; }

    int n = 0;
    for (int k = 0; k < kbod; ++k) n += nd[k];

    double rscal, rmin[2];
    double *zsd = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    ora_scale(n, x, y, zsd, &rscal, rmin);

    cplx *zs = (cplx *)malloc(sizeof(cplx) * n), *qa = (cplx *)malloc(sizeof(cplx) * n);
    cplx *qb = (cplx *)malloc(sizeof(cplx) * n), *qc = (cplx *)malloc(sizeof(cplx) * n);
    cplx *gw = (cplx *)calloc((size_t)n, sizeof(cplx));
    int j = 0;
    for (int k = 0; k < kbod; ++k)
        for (int nk = 0; nk < nd[k]; ++nk, ++j) {
            zs[j] = CMPLX(zsd[2 * j], zsd[2 * j + 1]);
            cplx d = ld(dens, j), w = ld(dzdt, j);
            qa[j] = z2pi * d * w * h[k];
            qb[j] = (sqb * conj(d) * w + d * conj(w)) * z2pi * h[k];
            qc[j] = -(z2pi * conj(zs[j]) * d * w * h[k]);
        }

    /* Pascal table, cadapdrop2.f:53-63 (column-major, 1-based). */
    double *C = (double *)malloc(sizeof(double) * 120 * 120);
#define CI(i, jj) C[((jj) - 1) * 120 + ((i) - 1)]
    CI(1, 1) = 1.0;
    for (int i = 2; i <= 120; ++i) { CI(1, i) = 0.0; CI(i, 1) = 1.0; }
    for (int i = 2; i <= 120; ++i)
        for (int jj = 2; jj <= 120; ++jj) CI(i, jj) = CI(i - 1, jj - 1) + CI(i - 1, jj);
#undef CI

    const int p = nterms, nlmax = nterms / 2;
    int *ibox = (int *)malloc(sizeof(int) * n), *iatadr = (int *)malloc(sizeof(int) * n);
    cplx *zs2 = (cplx *)malloc(sizeof(cplx) * n), *qa2 = (cplx *)malloc(sizeof(cplx) * n);
    cplx *qb2 = (cplx *)malloc(sizeof(cplx) * n), *qc2 = (cplx *)malloc(sizeof(cplx) * n);
    cplx m1[60], m2[60], b1[61], b2[61];
    int intact[30], nintac, nbors[30], nnbors;
    int *icnt = NULL, *ioffst = NULL, *iloc = NULL;
    int ilev = 1;
    for (ilev = 1; ilev <= levmax; ++ilev) {
        inform[0] = ilev;
        const int nside = 1 << ilev, nboxes = nside * nside;
        const double hb = 64.0 / nside;
        free(icnt); free(ioffst); free(iloc);
        icnt = (int *)malloc(sizeof(int) * nboxes);
        ioffst = (int *)malloc(sizeof(int) * (nboxes + 1));
        iloc = (int *)malloc(sizeof(int) * nboxes);
        int ntbox, nsmax;
        ora_assign1(ilev, n, zsd, ibox, icnt, ioffst, iatadr, iloc, &ntbox, &nsmax);

        /* DOINT0DR (cadapdrop2.f:2-249) */
        cplx *loc1 = (cplx *)calloc((size_t)(p + 1) * ntbox, sizeof(cplx));
        cplx *loc2 = (cplx *)calloc((size_t)(p + 1) * ntbox, sizeof(cplx));
        for (int ib = 1; ib <= nboxes; ++ib) {
            int ioff = ioffst[ib - 1], nin = ioffst[ib] - ioff;
            for (int kk = 0; kk < nin; ++kk) {
                int a = iatadr[ioff + kk] - 1;
                zs2[kk] = zs[a]; qa2[kk] = qa[a]; qb2[kk] = qb[a]; qc2[kk] = qc[a];
            }
            if (nin <= 0) continue;
            int icol = 1 + (ib - 1) % nside, irow = 1 + (ib - 1) / nside;
            cplx zc = CMPLX(icol * hb - hb / 2, irow * hb - hb / 2);
            slmkxp(zc, zs2, qa2, qb2, qc2, nin, m1, m2, p);
            ora_mkintl(ilev, ib, intact, &nintac);
            for (int jl = 0; jl < nintac; ++jl) {
                int iadr = intact[jl];
                int inoff = ioffst[iadr - 1], ninnbr = ioffst[iadr] - inoff;
                if (ninnbr <= nlmax) {
                    for (int kk = 0; kk < ninnbr; ++kk) {
                        int jt = iatadr[inoff + kk] - 1;
                        slmval(sgn, m1, m2, zc, p, zs[jt], &gw[jt]);
                        ++nslmval;
                    }
                } else {
                    int icol2 = 1 + (iadr - 1) % nside, irow2 = 1 + (iadr - 1) / nside;
                    cplx zc2 = CMPLX(icol2 * hb - hb / 2, irow2 * hb - hb / 2);
                    int iadloc = iloc[iadr - 1] - 1;
                    slshift(m1, m2, zc, p, zc2, b1, b2, C);
                    ++nshift;
                    for (int t = 0; t <= p; ++t) {
                        loc1[(size_t)iadloc * (p + 1) + t] += b1[t];
                        loc2[(size_t)iadloc * (p + 1) + t] += b2[t];
                    }
                }
            }
        }
        for (int ib = 1; ib <= nboxes; ++ib) {
            int inoff = ioffst[ib - 1], nin = ioffst[ib] - inoff;
            if (nin <= nlmax) continue;
            int icol = 1 + (ib - 1) % nside, irow = 1 + (ib - 1) / nside;
            cplx zc = CMPLX(icol * hb - hb / 2, irow * hb - hb / 2);
            int jadr = iloc[ib - 1] - 1;
            for (int kk = 0; kk < nin; ++kk) {
                int jt = iatadr[inoff + kk] - 1;
                sllval(sgn, &loc1[(size_t)jadr * (p + 1)], &loc2[(size_t)jadr * (p + 1)], p, zc, zs[jt], &gw[jt]);
                ++nl2p;
            }
        }
        free(loc1); free(loc2);
        if (ilev == levmax) break;
    }

    /* DONN0DR (cadapdrop2.f:623-722) at the last level */
    {
        const int nside = 1 << ilev, nboxes = nside * nside;
        for (int ib = 1; ib <= nboxes; ++ib) {
            int ioff = ioffst[ib - 1], nin = ioffst[ib] - ioff;
            if (nin <= 0) continue;
            for (int kk = 0; kk < nin; ++kk) {
                int a = iatadr[ioff + kk] - 1;
                zs2[kk] = zs[a]; qa2[kk] = qa[a]; qb2[kk] = qb[a]; qc2[kk] = qc[a];
            }
            slself(sgn, zs2, ioff, iatadr, nin, qa2, qb2, qc2, gw);
            npair += (long long)nin * (nin - 1);
            ora_mknbor(ilev, ib, nbors, &nnbors);
            for (int jn = 0; jn < nnbors; ++jn) {
                int iadr = nbors[jn];
                int inoff = ioffst[iadr - 1], ninnbr = ioffst[iadr] - inoff;
                if (iadr == ib) continue;
                for (int kk = 0; kk < ninnbr; ++kk) {
                    int jt = iatadr[inoff + kk] - 1;
                    sldir(sgn, zs[jt], zs2, nin, qa2, qb2, qc2, &gw[jt]);
                }
                npair += (long long)nin * ninnbr;
            }
        }
    }
    for (int i = 0; i < n; ++i) st(gradw, i, gw[i] * rscal);

    if (counts) { counts[0] = nshift; counts[1] = nslmval; counts[2] = nl2p; counts[3] = npair; }
    free(icnt); free(ioffst); free(iloc); free(ibox); free(iatadr);
    free(zs2); free(qa2); free(qb2); free(qc2); free(C);
    free(zs); free(qa); free(qb); free(qc); free(gw); free(zsd);
    return 0;
}

/* ------------------------------------------------------------------------------------- */
/* Direct sum (closed form of the cadap kernel), long double                             */
/* ------------------------------------------------------------------------------------- */
typedef long double complex lcplx;

/* The target loop runs on up to ORA_THREADS host threads (environment variable, default the
 * online cores, at most 64); each target's sum over the sources runs in source order on one
 * thread, so the result does not depend on the thread count. */
typedef struct {
    const lcplx *z, *qa, *qb;
    int n, ntgt, kind;       /* kind 0: cadap kernel with sign sgn; 1: Cauchy sum; 2, 3: transposes */
    long double sgn;
    const int *tgt;
    double *out;
    int t0, t1;
    long double sqb;         /* kind 2 */
} ora_sum_job;

/* kind 2 (the CADAP family) and 3 (Hilbert): qa holds the weights wdz, qb the multipliers lam.
 * The forward's coefficients of dens_j and conj(dens_j) in GW_i, with d = z_j - z_i (source j,
 * target i), alpha = wdz/(2 pi i), gamma = conj(wdz)/(2 pi i):
 *   c_ij = alpha_j / d + sgn sqb conj(alpha_j) / conj(d)
 *   d_ij = sgn conj(gamma_j) / conj(d) - sgn conj(alpha_j) d / conj(d)^2
 * (from qa = alpha dens, qb = sqb alpha conj(dens) + gamma dens in the closed form below), and
 * for the Hilbert product c_ij = wdz_j / (z_i - z_j), d_ij = 0.  Output j sums over i. */
static void ora_tsum_target(const ora_sum_job *J, int t)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    const lcplx z2pi = 1.0L / (2.0L * pi * I);
    const int j = J->tgt[t];
    const lcplx w = J->qa[j];
    const lcplx al = w * z2pi, ga = conjl(w) * z2pi;
    lcplx acc = 0;
    for (int i = 0; i < J->n; ++i) {
        if (i == j) continue;
        const lcplx d = J->z[j] - J->z[i];
        const lcplx lam = J->qb[i];
        lcplx c, dc;
        if (J->kind == 3) {
            c = w / (J->z[i] - J->z[j]);
            dc = 0;
        } else {
            const lcplx cd = conjl(d);
            c = al / d + J->sgn * J->sqb * conjl(al) / cd;
            dc = J->sgn * conjl(ga) / cd - J->sgn * conjl(al) * d / (cd * cd);
        }
        acc += conjl(c) * lam + dc * conjl(lam);
    }
    J->out[2 * t] = (double)creall(acc);
    J->out[2 * t + 1] = (double)cimagl(acc);
}

static void *ora_sum_run(void *arg)
{
    ora_sum_job *J = (ora_sum_job *)arg;
    for (int t = J->t0; t < J->t1; ++t) {
        if (J->kind >= 2) {
            ora_tsum_target(J, t);
            continue;
        }
        const int i = J->tgt[t];
        lcplx acc = 0;
        for (int s = 0; s < J->n; ++s) {
            if (s == i) continue;
            lcplx dd = J->z[s] - J->z[i];
            long double r2 = creall(dd) * creall(dd) + cimagl(dd) * cimagl(dd);
            lcplx inv = conjl(dd) / r2;           /* 1/d */
            if (J->kind == 1) {
                acc -= J->qa[s] * inv;            /* qa_s / (z_i - z_s) */
                continue;
            }
            lcplx cinv = dd / r2;                 /* 1/conj(d) */
            acc += J->qa[s] * inv + J->sgn * (conjl(J->qb[s]) * cinv - conjl(J->qa[s]) * dd * cinv * cinv);
        }
        J->out[2 * t] = (double)creall(acc);
        J->out[2 * t + 1] = (double)cimagl(acc);
    }
    return NULL;
}

static void ora_sum_parallel(ora_sum_job base)
{
    int nth = 1;
    const char *env = getenv("ORA_THREADS");
    if (env && atoi(env) > 0) nth = atoi(env);
    else {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        nth = c > 0 ? (int)c : 1;
    }
    if (nth > 64) nth = 64;
    if (nth > base.ntgt) nth = base.ntgt > 0 ? base.ntgt : 1;
    /* below some 10^6 pairs the threads cost more than they save */
    if ((long long)base.ntgt * base.n < 4000000LL) nth = 1;
    ora_sum_job jobs[64];
    pthread_t th[64];
    int started[64] = {0};
    for (int k = 0; k < nth; ++k) {
        jobs[k] = base;
        jobs[k].t0 = (int)((long long)base.ntgt * k / nth);
        jobs[k].t1 = (int)((long long)base.ntgt * (k + 1) / nth);
        if (k > 0 && pthread_create(&th[k], NULL, ora_sum_run, &jobs[k]) == 0) started[k] = 1;
    }
    ora_sum_run(&jobs[0]);
    for (int k = 1; k < nth; ++k) {
        if (started[k]) pthread_join(th[k], NULL);
        else ora_sum_run(&jobs[k]);
    }
}

void ora_direct(int variant, int kbod, const int *nd, const double *x, const double *y,
                const double *dzdt, const double *dens, const double *h,
                int ntgt, const int *tgt, double *gradw)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    const lcplx z2pi = 1.0L / (2.0L * pi * I);
    const long double sgn = (variant == ORA_DR) ? -1.0L : 1.0L;
    const long double sqb = (variant == ORA_PLUS) ? 1.0L : -1.0L;
    int n = 0;
    for (int k = 0; k < kbod; ++k) n += nd[k];
    lcplx *qa = (lcplx *)malloc(sizeof(lcplx) * n), *qb = (lcplx *)malloc(sizeof(lcplx) * n);
    lcplx *z = (lcplx *)malloc(sizeof(lcplx) * n);
    int j = 0;
    for (int k = 0; k < kbod; ++k)
        for (int nk = 0; nk < nd[k]; ++nk, ++j) {
            lcplx d = (long double)dens[2 * j] + I * (long double)dens[2 * j + 1];
            lcplx w = (long double)dzdt[2 * j] + I * (long double)dzdt[2 * j + 1];
            z[j] = (long double)x[j] + I * (long double)y[j];
            qa[j] = z2pi * d * w * (long double)h[k];
            qb[j] = (sqb * conjl(d) * w + d * conjl(w)) * z2pi * (long double)h[k];
        }
    ora_sum_job J = {z, qa, qb, n, ntgt, 0, sgn, tgt, gradw, 0, 0};
    ora_sum_parallel(J);
    free(qa); free(qb); free(z);
}

void ora_direct_transpose(int variant, int n, const double *x, const double *y, const double *wdz,
                          const double *lam, int ntgt, const int *tgt, double *mu)
{
    lcplx *z = (lcplx *)malloc(sizeof(lcplx) * n), *w = (lcplx *)malloc(sizeof(lcplx) * n);
    lcplx *l = (lcplx *)malloc(sizeof(lcplx) * n);
    for (int j = 0; j < n; ++j) {
        z[j] = (long double)x[j] + I * (long double)y[j];
        w[j] = (long double)wdz[2 * j] + I * (long double)wdz[2 * j + 1];
        l[j] = (long double)lam[2 * j] + I * (long double)lam[2 * j + 1];
    }
    ora_sum_job J = {z, w, l, n, ntgt, variant == 3 ? 3 : 2, (variant == ORA_DR) ? -1.0L : 1.0L, tgt, mu, 0, 0,
                     (variant == ORA_PLUS) ? 1.0L : -1.0L};
    ora_sum_parallel(J);
    free(z); free(w); free(l);
}

/* ------------------------------------------------------------------------------------- */
/* Hilbert-matrix product of DAPIF2 with IFLAG7 = 3 (dapif1.f:540-558, the direct form    */
/* DADIRE dapif1.f:104-311): FIELD_i = sum_{j != i} qa_j / (z_i - z_j), long double.      */
/* ------------------------------------------------------------------------------------- */
void ora_cauchy(int n, const double *x, const double *y, const double *qa, int ntgt, const int *tgt, double *field)
{
    lcplx *q = (lcplx *)malloc(sizeof(lcplx) * n), *z = (lcplx *)malloc(sizeof(lcplx) * n);
    for (int j = 0; j < n; ++j) {
        z[j] = (long double)x[j] + I * (long double)y[j];
        q[j] = (long double)qa[2 * j] + I * (long double)qa[2 * j + 1];
    }
    ora_sum_job J = {z, q, NULL, n, ntgt, 1, 0.0L, tgt, field, 0, 0};
    ora_sum_parallel(J);
    free(q); free(z);
}

/* ------------------------------------------------------------------------------------- */
/* FASMVP epilogue (stokesbubb.f:527-549, 585-611)                                        */
/* ------------------------------------------------------------------------------------- */
void ora_fasmvp_epilogue(int k, const int *nd, const double *x, const double *y, const double *dz,
                         const double *dsdth, const double *rkappa, const double *zk,
                         const double *u, const double *stress, double *w, double *za, double *pk)
{
    const double pi = 4.0 * atan(1.0);
    const cplx eye = I;
    int nbk = 0;
    for (int kb = 0; kb < k; ++kb) nbk += nd[kb];
    cplx *dens = (cplx *)malloc(sizeof(cplx) * nbk);
    for (int i = 0; i < nbk; ++i) dens[i] = CMPLX(u[2 * i], u[2 * i + 1]);
    int istart = 0;
    cplx *zav = (cplx *)malloc(sizeof(cplx) * k);
    for (int kb = 0; kb < k; ++kb) {
        double hk = 2.0 * pi / nd[kb];
        zav[kb] = 0; pk[kb] = 0;
        for (int i = 0; i < nd[kb]; ++i) {
            cplx z = CMPLX(x[istart + i], y[istart + i]);
            cplx zdis = z - ld(zk, kb);
            double zx = creal(dens[istart + i] * ld(dz, istart + i) / (zdis * zdis));
            pk[kb] = pk[kb] + 2.0 * hk * zx / pi;
            zav[kb] = zav[kb] + hk * dens[istart + i];
        }
        istart += nd[kb];
    }
    double pinf = -pk[0];
    for (int kb = 0; kb < k; ++kb) pk[kb] += pinf;
    istart = 0;
    for (int kb = 0; kb < k; ++kb) {
        double qwi = 2.0 * pi / nd[kb];
        for (int i = 0; i < nd[kb]; ++i) {
            int g = istart + i;
            cplx zu = dens[g];
            cplx z = CMPLX(x[g], y[g]);
            cplx zn = -eye * ld(dz, g) / dsdth[g];
            cplx zself = zu * qwi * dsdth[g] * 0.5 * rkappa[g] / pi;
            zself = zself - conj(zu) * qwi * dsdth[g] * 0.5 * rkappa[g] * zn * zn / pi;
            cplx s = ld(stress, g) + zself - 0.5 * eye * (pinf - pk[kb]) * z;
            cplx out = dens[g] + zav[kb] + s;
            w[2 * g] = creal(out);
            w[2 * g + 1] = cimag(out);
        }
        istart += nd[kb];
    }
    for (int kb = 0; kb < k; ++kb) st(za, kb, zav[kb]);
    free(dens); free(zav);
}

/* ------------------------------------------------------------------------------------- */
/* Spectral                                                                               */
/* ------------------------------------------------------------------------------------- */
void ora_dft(int n, double *c, int sign)
{
    const long double twopi = 6.28318530717958647692528676655900577L;
    lcplx *tw = (lcplx *)malloc(sizeof(lcplx) * n), *out = (lcplx *)malloc(sizeof(lcplx) * n);
    for (int m = 0; m < n; ++m) {
        long double a = twopi * (long double)m / (long double)n;
        tw[m] = cosl(a) + (long double)sign * I * sinl(a);
    }
    for (int j = 0; j < n; ++j) {
        lcplx acc = 0;
        for (int k = 0; k < n; ++k) {
            int m = (int)(((long long)j * k) % n);
            acc += ((long double)c[2 * k] + I * (long double)c[2 * k + 1]) * tw[m];
        }
        out[j] = acc;
    }
    for (int j = 0; j < n; ++j) { c[2 * j] = (double)creall(out[j]); c[2 * j + 1] = (double)cimagl(out[j]); }
    free(tw); free(out);
}

void ora_fdifff(int n, const double *f, double *df)
{
    memcpy(df, f, sizeof(double) * 2 * n);
    ora_dft(n, df, -1);
    for (int i = 0; i < n; ++i) st(df, i, ld(df, i) / (double)n);
    int nsc, nscr;
    if (n % 2 == 0) { nsc = (n - 2) / 2; nscr = nsc + 1; }
    else            { nsc = (n - 1) / 2; nscr = nsc; }
    for (int k = -nsc; k <= -1; ++k) st(df, n + k, (double)k * I * ld(df, n + k));
    st(df, 0, 0);
    for (int k = 1; k <= nscr; ++k) st(df, k, (double)k * I * ld(df, k));
    if (n % 2 == 0) st(df, nscr, 0);
    ora_dft(n, df, +1);
}

void ora_extractdat(int k, const int *nd, const double *rku, double *x, double *y, double *zk,
                    double *dz, double *tha, double *dsda, double *rkappa, double *ckern,
                    double *area, double *xleft, double *slength)
{
    const double pi = 4.0 * atan(1.0);
    int nbk = 0, ndm = 0;
    for (int kb = 0; kb < k; ++kb) { nbk += nd[kb]; if (nd[kb] > ndm) ndm = nd[kb]; }
    double *za = (double *)malloc(sizeof(double) * 2 * ndm), *zai = (double *)malloc(sizeof(double) * 2 * ndm);
    double xmin = 1e10, xmax = -1e10, ymin = 1e10, ymax = -1e10;
    int istart = 0;
    for (int kb = 0; kb < k; ++kb) {
        int m = nd[kb];
        for (int i = 0; i < m; ++i) {
            int g = istart + i;
            x[g] = rku[2 * g]; y[g] = rku[2 * g + 1];
            xmin = fmin(xmin, x[g]); xmax = fmax(xmax, x[g]);
            ymin = fmin(ymin, y[g]); ymax = fmax(ymax, y[g]);
            za[2 * i] = x[g]; za[2 * i + 1] = y[g];
        }
        ora_dft(m, za, -1);
        zk[2 * kb] = za[0] / m; zk[2 * kb + 1] = za[1] / m;
        ora_fdifff(m, &rku[2 * istart], za);
        ora_fdifff(m, za, zai);
        for (int i = 0; i < m; ++i) {
            int g = istart + i;
            double xdot = za[2 * i], ydot = za[2 * i + 1], xddot = zai[2 * i], yddot = zai[2 * i + 1];
            double dsdt = sqrt(xdot * xdot + ydot * ydot);
            rkappa[g] = -(xdot * yddot - ydot * xddot) / (dsdt * dsdt * dsdt);
            cplx ck = -CMPLX(xddot, yddot) / CMPLX(xdot, ydot);
            st(ckern, g, ck);
            tha[g] = -rkappa[g] * dsdt;
            dz[2 * g] = -xdot; dz[2 * g + 1] = -ydot;
            dsda[g] = cabs(CMPLX(-xdot, -ydot));
            za[2 * i] = y[g] * (-xdot); za[2 * i + 1] = 0.0;
        }
        ora_dft(m, za, -1);
        area[kb] = 2.0 * pi * (za[0] / m);
        istart += m;
    }
    *xleft = xmin;
    *slength = fmax(xmax - xmin, ymax);
    free(za); free(zai);
}

/* FOURINT (electric.f:1307-1331) */
static void fourint(int nd, const double *zf, double *zfi)
{
    int n = nd / 2;
    st(zfi, 0, 0);
    for (int i = 1; i <= n - 1; ++i) st(zfi, i, ld(zf, i) / ((double)i * I));
    st(zfi, n, 0);
    for (int i = 1; i <= n - 1; ++i) st(zfi, nd - i, -ld(zf, nd - i) / ((double)i * I));
}

void ora_equit(int nd, const double *tha, const double *un, double *t)
{
    const double pi = 4.0 * atan(1.0);
    double dalph = 2.0 * pi / nd;
    int n = nd / 2;
    double *zf = (double *)malloc(sizeof(double) * 2 * nd), *zfi = (double *)malloc(sizeof(double) * 2 * nd);
    for (int i = 0; i < nd; ++i) { zf[2 * i] = -tha[i] * un[i]; zf[2 * i + 1] = 0.0; }
    ora_dft(nd, zf, -1);
    for (int i = 0; i < nd; ++i) st(zf, i, ld(zf, i) / (double)nd);
    st(zf, n, 0);
    double dadt = zf[0];
    fourint(nd, zf, zfi);
    ora_dft(nd, zfi, +1);
    for (int i = 0; i < nd; ++i) {
        double alpha = dalph * (double)i;
        t[i] = -creal(ld(zf, 0) * alpha + ld(zfi, i)) + alpha * dadt;
    }
    free(zf); free(zfi);
}

static double filt_factor(int i, int n)
{
    double ratio = (double)((float)i / (float)n);   /* single-precision division, electric.f:1163 */
    double r = 1.0, b = ratio;                      /* ratio**25 */
    for (int e = 25; e > 0; e >>= 1) { if (e & 1) r *= b; b *= b; }
    return exp(-10.0 * r);
}

void ora_filter(int nd, double *cn)
{
    const int n = nd / 2;
    const double eps = 1e-13;
    for (int i = 0; i <= n; ++i) {
        double f = filt_factor(i, n);
        cn[2 * i] *= f; cn[2 * i + 1] *= f;
        if (fabs(cn[2 * i]) < eps) cn[2 * i] = 0.0;
        if (fabs(cn[2 * i + 1]) < eps) cn[2 * i + 1] = 0.0;
    }
    cn[2 * n] = 0.0; cn[2 * n + 1] = 0.0;
    for (int i = 1; i <= n - 1; ++i) {
        int q = nd - i;
        double f = filt_factor(i, n);
        cn[2 * q] *= f; cn[2 * q + 1] *= f;
        if (fabs(cn[2 * q]) < eps) cn[2 * q] = 0.0;
        if (fabs(cn[2 * q + 1]) < eps) cn[2 * q + 1] = 0.0;
    }
}

void ora_ftail(int k, const int *nd, const double *tha, const double *un, const double *dz,
               const double *dsda, double *rkup)
{
    int istart = 0, ndm = 0;
    for (int kb = 0; kb < k; ++kb) if (nd[kb] > ndm) ndm = nd[kb];
    double *t = (double *)malloc(sizeof(double) * ndm), *zf = (double *)malloc(sizeof(double) * 2 * ndm);
    for (int kb = 0; kb < k; ++kb) {
        int m = nd[kb];
        ora_equit(m, &tha[istart], &un[istart], t);
        for (int i = 0; i < m; ++i) {
            int g = istart + i;
            cplx dzi = ld(dz, g);
            cplx v = -(un[g] * I * dzi) - t[i] * dzi;
            st(zf, i, v / dsda[g]);
        }
        ora_dft(m, zf, -1);
        for (int i = 0; i < m; ++i) st(zf, i, ld(zf, i) / (double)m);
        ora_filter(m, zf);
        ora_dft(m, zf, +1);
        memcpy(&rkup[2 * istart], zf, sizeof(double) * 2 * m);
        istart += m;
    }
    free(t); free(zf);
}

void ora_rk2_stage1(int neq, double dt, const double *ybar, const double *dy, double *phi0, double *y)
{
    for (int i = 0; i < neq; ++i) { phi0[i] = dt * dy[i]; y[i] = ybar[i] + phi0[i]; }
}

void ora_rk2_stage2(int neq, double dt, const double *phi0, const double *dy, double *ybar)
{
    for (int i = 0; i < neq; ++i) {
        double phi1 = dt * dy[i];
        ybar[i] = ybar[i] + 0.5 * phi0[i] + 0.5 * phi1;
    }
}
