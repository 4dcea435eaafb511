/*
 * eb_oracle_atree.c -- CPU restatement of the adaptive quadtree of DAPIF2 (the tree that the
 * Laplace FMM builds, used here as the adaptive tree definition of the FMM-alone workload).
 *
 * TEST INFRASTRUCTURE ONLY (see eb_oracle.h).  Follows, in operation order:
 *   scaling             dapif1.f:740-770 (centre, RSCAL = 128/max range * REAL*4 0.99)
 *   DAASSI/DAINIA/DAIBNB/DAFAMI  dapif1.f:1165-1453 (split boxes with > NAPB points;
 *                       children NW, NE, SW, SE; level-by-level numbering)
 *   DACOMP              dapif1.f:1456-1539 (drop empty boxes)
 *   DAINDB              dapif1.f:1542-1606 (childless flags, MAXP)
 *   DAIJAT              dapif1.f:1609-1680 (per-leaf particle lists)
 *   DALI12/DAIN12       dapif2.f:3-210 (interaction list #1 and colleagues #2)
 *   DALI34/DASCAN       dapif2.f:212-474 (lists #3 and #4 of childless boxes)
 * All index tables hold the reference's 1-based values.  Internally the arrays are used
 * with 1-based subscripts (element 0 unused) so that each step reads like its source.
 *
 * One out-of-bounds read of the reference is given a defined value: DACOMP evaluates
 * SHIFT(NCHI(I)) and SHIFT(NPAR(I)) with NCHI = 0 (childless) and NPAR = 0 (first level),
 * i.e. SHIFT(0), the workspace word before SHIFT.  DAPIF2 zeroes its whole workspace on
 * entry (dapif1.f:690-695) and that word is only ever written when NBOX reaches NBOXMX at a
 * childless box, so it reads 0 and the entries stay 0.  Here shift[0] = 0.
 *
 * Pinned by tests/test_oracle_cpu.py against the reference routines themselves (DAASSI,
 * DACOMP, DAINDB, DAIJAT, DALI12, DALI34 called from oracle/_ref/libref.so on the scaled
 * coordinates that DAPIF2 leaves in its workspace), through tests/golden/atree_*.npz.
 */
#include "eb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int *a;
    long n, cap;
} ivec;

static void iv_reserve(ivec *v, long need)
{
    if (need <= v->cap) return;
    long c = v->cap ? v->cap : 64;
    while (c < need) c *= 2;
    v->a = (int *)realloc(v->a, sizeof(int) * c);
    memset(v->a + v->cap, 0, sizeof(int) * (c - v->cap));
    v->cap = c;
}

struct ora_atree_s {
    int nat, napb, ierr, nbox, nlev, len, len3, len4, maxp, jatlen;
    double geom[3]; /* rscal, xonew, yonew */
    double *xs, *ys;
    int nl[128];
    int *ibox;
    double *xb, *yb;
    int *nbat, *npar, *nchi, *indb, *iat, *jat, *l1, *l2, *jl, *l3, *l4, *jl3, *jl4;
};

static void *zcalloc(long n, long sz) { return calloc((size_t)(n > 0 ? n : 1), (size_t)sz); }

ora_atree *ora_atree_build(int nat, const double *xa, const double *ya, int napb, int lmx, double eps7)
{
    ora_atree *t = (ora_atree *)calloc(1, sizeof(ora_atree));
    t->nat = nat;
    t->napb = napb;
    if (nat <= napb) { /* dapif1.f:711-717 */
        t->ierr = 16;
        return t;
    }
    if (lmx > 120) lmx = 120;

    /* ---- scaling, dapif1.f:745-770 ---- */
    double rminx = (double)1.0e30f, rmaxx = -(double)1.0e30f, rminy = (double)1.0e30f, rmaxy = -(double)1.0e30f;
    for (int i = 0; i < nat; ++i) {
        if (xa[i] > rmaxx) rmaxx = xa[i];
        if (ya[i] > rmaxy) rmaxy = ya[i];
        if (xa[i] < rminx) rminx = xa[i];
        if (ya[i] < rminy) rminy = ya[i];
    }
    const double rsx = rmaxx - rminx, rsy = rmaxy - rminy;
    const double xo = (rmaxx + rminx) / 2.0, yo = (rmaxy + rminy) / 2.0;
    const double rs = (rsx >= rsy) ? rsx : rsy;
    const double rscal = 128.0 / rs * (double)0.99f;
    t->geom[0] = rscal;
    t->geom[1] = xo;
    t->geom[2] = yo;
    double *xs = (double *)zcalloc(nat + 1, sizeof(double));
    double *ys = (double *)zcalloc(nat + 1, sizeof(double));
    for (int i = 1; i <= nat; ++i) {
        xs[i] = (xa[i - 1] - xo) * rscal;
        ys[i] = (ya[i - 1] - yo) * rscal;
    }

    /* ---- DAINIA, dapif1.f:1262-1300 ---- */
    long cap = 64;
    double *xb = (double *)zcalloc(cap, sizeof(double)), *yb = (double *)zcalloc(cap, sizeof(double));
    int *nbat = (int *)zcalloc(cap, sizeof(int)), *npar = (int *)zcalloc(cap, sizeof(int)),
        *nchi = (int *)zcalloc(cap, sizeof(int));
    int *ibox = (int *)zcalloc(nat + 1, sizeof(int));
    int *nl = t->nl;
    int nbox = 4, L = 1, maxc = nat;
    nl[1] = 1;
    nl[2] = 5;
    {
        const double d = 0.25 * 128.0;
        const double cx[4] = {-d, d, -d, d}, cy[4] = {d, d, -d, -d};
        for (int b = 1; b <= 4; ++b) {
            xb[b] = cx[b - 1];
            yb[b] = cy[b - 1];
        }
    }

    /* ---- DAASSI loop, dapif1.f:1218-1256 ---- */
    for (;;) {
        if (L > lmx) {
            t->ierr = 32;
            break;
        }
        const double ssb = ldexp(1.0, -(L - 1));
        if (ssb <= eps7) {
            t->ierr = 1;
            break;
        }
        if (maxc <= napb) break;

        /* DAIBNB, dapif1.f:1303-1370: counts of level L are cleared, points of boxes with
         * more than NAPB points move to the child quadrant of their centre comparison. */
        for (int b = nl[L]; b <= nl[L + 1]; ++b) nbat[b] = 0;
        for (int i = 1; i <= nat; ++i) {
            const int ib = ibox[i];
            double xp, yp;
            int nc;
            if (ib == 0) {
                xp = 0.0;
                yp = 0.0;
                nc = 1;
            } else {
                if (nbat[ib] <= napb) continue;
                xp = xb[ib];
                yp = yb[ib];
                nc = nchi[ib];
            }
            int q;
            if (xs[i] < xp) q = (ys[i] < yp) ? 2 : 0;
            else q = (ys[i] < yp) ? 3 : 1;
            ibox[i] = nc + q;
            nbat[nc + q] += 1;
        }

        /* DAFAMI, dapif1.f:1373-1453: children of every box of level L with > NAPB points,
         * numbered after the current last box in the order of their parents. */
        const double rl2 = ldexp(1.0, L);
        const double d = 0.25 / rl2 * 128.0;
        const int nll = nl[L], nbl = nbox;
        maxc = 0;
        for (int b = nll; b <= nbl; ++b) {
            const int inb = nbat[b];
            if (inb > maxc) maxc = inb;
            nchi[b] = 0;
            if (inb <= napb) continue;
            if (nbox + 6 > cap) {
                long nc2 = cap * 2;
                xb = (double *)realloc(xb, sizeof(double) * nc2);
                yb = (double *)realloc(yb, sizeof(double) * nc2);
                nbat = (int *)realloc(nbat, sizeof(int) * nc2);
                npar = (int *)realloc(npar, sizeof(int) * nc2);
                nchi = (int *)realloc(nchi, sizeof(int) * nc2);
                memset(xb + cap, 0, sizeof(double) * (nc2 - cap));
                memset(yb + cap, 0, sizeof(double) * (nc2 - cap));
                memset(nbat + cap, 0, sizeof(int) * (nc2 - cap));
                memset(npar + cap, 0, sizeof(int) * (nc2 - cap));
                memset(nchi + cap, 0, sizeof(int) * (nc2 - cap));
                cap = nc2;
            }
            nchi[b] = nbox + 1;
            const double ox[4] = {-d, d, -d, d}, oy[4] = {d, d, -d, -d};
            for (int j = 1; j <= 4; ++j) {
                const int jj = nbox + j;
                npar[jj] = b;
                xb[jj] = xb[b] + ox[j - 1];
                yb[jj] = yb[b] + oy[j - 1];
            }
            nbox += 4;
        }
        L += 1;
        nl[L + 1] = nbox + 1;
    }
    if (t->ierr > 1) {
        free(xs); free(ys); free(xb); free(yb); free(nbat); free(npar); free(nchi); free(ibox);
        return t;
    }
    if (t->ierr == 1) {
        /* The reference goes on with boxes whose counts were never formed; this restatement
         * stops here and reports the condition. */
        free(xs); free(ys); free(xb); free(yb); free(nbat); free(npar); free(nchi); free(ibox);
        return t;
    }

    /* ---- DACOMP, dapif1.f:1456-1539 ---- */
    int *shift = (int *)zcalloc(nbox + 2, sizeof(int));
    {
        int nz = 0;
        shift[0] = 0;
        for (int b = 1; b <= nbox; ++b) {
            shift[b] = nz;
            if (nbat[b] == 0) nz += 1;
        }
        shift[nbox + 1] = nz;
        for (int b = 1; b <= nbox; ++b) {
            if (nbat[b] <= 0) continue;
            const int nb = b - shift[b];
            xb[nb] = xb[b];
            yb[nb] = yb[b];
            nbat[nb] = nbat[b];
            npar[nb] = npar[b] - shift[npar[b]];
            nchi[nb] = nchi[b] - shift[nchi[b]];
        }
        for (int i = 1; i <= nat; ++i) ibox[i] -= shift[ibox[i]];
        for (int lev = 1; lev <= L + 1; ++lev) nl[lev] -= shift[nl[lev]];
        nbox -= nz;
        npar[nbox + 1] = 0;
    }
    free(shift);

    /* ---- DAINDB, dapif1.f:1542-1606 ---- */
    int *indb = (int *)zcalloc(nbox + 2, sizeof(int));
    int maxp = 0;
    for (int lev = 1; lev <= L - 1; ++lev) {
        for (int b = nl[lev]; b <= nl[lev + 1] - 1; ++b) {
            const int nb = nbat[b];
            if (lev == L - 1) {
                indb[b] = 1;
                nchi[b] = 0;
                if (nb > maxp) maxp = nb;
            } else if (nb <= napb) {
                indb[b] = 1;
                if (nb > maxp) maxp = nb;
            } else {
                indb[b] = 0;
            }
        }
    }

    /* ---- DAIJAT, dapif1.f:1609-1680 ---- */
    int *iat = (int *)zcalloc(nbox + 2, sizeof(int));
    int *jat = (int *)zcalloc(nbox + nat + 3, sizeof(int));
    int jatlen = 0;
    {
        int nacc = 1, nc = 0;
        for (int b = 1; b <= nbox; ++b) {
            const int nb = nbat[b];
            if (indb[b] == 0) {
                iat[b] = -nb;
            } else {
                iat[b] = nacc + nc;
                nc = nb + 1;
                nacc = iat[b];
                jat[nacc] = 0;
                jatlen = nacc + nb;
            }
        }
        for (int i = 1; i <= nat; ++i) {
            const int ia = iat[ibox[i]];
            jat[ia] += 1;
            jat[ia + jat[ia]] = i;
        }
    }

    /* ---- DALI12 + DAIN12, dapif2.f:3-210 ---- */
    int *l1 = (int *)zcalloc(nbox + 2, sizeof(int)), *l2 = (int *)zcalloc(nbox + 2, sizeof(int));
    ivec jl = {0, 0, 0};
    iv_reserve(&jl, 36L * (nbox + 1) + 2);
    {
        const int imin = nl[1], imax = nl[2] - 1;
        l2[1] = 1;
        l1[1] = 1;
        for (int i = imin; i <= imax; ++i) {
            int lc = l2[i];
            for (int j = imin; j <= imax; ++j) {
                if (i == j) continue;
                jl.a[lc++] = j;
            }
            l2[i + 1] = lc;
            l1[i + 1] = lc;
        }
        double d = 0.5 * 1.25 * 128.0;
        for (int lev = 2; lev <= L - 1; ++lev) {
            d = d / 2;
            for (int i = nl[lev]; i <= nl[lev + 1] - 1; ++i) {
                const int np = npar[i];
                const int lpmin = l2[np], lpmax = l1[np + 1] - 1;
                const double x = xb[i], y = yb[i];
                int inli[27], col[8], nli = 0, nco = 0;
                for (int jco = lpmin; jco <= lpmax; ++jco) {
                    const int ico = jl.a[jco];
                    int icok = nchi[ico];
                    if (icok == 0) continue;
                    while (npar[icok] == ico) {
                        if (fabs(x - xb[icok]) <= d && fabs(y - yb[icok]) <= d) col[nco++] = icok;
                        else inli[nli++] = icok;
                        icok += 1;
                    }
                }
                for (int nbro = nchi[np]; npar[nbro] == np; ++nbro)
                    if (nbro != i) col[nco++] = nbro;
                int lr = l1[i];
                for (int j = 0; j < nli; ++j) jl.a[lr + j] = inli[j];
                l2[i] = lr + nli;
                lr = l2[i];
                for (int j = 0; j < nco; ++j) jl.a[lr + j] = col[j];
                l1[i + 1] = lr + nco;
            }
        }
    }
    const int len = l1[nbox + 1] - 1;

    /* ---- DALI34 + DASCAN, dapif2.f:212-474 ---- */
    int *l3 = (int *)zcalloc(nbox + 2, sizeof(int)), *l4 = (int *)zcalloc(nbox + 2, sizeof(int));
    ivec jl3 = {0, 0, 0}, jl4 = {0, 0, 0}, idesc = {0, 0, 0};
    int len3 = 0, len4 = 0;
    l3[1] = 1;
    l4[1] = 1;
    for (int i = 1; i <= nbox; ++i) {
        const int l2i = l2[i], l1i = l1[i + 1] - 1;
        if (l1i < l2i || indb[i] == 0) {
            l3[i + 1] = l3[i];
            l4[i + 1] = l4[i];
            continue;
        }
        int ista = 1, iend = 0, nnbar = 0;
        iv_reserve(&idesc, (l1i - l2i + 1) + 2);
        for (int j = l2i; j <= l1i; ++j) {
            idesc.a[++iend] = jl.a[j];
            if (indb[jl.a[j]] == 0) nnbar = 1;
        }
        if (!nnbar) {
            l3[i + 1] = l3[i];
            l4[i + 1] = l4[i];
            continue;
        }
        const double x = xb[i], y = yb[i];
        const double xc = xb[jl.a[l2i]], yc = yb[jl.a[l2i]];
        const double xd0 = fabs(x - xc), yd0 = fabs(y - yc);
        const double d1 = (xd0 >= yd0) ? xd0 : yd0;
        int iloop = 0;
        while (iend >= ista) {
            iloop += 1;
            /* DASCAN, dapif2.f:416-474; the factor is REAL*4 arithmetic in the reference */
            const float fac = 0.5f + 1.25f * ldexpf(1.0f, -iloop);
            const double d = d1 * (double)fac;
            int iendn = iend;
            const int istan = iend + 1;
            for (int k = ista; k <= iend; ++k) {
                const int ibo = idesc.a[k];
                const double xd = fabs(x - xb[ibo]), yd = fabs(y - yb[ibo]);
                if (xd > d || yd > d) {
                    iv_reserve(&jl4, len4 + 2);
                    jl4.a[++len4] = ibo;
                } else if (indb[ibo] == 1) {
                    iv_reserve(&jl3, len3 + 2);
                    jl3.a[++len3] = ibo;
                } else {
                    for (int ikid = nchi[ibo]; npar[ikid] == ibo; ++ikid) {
                        iv_reserve(&idesc, iendn + 2);
                        idesc.a[++iendn] = ikid;
                    }
                }
            }
            ista = istan;
            iend = iendn;
        }
        l3[i + 1] = len3 + 1;
        l4[i + 1] = len4 + 1;
    }
    free(idesc.a);

    t->xs = xs;
    t->ys = ys;
    t->ibox = ibox;
    t->xb = xb;
    t->yb = yb;
    t->nbat = nbat;
    t->npar = npar;
    t->nchi = nchi;
    t->indb = indb;
    t->iat = iat;
    t->jat = jat;
    t->jatlen = jatlen;
    t->l1 = l1;
    t->l2 = l2;
    t->jl = jl.a;
    t->l3 = l3;
    t->l4 = l4;
    t->jl3 = jl3.a;
    t->jl4 = jl4.a;
    t->nbox = nbox;
    t->nlev = L - 1;
    t->len = len;
    t->len3 = len3;
    t->len4 = len4;
    t->maxp = maxp;
    return t;
}

void ora_atree_sizes(const ora_atree *t, int *out)
{
    out[0] = t->ierr;
    out[1] = t->nbox;
    out[2] = t->nlev;
    out[3] = t->len;
    out[4] = t->len3;
    out[5] = t->len4;
    out[6] = t->maxp;
    out[7] = t->jatlen;
    out[8] = t->nat;
    out[9] = t->napb;
}

static void cpy_i(int *dst, const int *src1, long n)
{
    if (dst && src1 && n > 0) memcpy(dst, src1 + 1, sizeof(int) * (size_t)n);
}
static void cpy_d(double *dst, const double *src1, long n)
{
    if (dst && src1 && n > 0) memcpy(dst, src1 + 1, sizeof(double) * (size_t)n);
}

void ora_atree_copy(const ora_atree *t, double *xs, double *ys, double *geom, int *nl, int *ibox, double *xb,
                    double *yb, int *nbat, int *npar, int *nchi, int *indb, int *iat, int *jat, int *l1, int *l2,
                    int *jl, int *l3, int *l4, int *jl3, int *jl4)
{
    if (t->ierr > 1 || t->ierr == 1) return;
    const long nb = t->nbox;
    if (geom) memcpy(geom, t->geom, sizeof(t->geom));
    cpy_d(xs, t->xs, t->nat);
    cpy_d(ys, t->ys, t->nat);
    cpy_i(nl, t->nl, t->nlev + 2);
    cpy_i(ibox, t->ibox, t->nat);
    cpy_d(xb, t->xb, nb);
    cpy_d(yb, t->yb, nb);
    cpy_i(nbat, t->nbat, nb);
    cpy_i(npar, t->npar, nb);
    cpy_i(nchi, t->nchi, nb);
    cpy_i(indb, t->indb, nb);
    cpy_i(iat, t->iat, nb);
    cpy_i(jat, t->jat, t->jatlen);
    cpy_i(l1, t->l1, nb + 1);
    cpy_i(l2, t->l2, nb + 1);
    cpy_i(jl, t->jl, t->len);
    cpy_i(l3, t->l3, nb + 1);
    cpy_i(l4, t->l4, nb + 1);
    cpy_i(jl3, t->jl3, t->len3);
    cpy_i(jl4, t->jl4, t->len4);
}

void ora_atree_free(ora_atree *t)
{
    if (!t) return;
    void *p[] = {t->xs, t->ys, t->ibox, t->xb, t->yb, t->nbat, t->npar, t->nchi, t->indb, t->iat, t->jat,
                 t->l1, t->l2, t->jl, t->l3, t->l4, t->jl3, t->jl4};
    for (unsigned i = 0; i < sizeof(p) / sizeof(p[0]); ++i) free(p[i]);
    free(t);
}
