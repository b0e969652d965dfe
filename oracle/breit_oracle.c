/*
 * breit_oracle.c -- CPU restatement of rci90's transverse Breit radial
 * integrals and their place in the SETHAM element loop.
 *
 * TEST INFRASTRUCTURE ONLY (part of liboracle.so, the checker of the HIP
 * engine; the product path never links or calls it).
 *
 * Routines restated (rci90/): BESSEL (bessel.f90:46-241) with its KEEP
 * cache and WIJ frequencies, ZKF (zkf.f90:33-83), RKINT (rkint.f90:43-81),
 * SKINT (skint.f90:44-109), BRRA (brra.f90:41-124), BRINTF (brintf.f90:
 * 44-96), TRIANGBREIT1/2, GENINTBREIT1/2 (genintbreit1.f90:45-104),
 * BRINT1/2 (table lookups, brint1.f90:27-67) and BRINT3..6 (memoised
 * BRINTF, brint3.f90:30-107), and the Breit part of SETHAM
 * (setham_gg.f90:298-357).  All of it is state-dependent the way the
 * reference is: BESSEL's cached Bessel functions survive from call to call,
 * and BRINT3..6 keep the first value they computed per key.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "graspci_oracle.h"

#define F1(a, i) ((a)[(i) - 1])
#define KEYORB 215

struct orc_breit {
    int n;
    const double *r, *rp, *rpor;
    double ncc[4], cnc5c[12], accy;
    int nw, ld;
    const double *P, *Q, *e;
    const int *mf, *nak, *nkl, *nkj;
    double c, wfact;
    /* bess_C / stor_C: KEEP(IK,IW), WIJ(IW), BESSJ/BESSN(IK,IW,1:N) */
    int keep[2][2];
    double wij[2];
    double *bessj[2][2], *bessn[2][2];
    /* tatb_C */
    double *ta, *tb;
    int mtp;
    /* GENINTBREIT1/2 tables (bilst_C, kkstartbreit_C) */
    int ntab[2];
    int kstart[2][22];
    int32_t *tkey[2];
    double *tval[2];
    /* BRINT3..6 memo: sorted keys and values per type */
    int nmemo[4], capmemo[4];
    int64_t *mkey[4];
    double *mval[4];
};

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* Fortran REAL**INTEGER: repeated squaring, a reciprocal for a negative
 * exponent */
static double ipow_b(double a, int k)
{
    const int recip = k < 0;
    double res = 1.0;
    for (;;) {
        if (k & 1) res *= a;
        k /= 2;
        if (!k) break;
        a *= a;
    }
    return recip ? 1.0 / res : res;
}

orc_breit *orc_breit_new(int n, const double *r, const double *rp, const double *rpor, const double *ncc,
                         const double *cnc5c, double accy, int nw, int ld, const double *P, const double *Q,
                         const int *mf, const int *nak, const int *nkl, const int *nkj, const double *e,
                         double c, double wfact)
{
    orc_breit *s = calloc(1, sizeof(orc_breit));
    s->n = n;
    s->r = r;
    s->rp = rp;
    s->rpor = rpor;
    memcpy(s->ncc, ncc, sizeof(s->ncc));
    memcpy(s->cnc5c, cnc5c, sizeof(s->cnc5c));
    s->accy = accy;
    s->nw = nw;
    s->ld = ld;
    s->P = P;
    s->Q = Q;
    s->mf = mf;
    s->nak = nak;
    s->nkl = nkl;
    s->nkj = nkj;
    s->e = e;
    s->c = c;
    s->wfact = wfact;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            s->bessj[a][b] = calloc(n + 16, sizeof(double));
            s->bessn[a][b] = calloc(n + 16, sizeof(double));
        }
    s->ta = calloc(n + 16, sizeof(double));
    s->tb = calloc(n + 16, sizeof(double));
    return s;
}

void orc_breit_free(orc_breit *s)
{
    if (!s) return;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            free(s->bessj[a][b]);
            free(s->bessn[a][b]);
        }
    free(s->ta);
    free(s->tb);
    for (int t = 0; t < 2; ++t) {
        free(s->tkey[t]);
        free(s->tval[t]);
    }
    for (int t = 0; t < 4; ++t) {
        free(s->mkey[t]);
        free(s->mval[t]);
    }
    free(s);
}

#define PF(i, j) ((s->P)[(size_t)((j) - 1) * s->ld + (i) - 1])
#define QF(i, j) ((s->Q)[(size_t)((j) - 1) * s->ld + (i) - 1])
#define BJ(ik, iw) (s->bessj[(ik) - 1][(iw) - 1])
#define BN(ik, iw) (s->bessn[(ik) - 1][(iw) - 1])

/* bessel.f90:46-241 */
static void bessel(orc_breit *s, int ia, int ib, int ik, int iw, int k)
{
    const int N = s->n;
    const double epsi = sqrt(0.1 * s->accy);
    const int icode = imax(ia, ib) + KEYORB * (imin(ia, ib) + KEYORB * k);
    if (icode == s->keep[ik - 1][iw - 1]) return;
    const double w = s->wfact * fabs(F1(s->e, ia) - F1(s->e, ib)) / s->c;
    s->wij[iw - 1] = w;
    for (int iwkp = 1; iwkp <= 2; ++iwkp)
        for (int ikkp = 1; ikkp <= 2; ++ikkp) {
            if (s->keep[ikkp - 1][iwkp - 1] != icode) continue;
            s->keep[ik - 1][iw - 1] = icode;
            memcpy(BJ(ik, iw), BJ(ikkp, iwkp), sizeof(double) * (size_t)(N + 1));
            memcpy(BN(ik, iw), BN(ikkp, iwkp), sizeof(double) * (size_t)(N + 1));
            return;
        }
    s->keep[ik - 1][iw - 1] = icode;
    double *bj = BJ(ik, iw), *bn = BN(ik, iw);
    if (w < epsi * epsi) {
        for (int j = 1; j <= N; ++j) {
            F1(bj, j) = 0.0;
            F1(bn, j) = 0.0;
        }
        return;
    }
    const int nn = k;
    F1(bj, 1) = 0.0;
    F1(bn, 1) = 0.0;
    int jchan = -1;
    for (int j = 2; j <= N; ++j) {
        const double rw = F1(s->r, j) * w;
        const double wa = -0.5 * (rw * rw);
        double xb1 = 1.0, xb2 = 1.0, s1 = 0.0, s2 = 0.0;
        int done = 0;
        for (int i = 1; i <= 4; ++i) {
            xb1 = xb1 * wa / (double)(i * (2 * (nn + i) + 1));
            xb2 = xb2 * wa / (double)(i * (2 * (i - nn) - 1));
            s1 = s1 + xb1;
            s2 = s2 + xb2;
            if (fabs(xb1) >= fabs(s1) * epsi || fabs(xb2) >= fabs(s2) * epsi) continue;
            F1(bj, j) = s1;
            F1(bn, j) = s2;
            done = 1;
            break;
        }
        if (!done) {
            jchan = j;
            break;
        }
    }
    if (jchan < 0) return;
    double dfnm, dfn;
    if (nn == 0) {
        dfnm = 1.0;
        dfn = 1.0;
    } else {
        dfnm = 1.0;
        for (int i = 3; i <= 2 * nn - 1; i += 2) dfnm = dfnm * (double)i;
        dfn = dfnm * (double)(2 * nn + 1);
    }
    dfnm = 1.0 / dfnm;
    double ssn, scn;
    int iswap;
    switch (nn % 4) {
        case 1: ssn = -1.0; scn = 1.0; iswap = 1; break;
        case 2: ssn = -1.0; scn = -1.0; iswap = 0; break;
        case 3: ssn = 1.0; scn = -1.0; iswap = 1; break;
        default: ssn = 1.0; scn = 1.0; iswap = 0; break;
    }
    for (int j = jchan; j <= N; ++j) {
        const double wa = w * F1(s->r, j);
        double sn, cn;
        if (iswap == 0) {
            sn = ssn * sin(wa);
            cn = scn * cos(wa);
        } else {
            sn = ssn * cos(wa);
            cn = scn * sin(wa);
        }
        const double obwa = 1.0 / wa;
        double b = obwa;
        double s1 = b * sn, s2 = b * cn;
        for (int i = 1; i <= nn; ++i) {
            const double skeep = sn;
            sn = cn;
            cn = -skeep;
            b = b * obwa * (double)((nn + i) * (nn - i + 1)) / (double)(2 * i);
            s1 = s1 + b * sn;
            s2 = s2 + b * cn;
        }
        s1 = s1 * dfn / ipow_b(wa, nn) - 1.0;
        s2 = s2 * ipow_b(wa, nn + 1) * dfnm - 1.0;
        F1(bj, j) = s1;
        F1(bn, j) = s2;
    }
}

/* zkf.f90:33-83: ZK (TB) from TA(1:MTP) */
static void zkf(orc_breit *s, int k)
{
    const int N = s->n, mtp = s->mtp;
    const int mtpp1 = mtp + 1, mtpp3 = mtp + 3, mtpp4 = mtp + 4;
    double *ta = s->ta, *zk = s->tb;
    double *temp = calloc(N + 16, sizeof(double));
    double *rttk = calloc(N + 16, sizeof(double));
    const double C1 = s->ncc[0], C2 = s->ncc[1], C3 = s->ncc[2];
    /* zkf.f90:46-48 tabulates R**K only for K > 0 and reads RTTK for every
     * K /= 0; SKINT's ZKF(K - 1) with K = 0 (GENINTBREIT2's k = 0 entries)
     * therefore reads an unset array in the reference.  Here RTTK = R**K
     * for K < 0 as well. */
    if (k != 0)
        for (int i = 2; i <= N; ++i) F1(rttk, i) = ipow_b(F1(s->r, i), k);
    F1(temp, 1) = 0.0;
    for (int i = 2; i <= mtp; ++i) {
        const double rhop = F1(s->rp, i) * F1(ta, i);
        F1(temp, i) = (k == 0) ? rhop : F1(rttk, i) * rhop;
    }
    for (int i = mtpp1; i <= mtpp4; ++i) F1(temp, i) = 0.0;
    F1(zk, 1) = 0.0;
    for (int ii = 2; ii <= 4; ++ii) {
        double sum = 0.0;
        for (int kk = 2; kk <= 5; ++kk) sum = sum + s->cnc5c[(ii - 2) * 4 + (kk - 2)] * F1(temp, kk);
        F1(zk, ii) = sum;
    }
    for (int ii = 5; ii <= mtpp3; ++ii)
        F1(zk, ii) = F1(zk, ii - 4) + C1 * (F1(temp, ii - 4) + F1(temp, ii)) +
                     C2 * (F1(temp, ii - 3) + F1(temp, ii - 1)) + C3 * F1(temp, ii - 2);
    const double zklim = F1(zk, mtpp3);
    if (k == 0) {
        for (int i = mtpp4; i <= N; ++i) F1(zk, i) = zklim;
    } else {
        for (int i = 2; i <= mtpp3; ++i) F1(zk, i) = F1(zk, i) / F1(rttk, i);
        for (int i = mtpp4; i <= N; ++i) F1(zk, i) = zklim / F1(rttk, i);
    }
    free(temp);
    free(rttk);
}

static double quad_ta(orc_breit *s) { return orc_quad(s->r, s->rp, s->ncc, s->ta, s->mtp); }

/* rkint.f90:43-81 */
static double rkint(orc_breit *s, const double *rac, int ia, int ic, const double *rbd, int ib, int id, int k, int iw)
{
    const int mxrbd = imin(F1(s->mf, ib), F1(s->mf, id));
    const int mxrac = imin(F1(s->mf, ia), F1(s->mf, ic));
    double *ta = s->ta, *tb = s->tb;
    if (iw == 0) {
        for (int i = 1; i <= mxrac; ++i) F1(ta, i) = F1(rac, i);
        s->mtp = mxrac;
        zkf(s, k);
        s->mtp = imin(s->mtp, mxrbd);
        F1(ta, 1) = 0.0;
        for (int i = 2; i <= s->mtp; ++i) F1(ta, i) = F1(rbd, i) * F1(tb, i) * F1(s->rpor, i);
    } else {
        const double *bj = BJ(1, iw), *bn = BN(1, iw);
        for (int i = 1; i <= mxrac; ++i) F1(ta, i) = F1(rac, i) * (1.0 + F1(bj, i));
        s->mtp = mxrac;
        zkf(s, k);
        s->mtp = imin(s->mtp, mxrbd);
        F1(ta, 1) = 0.0;
        for (int i = 2; i <= s->mtp; ++i)
            F1(ta, i) = F1(rbd, i) * (1.0 + F1(bn, i)) * F1(tb, i) * F1(s->rpor, i);
    }
    return quad_ta(s);
}

/* skint.f90:44-109 */
static double skint(orc_breit *s, const double *rac, int ia, int ic, const double *rbd, int ib, int id, int k, int iw)
{
    const int N = s->n;
    const double epsi = 1.0e-10;
    const double w = s->wij[iw - 1];
    const double wk = (double)(k + k + 1);
    const int mxrac = imin(F1(s->mf, ia), F1(s->mf, ic));
    const int mxrbd = imin(F1(s->mf, ib), F1(s->mf, id));
    double *ta = s->ta, *tb = s->tb;
    double *tkeep = calloc(N + 16, sizeof(double));
    double res;
    for (int i = 1; i <= mxrac; ++i) F1(ta, i) = F1(rac, i);
    s->mtp = mxrac;
    zkf(s, k - 1);
    if (fabs(w) < epsi) {
        for (int i = 1; i <= s->mtp; ++i) F1(tkeep, i) = F1(tb, i);
        zkf(s, k + 1);
        s->mtp = imin(s->mtp, mxrbd);
        F1(ta, 1) = 0.0;
        for (int i = 2; i <= s->mtp; ++i) F1(ta, i) = F1(rbd, i) * F1(s->rpor, i) * (F1(tb, i) - F1(tkeep, i));
        const double valu = quad_ta(s);
        res = wk * valu * 0.5;
    } else {
        const double *bj1 = BJ(1, iw), *bn2 = BN(2, iw), *bj2 = BJ(2, iw), *bn1 = BN(1, iw);
        for (int i = 1; i <= s->mtp; ++i) F1(tkeep, i) = F1(tb, i);
        for (int i = 1; i <= s->mtp; ++i) F1(ta, i) = -F1(ta, i) * F1(bj1, i);
        zkf(s, k - 1);
        s->mtp = imin(s->mtp, mxrbd);
        F1(ta, 1) = 0.0;
        for (int i = 2; i <= s->mtp; ++i) {
            const double r = F1(s->r, i);
            F1(ta, i) = ((1.0 + F1(bn2, i)) * F1(tb, i) - F1(tkeep, i) * F1(bn2, i)) * F1(rbd, i) / (r * r) *
                        F1(s->rpor, i);
        }
        double valu = quad_ta(s);
        res = (wk / w) * (wk / w) * valu;
        for (int i = 1; i <= mxrbd; ++i) F1(ta, i) = F1(rbd, i) * (1.0 + F1(bj2, i));
        s->mtp = mxrbd;
        zkf(s, k + 1);
        s->mtp = imin(s->mtp, mxrac);
        F1(ta, 1) = 0.0;
        for (int i = 2; i <= s->mtp; ++i)
            F1(ta, i) = F1(rac, i) * (1.0 + F1(bn1, i)) * F1(tb, i) * F1(s->r, i) * F1(s->rp, i);
        valu = quad_ta(s);
        res = res - valu * w * w / (double)((2 * k + 3) * (2 * k - 1));
    }
    free(tkeep);
    return res;
}

/* brra.f90:41-124; arguments in BRRA's own order (IA, IC, IB, ID) */
static double brra(orc_breit *s, int itype, int ia, int ic, int ib, int id, int k)
{
    const int N = s->n;
    double *rac = calloc(N + 16, sizeof(double)), *rbd = calloc(N + 16, sizeof(double));
    const int mxrac = imin(F1(s->mf, ia), F1(s->mf, ic));
    for (int i = 1; i <= mxrac; ++i) F1(rac, i) = PF(i, ia) * QF(i, ic);
    const int mxrbd = imin(F1(s->mf, ib), F1(s->mf, id));
    for (int i = 1; i <= mxrbd; ++i) F1(rbd, i) = PF(i, ib) * QF(i, id);
    double v;
    switch (itype) {
        case 1:
            if (ia == ib && ic == id) goto l9;
            if (ia == id && ic == ib) goto l10;
            v = (rkint(s, rac, ia, ic, rbd, ib, id, k, 1) + rkint(s, rac, ia, ic, rbd, ib, id, k, 2) +
                 rkint(s, rbd, ib, id, rac, ia, ic, k, 1) + rkint(s, rbd, ib, id, rac, ia, ic, k, 2)) *
                0.5;
            goto out;
        case 2:
            if (ia == ib && ic == id) goto l26;
            if (ia == id && ic == ib) goto l26;
            v = (skint(s, rac, ia, ic, rbd, ib, id, k, 1) + skint(s, rac, ia, ic, rbd, ib, id, k, 2)) * 0.5;
            goto out;
        case 3:
            if (ia == ic) {
                for (int i = 1; i <= mxrbd; ++i) F1(rbd, i) = F1(rbd, i) + PF(i, id) * QF(i, ib);
                v = (rkint(s, rac, ia, ic, rbd, ib, id, k, 0) + rkint(s, rbd, ib, id, rac, ia, ic, k, 0) +
                     rkint(s, rac, ia, ic, rbd, ib, id, k, 2) + rkint(s, rbd, ib, id, rac, ia, ic, k, 2)) *
                    0.5;
                goto out;
            }
            for (int i = 1; i <= mxrac; ++i) F1(rac, i) = F1(rac, i) + PF(i, ic) * QF(i, ia);
            v = (rkint(s, rac, ia, ic, rbd, ib, id, k, 1) + rkint(s, rbd, ib, id, rac, ia, ic, k, 1) +
                 rkint(s, rac, ia, ic, rbd, ib, id, k, 0) + rkint(s, rbd, ib, id, rac, ia, ic, k, 0)) *
                0.5;
            goto out;
        case 4:
            v = rkint(s, rac, ia, ic, rbd, ib, id, k, 0) + rkint(s, rbd, ib, id, rac, ia, ic, k, 0);
            goto out;
        case 5:
            if (ia == id && ic == ib) goto l10;
            goto l9;
        default:
            goto l26;
    }
l9:
    v = 2.0 * rkint(s, rac, ia, ic, rbd, ib, id, k, 1);
    goto out;
l10:
    v = rkint(s, rac, ia, ic, rbd, ib, id, k, 1) + rkint(s, rbd, ib, id, rac, ia, ic, k, 1);
    goto out;
l26:
    v = skint(s, rac, ia, ic, rbd, ib, id, k, 1);
out:
    free(rac);
    free(rbd);
    return v;
}

/* brintf.f90:44-96 */
double orc_brintf(orc_breit *s, int itype, int ia, int ib, int ic, int id, int k)
{
    const int KEY2 = KEYORB * KEYORB;
    switch (itype) {
        case 1:
            if (ia != ic || ib != id || ia != id || ic == ib) bessel(s, ic, id, 1, 2, k);
            bessel(s, ia, ib, 1, 1, k);
            break;
        case 5:
            bessel(s, ia, ib, 1, 1, k);
            break;
        case 3:
            if (ia != ib) bessel(s, ia, ib, 1, 1, k);
            if (ic != id) bessel(s, ic, id, 1, 2, k);
            break;
        case 2:
            if (ia != ic || ib != id || ia != id || ic != ib) {
                int icod = imax(ic, id) + KEYORB * imin(ic, id);
                const int icod1 = icod + KEY2 * (k - 1), icod2 = icod + KEY2 * (k + 1);
                icod = imax(ia, ib) + KEYORB * imin(ia, ib);
                const int icod3 = icod + KEY2 * (k - 1), icod4 = icod + KEY2 * (k + 1);
                if (icod1 == s->keep[0][1] && icod2 == s->keep[1][1] && icod3 == s->keep[0][0] &&
                    icod4 == s->keep[1][0])
                    break;
                if (icod1 == s->keep[0][0] && icod2 == s->keep[1][0] && icod3 == s->keep[0][1] &&
                    icod4 == s->keep[1][1])
                    break;
                bessel(s, ic, id, 1, 2, k - 1);
                bessel(s, ic, id, 2, 2, k + 1);
            }
            bessel(s, ia, ib, 1, 1, k - 1);
            bessel(s, ia, ib, 2, 1, k + 1);
            break;
        case 6:
            bessel(s, ia, ib, 1, 1, k - 1);
            bessel(s, ia, ib, 2, 1, k + 1);
            break;
        default:
            break;
    }
    return brra(s, itype, ia, ib, ic, id, k);
}

/* triangbreit1.f90 / triangbreit2.f90 */
int orc_triangbreit(int type, int la, int ja, int lb, int jb, int lc, int jc, int ld, int jd, int k)
{
    if (type == 1) {
        for (int l = k - 1; l <= k + 1; ++l)
            if ((k + la + lb) % 2 == 1 && abs(ja - jb) <= 2 * l && abs(ja + jb) >= 2 * l &&
                (k + lc + ld) % 2 == 1 && abs(jc - jd) <= 2 * l && abs(jc + jd) >= 2 * l)
                return 1;
        return 0;
    }
    const int l = k;
    return (l + la + lb) % 2 == 0 && abs(ja - jb) <= 2 * l && abs(ja + jb) >= 2 * l && (l + lc + ld) % 2 == 0 &&
           abs(jc - jd) <= 2 * l && abs(jc + jd) >= 2 * l;
}

/* genintbreit1.f90:45-104 (type 1) and genintbreit2.f90 (type 2): the
 * count sweep, then the fill sweep with BRINTF in enumeration order.  The
 * table is kept in the state for BRINT1/2; kstart gets j2max + 2 entries. */
int64_t orc_genintbreit(orc_breit *s, int type, int *j2max_out, int *kstart, int32_t *keys, double *vals)
{
    const int nw = s->nw, key = nw + 1;
    int j2max = F1(s->nkj, 1);
    for (int i = 2; i <= nw; ++i)
        if (F1(s->nkj, i) > j2max) j2max = F1(s->nkj, i);
    if (j2max > 20) return -1;
    if (j2max_out) *j2max_out = j2max;
    int64_t nb = 0;
    int ks[22];
    ks[0] = 1;
    for (int gen = 0; gen < 2; ++gen) {
        nb = 0;
        for (int k = 0; k <= j2max; ++k) {
            for (int ia = 1; ia <= nw; ++ia)
                for (int ib = 1; ib <= nw; ++ib)
                    for (int ic = 1; ic <= nw; ++ic)
                        for (int id = 1; id <= nw; ++id) {
                            if (!orc_triangbreit(type, F1(s->nkl, ia), F1(s->nkj, ia), F1(s->nkl, ib),
                                                 F1(s->nkj, ib), F1(s->nkl, ic), F1(s->nkj, ic), F1(s->nkl, id),
                                                 F1(s->nkj, id), k))
                                continue;
                            ++nb;
                            if (gen) {
                                s->tkey[type - 1][nb - 1] = ((ia * key + ib) * key + ic) * key + id;
                                s->tval[type - 1][nb - 1] = orc_brintf(s, type, ia, ib, ic, id, k);
                            }
                        }
            ks[k + 1] = (int)nb + 1;
        }
        if (!gen) {
            free(s->tkey[type - 1]);
            free(s->tval[type - 1]);
            s->tkey[type - 1] = calloc(nb ? nb : 1, sizeof(int32_t));
            s->tval[type - 1] = calloc(nb ? nb : 1, sizeof(double));
        }
    }
    s->ntab[type - 1] = (int)nb;
    memcpy(s->kstart[type - 1], ks, sizeof(int) * (size_t)(j2max + 2));
    if (kstart) memcpy(kstart, ks, sizeof(int) * (size_t)(j2max + 2));
    if (keys) memcpy(keys, s->tkey[type - 1], sizeof(int32_t) * (size_t)nb);
    if (vals) memcpy(vals, s->tval[type - 1], sizeof(double) * (size_t)nb);
    return nb;
}

/* the table of GENINTBREIT type 1 or 2 held in the state (count returned) */
int64_t orc_breit_table(const orc_breit *s, int type, int32_t *keys, double *vals)
{
    const int nb = s->ntab[type - 1];
    if (keys) memcpy(keys, s->tkey[type - 1], sizeof(int32_t) * (size_t)nb);
    if (vals) memcpy(vals, s->tval[type - 1], sizeof(double) * (size_t)nb);
    return nb;
}

/* brint1.f90:27-67 (type 1, 2: table lookup) and brint3.f90:30-107 (types
 * 3..6: BRINTF on first use, then the stored value).  Returns 0, or -1 when
 * a table integral is missing (the reference STOPs). */
int orc_brint(orc_breit *s, int type, int ia, int ib, int ic, int id, int k, double *out)
{
    const int key = s->nw + 1;
    if (type <= 2) {
        const int t = type - 1;
        const int index = ((ia * key + ib) * key + ic) * key + id;
        int jl = s->kstart[t][k], ju = s->kstart[t][k + 1] - 1;
        if (ju < jl || index < s->tkey[t][jl - 1] || index > s->tkey[t][ju - 1]) return -1;
        while (ju - jl > 1) {
            const int jm = (ju + jl) / 2;
            if (s->tkey[t][jm - 1] > index) ju = jm;
            else jl = jm;
        }
        int loc;
        if (index == s->tkey[t][ju - 1]) loc = ju;
        else if (index == s->tkey[t][jl - 1]) loc = jl;
        else return -1;
        *out = s->tval[t][loc - 1];
        return 0;
    }
    const int t = type - 3;
    const int64_t index = ((((int64_t)k * key + ia) * key + ib) * key + ic) * key + id;
    int lo = 0, hi = s->nmemo[t];
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (s->mkey[t][mid] < index) lo = mid + 1;
        else hi = mid;
    }
    if (lo < s->nmemo[t] && s->mkey[t][lo] == index) {
        *out = s->mval[t][lo];
        return 0;
    }
    const double v = orc_brintf(s, type, ia, ib, ic, id, k);
    if (s->nmemo[t] == s->capmemo[t]) {
        s->capmemo[t] = s->capmemo[t] ? 2 * s->capmemo[t] : 1024;
        s->mkey[t] = realloc(s->mkey[t], sizeof(int64_t) * (size_t)s->capmemo[t]);
        s->mval[t] = realloc(s->mval[t], sizeof(double) * (size_t)s->capmemo[t]);
    }
    memmove(s->mkey[t] + lo + 1, s->mkey[t] + lo, sizeof(int64_t) * (size_t)(s->nmemo[t] - lo));
    memmove(s->mval[t] + lo + 1, s->mval[t] + lo, sizeof(double) * (size_t)(s->nmemo[t] - lo));
    s->mkey[t][lo] = index;
    s->mval[t][lo] = v;
    ++s->nmemo[t];
    *out = v;
    return 0;
}

/* The Breit part of SETHAM's element loop (setham_gg.f90:298-357) for one
 * pair: the entries [beg, end) of LABEL (6 per entry: a, b, c, d, k, type;
 * type < 0 for core terms) and COEFF, in list order.  Terms with
 * |COEFF| <= 1e-12 are skipped; core terms go to *elsto, the others to
 * *elemnt.  Returns 0, or -1 for a missing table integral. */
int orc_setham_breit_pair(orc_breit *s, int64_t beg, int64_t end, const int32_t *label, const double *coeff,
                          double *elemnt, double *elsto)
{
    for (int64_t i = beg; i < end; ++i) {
        if (!(fabs(coeff[i]) > 1.0e-12)) continue;
        const int32_t *lb = label + 6 * i;
        const int itype = abs(lb[5]);
        double tegral;
        if (orc_brint(s, itype, lb[0], lb[1], lb[2], lb[3], lb[4], &tegral)) return -1;
        const double contr = coeff[i] * tegral;
        if (lb[5] > 0) *elemnt = *elemnt + contr;
        else *elsto = *elsto + contr;
    }
    return 0;
}
