/*
 * graspci_oracle.c -- CPU restatement of GRASP2018's rci90 hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product library never does.  Built with -ffp-contract=off so that every
 * multiply-add rounds twice, as the reference's Fortran does on x86-64.
 *
 * Each routine cites the reference file:line it restates (paths relative to
 * the GRASP source tree src/).  Fortran 1-based indexing is kept inside the
 * routines through the F1() accessor to make the correspondence auditable.
 */
#include "graspci_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define F1(a, i) ((a)[(i) - 1])

/* ------------------------------------------------------------------ grid */

/* lib9290/radgrd.f90:35-60 (HP == 0 branch). */
void orc_radgrd(int n, double rnt, double h, double *r, double *rp, double *rpor)
{
    int np10 = n + 10;
    double eph = exp(h), ett = 1.0;
    F1(r, 1) = 0.0;
    F1(rpor, 1) = 0.0;
    F1(rp, 1) = rnt;
    for (int i = 2; i <= np10; ++i) {
        ett = eph * ett;
        double ettm1 = ett - 1.0;
        F1(r, i) = rnt * ettm1;
        F1(rp, i) = rnt * ett;
        F1(rpor, i) = ett / ettm1;
    }
}

/* Integer numerators of the 5-point Newton-Cotes weights over 720 used by
 * SETQIC (setqic.f90:249-258).  Derived here from first principles: weight
 * of node i (0..4, unit spacing) for the integral from node 0 to node t is
 * the integral of the Lagrange basis polynomial L_i over [0, t]; times 720
 * this is an integer.  t = 4 gives the closed rule (Boole), t = 1..3 the
 * starting formulas. */
static void nc5_numerators(int t, double num[5])
{
    for (int i = 0; i < 5; ++i) {
        /* polynomial coefficients of prod_{m != i} (x - m) */
        double c[5] = {1, 0, 0, 0, 0};
        int deg = 0;
        double den = 1.0;
        for (int m = 0; m < 5; ++m) {
            if (m == i) continue;
            /* multiply by (x - m) */
            for (int d = deg + 1; d >= 1; --d) c[d] = c[d - 1] - m * c[d];
            c[0] = -m * c[0];
            ++deg;
            den *= (double)(i - m);
        }
        double integ = 0.0;
        for (int d = 0; d <= 4; ++d) integ += c[d] * pow((double)t, d + 1) / (d + 1);
        num[i] = (double)llround(720.0 * integ / den);
    }
}

/* lib9290/setqic.f90:312-322: CNC5C(i,j) = H/720 * C5NUM(i,j), C1..C4. */
void orc_setqic(double h, double *ncc, double *cnc5c)
{
    double factor = h / 720.0;
    for (int j = 2; j <= 4; ++j) {
        double num[5];
        nc5_numerators(j - 1, num);
        for (int i = 2; i <= 5; ++i) cnc5c[(j - 2) * 4 + (i - 2)] = factor * num[i - 1];
    }
    double closed[5];
    nc5_numerators(4, closed);
    ncc[0] = factor * closed[0];  /* C1 = 224 H/720 */
    ncc[1] = factor * closed[1];  /* C2 = 1024 H/720 */
    ncc[2] = factor * closed[2];  /* C3 = 384 H/720 */
    ncc[3] = ncc[0] + ncc[0];     /* C4 = C1 + C1 */
}

/* 13-point Lagrange first-derivative table of SETQIC (setqic.f90:179-234,
 * 291-296): A(i,j) = B13(i,j) * (1/12!), B13(i,j) = 12! L_j'(x_i) on unit
 * spaced nodes 1..13, an integer; derived here from the Lagrange formula. */
void orc_lagrange13(double *a13)
{
    const double fact12 = 479001600.0;
    double factor = 1.0 / fact12;
    for (int i = 1; i <= 13; ++i) {
        for (int j = 1; j <= 13; ++j) {
            double b;
            if (i == j) {
                double s = 0.0;
                for (int m = 1; m <= 13; ++m)
                    if (m != j) s += fact12 / (double)(j - m);
                b = s;
            } else {
                double num = 1.0, den = 1.0;
                for (int m = 1; m <= 13; ++m) {
                    if (m == j) continue;
                    den *= (double)(j - m);
                    if (m != i) num *= (double)(i - m);
                }
                b = fact12 * num / den;
            }
            a13[(i - 1) * 13 + (j - 1)] = (double)llround(b) * factor;
        }
    }
}

/* ------------------------------------------------------------ quadrature */

/* lib9290/quad.f90:388-437. */
double orc_quad(const double *r, const double *rp, const double *ncc, double *ta, int mtp)
{
    const double C1 = ncc[0], C2 = ncc[1], C3 = ncc[2], C4 = ncc[3];
    int mtpm1 = mtp - 1;
    for (int i = 2; i <= mtpm1; ++i) {
        double tai = F1(ta, i);
        if (fabs(tai) <= 0.0) continue;
        int ip1 = i + 1;
        double taip1 = F1(ta, ip1);
        double quott = taip1 / tai;
        if (quott <= 0.0) continue;
        double frip1 = taip1 / F1(rp, ip1);
        double fri = tai / F1(rp, i);
        double ratio = frip1 / fri;
        double rip1 = F1(r, ip1), ri = F1(r, i);
        double sigma = log(ratio) / log(rip1 / ri);
        fri = ri * fri;
        double result = fri / (sigma + 1.0);
        F1(ta, mtp + 1) = 0.0;
        F1(ta, mtp + 2) = 0.0;
        F1(ta, mtp + 3) = 0.0;
        result = result + C1 * tai;
        double s = 0.0;
        for (int a = ip1; a <= mtp; a += 4)
            s += C2 * (F1(ta, a) + F1(ta, a + 2)) + C3 * F1(ta, a + 1) + C4 * F1(ta, a + 3);
        result = result + s;
        if ((mtp - i) % 4 == 0) result = result - C1 * F1(ta, mtp);
        return result;
    }
    return 0.0;
}

/* Integer power by repeated squaring (Fortran R**K with integer K). */
static double ipow(double a, int k)
{
    double res = 1.0;
    for (;;) {
        if (k & 1) res *= a;
        k >>= 1;
        if (!k) break;
        a *= a;
    }
    return res;
}

/* ------------------------------------------------------------------ YZK */

/* lib9290/yzk.f90:78-223.  Orbitals are 0-based indices i, j. */
int orc_yzk(int n, const double *r, const double *rp, const double *ncc,
            const double *cnc5c, double accy, int k,
            const double *P, const double *Q, int ld, const int *mf,
            int i, int j, double *yk)
{
    const double C1 = ncc[0], C2 = ncc[1], C3 = ncc[2];
    const double *Pi = P + (size_t)i * ld, *Pj = P + (size_t)j * ld;
    const double *Qi = Q + (size_t)i * ld, *Qj = Q + (size_t)j * ld;
    int sz = n + 16;
    double *rhop = calloc(sz, sizeof(double));
    double *temp = calloc(sz, sizeof(double));
    double *zk = calloc(sz, sizeof(double));
    double *wk = calloc(sz, sizeof(double));
    double *rttk = calloc(sz, sizeof(double)), *rttkm = calloc(sz, sizeof(double));
    double *rttk1 = calloc(sz, sizeof(double)), *rttkm1 = calloc(sz, sizeof(double));
    double *rm = calloc(sz, sizeof(double));
    /* yzk.f90:84-92: powers of r on 2..N */
    for (int ii = 2; ii <= n; ++ii) {
        if (k > 0) {
            F1(rttk, ii) = ipow(F1(r, ii), k);
            F1(rttkm, ii) = 1.0 / F1(rttk, ii);
            F1(rttk1, ii) = F1(rttk, ii) * F1(r, ii);
            F1(rttkm1, ii) = 1.0 / F1(rttk1, ii);
        }
        F1(rm, ii) = 1.0 / F1(r, ii);
    }
    int mtp = mf[i] < mf[j] ? mf[i] : mf[j];
    int mtpp1 = mtp + 1, mtpp3 = mtp + 3, mtpp4 = mtp + 4;
    for (int ii = 2; ii <= mtp; ++ii)
        F1(rhop, ii) = F1(rp, ii) * (F1(Pi, ii) * F1(Pj, ii) + F1(Qi, ii) * F1(Qj, ii));
    F1(temp, 1) = 0.0;
    for (int ii = 2; ii <= mtp; ++ii)
        F1(temp, ii) = (k == 0) ? F1(rhop, ii) : F1(rttk, ii) * F1(rhop, ii);
    for (int ii = mtpp1; ii <= mtpp4; ++ii) F1(temp, ii) = 0.0;
    /* yzk.f90:128-144: semi-open start, then the closed 5-point march */
    F1(zk, 1) = 0.0;
    for (int ii = 2; ii <= 4; ++ii) {
        double sum = 0.0;
        for (int kk = 2; kk <= 5; ++kk) sum = sum + cnc5c[(ii - 2) * 4 + (kk - 2)] * F1(temp, kk);
        F1(zk, ii) = sum;
    }
    for (int ii = 5; ii <= mtpp3; ++ii) {
        double rtmp = C1 * (F1(temp, ii - 4) + F1(temp, ii)) + C2 * (F1(temp, ii - 3) + F1(temp, ii - 1))
                      + C3 * F1(temp, ii - 2);
        F1(zk, ii) = F1(zk, ii - 4) + rtmp;
    }
    /* yzk.f90:154-173: asymptotic value; K = 0 correction chains */
    double zklim;
    if (k == 0) {
        zklim = (i == j) ? 1.0 : 0.0;
        for (int kk = mtpp3; kk >= mtp; --kk) {
            double dif = F1(zk, kk) - zklim;
            if (fabs(dif) <= accy) continue;
            for (int idx = kk; idx >= 2; idx -= 4) F1(zk, idx) = F1(zk, idx) - dif;
        }
    } else {
        zklim = F1(zk, mtpp3);
    }
    /* yzk.f90:177-187 */
    if (k == 0) {
        for (int ii = mtpp4; ii <= n; ++ii) F1(zk, ii) = zklim;
    } else {
        for (int ii = 2; ii <= mtpp3; ++ii) F1(zk, ii) = F1(zk, ii) * F1(rttkm, ii);
        for (int ii = mtpp4; ii <= n; ++ii) F1(zk, ii) = zklim * F1(rttkm, ii);
    }
    /* yzk.f90:191-218: inward march for W */
    int np4 = n + 4;
    for (int ii = np4; ii >= mtpp1; --ii) F1(wk, ii) = 0.0;
    F1(temp, 1) = 0.0;
    for (int ii = 2; ii <= mtp; ++ii)
        F1(temp, ii) = (k == 0) ? F1(rhop, ii) * F1(rm, ii) : F1(rhop, ii) * F1(rttkm1, ii);
    for (int ii = mtp; ii >= 2; --ii)
        F1(wk, ii) = F1(wk, ii + 4) + C1 * (F1(temp, ii) + F1(temp, ii + 4))
                     + C2 * (F1(temp, ii + 1) + F1(temp, ii + 3)) + C3 * F1(temp, ii + 2);
    F1(wk, 1) = 0.0;
    for (int ii = 2; ii <= mtp; ++ii)
        F1(wk, ii) = F1(wk, ii) * ((k == 0) ? F1(r, ii) : F1(rttk1, ii));
    /* yzk.f90:222-223 */
    F1(yk, 1) = 0.0;
    for (int ii = 2; ii <= n; ++ii) F1(yk, ii) = F1(zk, ii) + F1(wk, ii);
    free(rhop); free(temp); free(zk); free(wk);
    free(rttk); free(rttkm); free(rttk1); free(rttkm1); free(rm);
    return mtp;
}

/* lib9290/slater.f90:278-294. */
double orc_slater(int n, const double *r, const double *rp, const double *rpor,
                  const double *ncc, const double *cnc5c, double accy,
                  const double *P, const double *Q, int ld, const int *mf,
                  int ia, int ib, int ic, int id, int k)
{
    double *tb = calloc(n + 16, sizeof(double));
    double *ta = calloc(n + 16, sizeof(double));
    int mtp = orc_yzk(n, r, rp, ncc, cnc5c, accy, k, P, Q, ld, mf, ib - 1, id - 1, tb);
    int mfa = mf[ia - 1], mfc = mf[ic - 1];
    if (k == 0 && ib == id) {
        mtp = mfa < mfc ? mfa : mfc;
    } else {
        if (mfa < mtp) mtp = mfa;
        if (mfc < mtp) mtp = mfc;
    }
    const double *Pa = P + (size_t)(ia - 1) * ld, *Pc = P + (size_t)(ic - 1) * ld;
    const double *Qa = Q + (size_t)(ia - 1) * ld, *Qc = Q + (size_t)(ic - 1) * ld;
    F1(ta, 1) = 0.0;
    for (int i = 2; i <= mtp; ++i)
        F1(ta, i) = (F1(Pa, i) * F1(Pc, i) + F1(Qa, i) * F1(Qc, i)) * F1(rpor, i) * F1(tb, i);
    double res = orc_quad(r, rp, ncc, ta, mtp);
    free(tb);
    free(ta);
    return res;
}

/* -------------------------------------------------------- nuclear model */

/* lib9290/es.f90 */
void orc_es(double f, double *s2f, double *s3f)
{
    int nn = 0;
    double s2 = 0.0, s3 = 0.0, fase = 1.0, s2last;
    nn = nn + 1;
    double en = (double)nn, obn = 1.0 / en;
    fase = -fase;
    double enf = exp(en * f);
    double term2 = fase * enf * obn * obn, term3 = term2 * obn;
    s2last = s2;
    s2 = s2 + term2;
    s3 = s3 + term3;
    while (fabs(s2) != fabs(s2last)) {
        nn = nn + 1;
        en = (double)nn;
        obn = 1.0 / en;
        fase = -fase;
        enf = exp(en * f);
        term2 = fase * enf * obn * obn;
        term3 = term2 * obn;
        s2last = s2;
        s2 = s2 + term2;
        s3 = s3 + term3;
    }
    *s2f = s2;
    *s3f = s3;
}

/* lib9290/nucpot.f90 (NPARM = 0 and NPARM = 2 branches). */
void orc_nucpot(int n, const double *r, double z, int nparm, const double *parm, double *zz)
{
    if (nparm == 0) {
        for (int i = 0; i < n; ++i) zz[i] = z;
        return;
    }
    const double pi = 4.0 * atan(1.0);
    double c = parm[0], a = parm[1];
    double abc = a / c, tabc = 2.0 * abc, abc2 = abc * abc, thabc2 = 3.0 * abc2;
    double abc3 = abc2 * abc, cba = c / a, pi2 = pi * pi;
    double hpiac2 = 0.5 * pi2 * abc2, h3php = 1.5 + hpiac2;
    double s2mcba, s3mcba;
    orc_es(-cba, &s2mcba, &s3mcba);
    double sabc3 = 6.0 * abc3, dmsas = -sabc3 * s3mcba;
    double en = 1.0 + abc2 * pi2 + dmsas, zbn = z / en;
    for (int i = 1; i <= n; ++i) {
        double ri = F1(r, i), rmc = ri - c, rmcba = rmc / a, rbc = ri / c;
        double s2, s3;
        if (rbc <= 1.0) {
            orc_es(rmcba, &s2, &s3);
            F1(zz, i) = zbn * (dmsas + sabc3 * s3 + rbc * (h3php - thabc2 * s2 - 0.5 * rbc * rbc));
        } else {
            orc_es(-rmcba, &s2, &s3);
            F1(zz, i) = z * (1.0 + thabc2 * (rbc * s2 + tabc * s3) / en);
        }
    }
}

/* ------------------------------------------------------ one-body I(ab) */

/* lib9290/dpbdt.f90 */
void orc_dpbdt(int n, const double *a13, const double *P, const double *Q, int ld, int j,
               double *ta, double *tb)
{
#define A(i, jj) a13[((i) - 1) * 13 + ((jj) - 1)]
    const double *Pj = P + (size_t)j * ld, *Qj = Q + (size_t)j * ld;
    double a1 = A(7, 1), a2 = A(7, 2), a3 = A(7, 3), a4 = A(7, 4), a5 = A(7, 5), a6 = A(7, 6);
    for (int i = 1; i <= 6; ++i) {
        double hp = 0.0, hq = 0.0;
        for (int c = 1; c <= 13; ++c) {
            hp += A(i, c) * F1(Pj, c);
            hq += A(i, c) * F1(Qj, c);
        }
        F1(ta, i) = hp;
        F1(tb, i) = hq;
    }
    for (int i = 7; i <= n - 6; ++i) {
        F1(ta, i) = a1 * (F1(Pj, i - 6) - F1(Pj, i + 6)) + a2 * (F1(Pj, i - 5) - F1(Pj, i + 5))
                    + a3 * (F1(Pj, i - 4) - F1(Pj, i + 4)) + a4 * (F1(Pj, i - 3) - F1(Pj, i + 3))
                    + a5 * (F1(Pj, i - 2) - F1(Pj, i + 2)) + a6 * (F1(Pj, i - 1) - F1(Pj, i + 1));
        F1(tb, i) = a1 * (F1(Qj, i - 6) - F1(Qj, i + 6)) + a2 * (F1(Qj, i - 5) - F1(Qj, i + 5))
                    + a3 * (F1(Qj, i - 4) - F1(Qj, i + 4)) + a4 * (F1(Qj, i - 3) - F1(Qj, i + 3))
                    + a5 * (F1(Qj, i - 2) - F1(Qj, i + 2)) + a6 * (F1(Qj, i - 1) - F1(Qj, i + 1));
    }
    for (int i = n - 5; i <= n; ++i) {
        int irow = i - n + 13;
        double hp = 0.0, hq = 0.0;
        for (int c = 1; c <= 13; ++c) {
            hp += A(irow, c) * F1(Pj, n - 13 + c);
            hq += A(irow, c) * F1(Qj, n - 13 + c);
        }
        F1(ta, i) = hp;
        F1(tb, i) = hq;
    }
#undef A
}

/* lib9290/rinti.f90:49-103 (j, k 1-based). */
double orc_rinti(int n, const double *r, const double *rp, const double *rpor,
                 const double *ncc, const double *a13, double h, double c,
                 const double *zz, const double *P, const double *Q, int ld,
                 const int *mf, const int *nak, int j, int k, int mode)
{
    const double *Pj = P + (size_t)(j - 1) * ld, *Qj = Q + (size_t)(j - 1) * ld;
    const double *Pk = P + (size_t)(k - 1) * ld, *Qk = Q + (size_t)(k - 1) * ld;
    int mtp = mf[j - 1] > mf[k - 1] ? mf[j - 1] : mf[k - 1];
    double *ta = calloc(n + 16, sizeof(double)), *tb = calloc(n + 16, sizeof(double));
    orc_dpbdt(n, a13, P, Q, ld, k - 1, ta, tb);
    F1(ta, 1) = 0.0;
    for (int i = 2; i <= mtp; ++i) F1(ta, i) = F1(Qj, i) * F1(ta, i) - F1(Pj, i) * F1(tb, i);
    double p1 = orc_quad(r, rp, ncc, ta, mtp);
    p1 = c * p1 / h;
    F1(ta, 1) = 0.0;
    for (int i = 2; i <= mtp; ++i) F1(ta, i) = F1(rp, i) * F1(Qj, i) * F1(Qk, i);
    double p2 = orc_quad(r, rp, ncc, ta, mtp);
    p2 = -2.0 * c * c * p2;
    F1(ta, 1) = 0.0;
    for (int i = 2; i <= mtp; ++i)
        F1(ta, i) = F1(rpor, i) * (F1(Pj, i) * F1(Qk, i) + F1(Qj, i) * F1(Pk, i));
    double p3 = orc_quad(r, rp, ncc, ta, mtp);
    p3 = p3 * c * (double)nak[k - 1];
    double p4 = 0.0;
    if (mode == 0) {
        F1(ta, 1) = 0.0;
        for (int i = 2; i <= mtp; ++i)
            F1(ta, i) = F1(rpor, i) * F1(zz, i) * (F1(Pj, i) * F1(Pk, i) + F1(Qj, i) * F1(Qk, i));
        p4 = orc_quad(r, rp, ncc, ta, mtp);
        p4 = -p4;
    }
    free(ta);
    free(tb);
    return p1 + p2 + p3 + p4;
}

/* --------------------------------------------------- R^k table (GENINTRK) */

/* rci90/triangrk.f90 */
int orc_triangrk(int la, int k, int lb)
{
    if ((k + la + lb) % 2 != 0) return 0;
    if (abs(la - lb) > k) return 0;
    if (la + lb < k) return 0;
    return 1;
}

/* rci90/genintrk.f90:40-86. */
int64_t orc_genintrk_enum(int nw, const int *nkl, const int *nkj, int *j2max,
                          int32_t *kstart, int32_t *keys, int32_t *kval, int32_t *labels)
{
    int key = nw + 1;
    int jm = nkj[0];
    for (int i = 1; i < nw; ++i)
        if (nkj[i] > jm) jm = nkj[i];
    *j2max = jm;
    int64_t cnt = 0;
    if (kstart) kstart[0] = 1;
    for (int k = 0; k <= jm; ++k) {
        for (int ia = 1; ia <= nw; ++ia)
            for (int ib = ia; ib <= nw; ++ib)
                for (int ic = ia; ic <= nw; ++ic) {
                    if (!orc_triangrk(nkl[ia - 1], k, nkl[ic - 1])) continue;
                    for (int id = ib; id <= nw; ++id) {
                        if (!orc_triangrk(nkl[ib - 1], k, nkl[id - 1])) continue;
                        if (keys) keys[cnt] = ((ia * key + ib) * key + ic) * key + id;
                        if (kval) kval[cnt] = k;
                        if (labels) {
                            labels[4 * cnt + 0] = ia;
                            labels[4 * cnt + 1] = ib;
                            labels[4 * cnt + 2] = ic;
                            labels[4 * cnt + 3] = id;
                        }
                        ++cnt;
                    }
                }
        if (kstart) kstart[k + 1] = (int32_t)(cnt + 1);
    }
    return cnt;
}

/* rci90/rkintc.f90:40-103 (returns -1 where the reference STOPs). */
int orc_rkintc(int nw, const int32_t *kstart, const int32_t *keys, const double *vals,
               int ia, int ib, int ic, int id, int k, double *val)
{
    int key = nw + 1, iswap;
    if (ia > ic) { iswap = ic; ic = ia; ia = iswap; }
    if (ib > id) { iswap = id; id = ib; ib = iswap; }
    if (ia > ib) {
        iswap = ib; ib = ia; ia = iswap;
        iswap = id; id = ic; ic = iswap;
    }
    int32_t index = ((ia * key + ib) * key + ic) * key + id;
    int jl = kstart[k], ju = kstart[k + 1] - 1;
    if (ju < jl) return -1;
    if (index < F1(keys, jl) || index > F1(keys, ju)) return -1;
    while (ju - jl > 1) {
        int jm = (ju + jl) / 2;
        if (F1(keys, jm) > index) ju = jm;
        else jl = jm;
    }
    int loc;
    if (index == F1(keys, ju)) loc = ju;
    else if (index == F1(keys, jl)) loc = jl;
    else return -1;
    *val = F1(vals, loc);
    return 0;
}

/* ----------------------------------------------------- H assembly (SETHAM) */

/* rci90/setham_gg.f90:162-385 restricted to the Dirac-Coulomb terms (no
 * LNMS/LVP/LSMS/LTRANS), CUTOFF = 1e-12 (setham_gg.f90:105).  The one-body
 * integral lookup follows IABINT (rci90/iabint.f90): labels are swapped so
 * that IA <= IB before the lookup. */
int64_t orc_setham_tb(int ncf, int nw, int64_t npairs,
                      const int32_t *pair_ic, const int32_t *pair_ir,
                      const int32_t *one_ia, const int32_t *one_ib,
                      const int64_t *tsh_off, const double *tshell,
                      const int64_t *tb_beg, const int64_t *tb_end,
                      const int32_t *label, const double *coeff,
                      const double *iab,
                      const int32_t *kstart, const int32_t *keys, const double *vals,
                      const int64_t *bt_beg, const int64_t *bt_end,
                      const int32_t *blabel, const double *bcoeff, orc_breit *br,
                      int lfordr, int iccut,
                      int64_t *row_end, int32_t *col, double *val, double *eav_sum,
                      double *elsto)
{
    const double cutoff = 1.0e-12;
    int64_t nnz = 0, p = 0;
    double eav = 0.0;
    if (elsto) *elsto = 0.0;
    row_end[0] = 0;
    for (int ic = 1; ic <= ncf; ++ic) {
        int has_diag = 0;
        for (; p < npairs && pair_ic[p] == ic; ++p) {
            int ir = pair_ir[p];
            if (ir < 1 || ir > ic) return -1;
            if (lfordr && ir > iccut && ir != ic) continue;
            double elemnt = 0.0;
            int ia = one_ia[p], ib = one_ib[p];
            if (ia != 0) {
                const double *ts = tshell + tsh_off[p];
                if (ia == ib) {
                    for (int a = 1; a <= nw; ++a) {
                        double t = ts[a - 1];
                        if (fabs(t) > cutoff) elemnt = elemnt + iab[(a - 1) * nw + (a - 1)] * t;
                    }
                } else {
                    double t = ts[0];
                    if (fabs(t) > cutoff) {
                        int lo = ia < ib ? ia : ib, hi = ia < ib ? ib : ia;
                        elemnt = elemnt + iab[(lo - 1) * nw + (hi - 1)] * t;
                    }
                }
            }
            for (int64_t e = tb_beg[p]; e < tb_end[p]; ++e) {
                double v = coeff[e];
                if (fabs(v) > cutoff) {
                    const int32_t *lb = label + 5 * e;
                    double tegral;
                    if (orc_rkintc(nw, kstart, keys, vals, lb[0], lb[1], lb[2], lb[3], lb[4], &tegral))
                        return -1;
                    elemnt = elemnt + tegral * v;
                }
            }
            /* transverse Breit (setham_gg.f90:298-357) after the
             * Dirac-Coulomb terms */
            if (br && orc_setham_breit_pair(br, bt_beg[p], bt_end[p], blabel, bcoeff, &elemnt, elsto))
                return -1;
            if (ir == ic || fabs(elemnt) > cutoff) {
                col[nnz] = ir;
                val[nnz] = elemnt;
                ++nnz;
                if (ir == ic) {
                    has_diag = 1;
                    eav = eav + elemnt;
                }
            }
        }
        if (!has_diag) return -1; /* every row carries its diagonal, last */
        row_end[ic] = nnz;
    }
    if (p != npairs) return -1;
    *eav_sum = eav;
    return nnz;
}

int64_t orc_setham(int ncf, int nw, int64_t npairs,
                   const int32_t *pair_ic, const int32_t *pair_ir,
                   const int32_t *one_ia, const int32_t *one_ib,
                   const int64_t *tsh_off, const double *tshell,
                   const int64_t *tb_beg, const int64_t *tb_end,
                   const int32_t *label, const double *coeff,
                   const double *iab,
                   const int32_t *kstart, const int32_t *keys, const double *vals,
                   int lfordr, int iccut,
                   int64_t *row_end, int32_t *col, double *val, double *eav_sum)
{
    return orc_setham_tb(ncf, nw, npairs, pair_ic, pair_ir, one_ia, one_ib, tsh_off, tshell, tb_beg, tb_end,
                         label, coeff, iab, kstart, keys, vals, NULL, NULL, NULL, NULL, NULL, lfordr, iccut,
                         row_end, col, val, eav_sum, NULL);
}

/* ------------------------------------------------------------ SPICMV2 */

/* lib9290/spicmv2.f90:40-56 with DMERGE (dmerge.f90:36-40). */
void orc_spicmv2(int n, int m, const int64_t *row_end, const int32_t *col,
                 const double *val, const double *B, double *C)
{
    memset(C, 0, sizeof(double) * (size_t)n * m);
    int64_t ibeg = 1;
    for (int icol = 1; icol <= n; ++icol) {
        int64_t iend = row_end[icol];
        int64_t nelc = iend - ibeg + 1;
        for (int iv = 0; iv < m; ++iv) {
            const double *b = B + (size_t)iv * n;
            double *c = C + (size_t)iv * n;
            double diag = F1(c, icol) + F1(val, iend) * F1(b, icol);
            double dsum = 0.0, dconst = F1(b, icol);
            for (int64_t t = 0; t < nelc - 1; ++t) {
                int irow = col[ibeg - 1 + t];
                double da = val[ibeg - 1 + t];
                dsum = dsum + da * F1(b, irow);
                F1(c, irow) = F1(c, irow) + dconst * da;
            }
            F1(c, icol) = diag + dsum;
        }
        ibeg = iend + 1;
    }
}

/* mpi90/spicmvmpi.f90:61-75 pattern: row-cyclic rows per worker, private
 * output, then the sum the MPI_Allreduce performs. */
void orc_spicmv2_mt(int n, int m, const int64_t *row_end, const int32_t *col,
                    const double *val, const double *B, double *C, int nthreads)
{
    if (nthreads <= 1) {
        orc_spicmv2(n, m, row_end, col, val, B, C);
        return;
    }
    size_t nm = (size_t)n * m;
    double *priv = calloc(nm * nthreads, sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        double *Ct = priv + nm * t;
        for (int icol = t + 1; icol <= n; icol += nthreads) {
            int64_t ibeg = row_end[icol - 1] + 1, iend = row_end[icol];
            for (int iv = 0; iv < m; ++iv) {
                const double *b = B + (size_t)iv * n;
                double *c = Ct + (size_t)iv * n;
                double dsum = 0.0, dconst = F1(b, icol);
                for (int64_t e = ibeg; e < iend; ++e) {
                    int irow = F1(col, e);
                    double da = F1(val, e);
                    dsum = dsum + da * F1(b, irow);
                    F1(c, irow) = F1(c, irow) + dconst * da;
                }
                F1(c, icol) = F1(c, icol) + F1(val, iend) * dconst + dsum;
            }
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (size_t q = 0; q < nm; ++q) {
            double s = 0.0;
            for (int tt = 0; tt < nthreads; ++tt) s += priv[nm * tt + q];
            C[q] = s;
        }
    }
    free(priv);
}

/* ------------------------------------------------ dense eigensolver */

/* Householder reduction to tridiagonal form followed by the implicit QL
 * iteration with accumulated transformations: the tred2/tql2 procedures of
 * Wilkinson & Reinsch (Handbook for Automatic Computation, vol. II).
 * a: n x n symmetric, column-major, lower triangle used; on exit the
 * eigenvectors (column j for w[j]); eigenvalues ascending. */
int orc_syev(int n, double *a, double *w)
{
    if (n <= 0) return 0;
    double *e = calloc(n, sizeof(double));
    double *d = w;
#define Z(i, j) a[(size_t)(j) * n + (i)]
    /* tred2 */
    for (int i = n - 1; i > 0; --i) {
        int l = i - 1;
        double h = 0.0, scale = 0.0;
        if (l > 0) {
            for (int k = 0; k <= l; ++k) scale += fabs(Z(i, k));
            if (scale == 0.0) {
                e[i] = Z(i, l);
            } else {
                for (int k = 0; k <= l; ++k) {
                    Z(i, k) /= scale;
                    h += Z(i, k) * Z(i, k);
                }
                double f = Z(i, l);
                double g = (f >= 0.0) ? -sqrt(h) : sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                Z(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; ++j) {
                    Z(j, i) = Z(i, j) / h;
                    g = 0.0;
                    for (int k = 0; k <= j; ++k) g += Z(j, k) * Z(i, k);
                    for (int k = j + 1; k <= l; ++k) g += Z(k, j) * Z(i, k);
                    e[j] = g / h;
                    f += e[j] * Z(i, j);
                }
                double hh = f / (h + h);
                for (int j = 0; j <= l; ++j) {
                    f = Z(i, j);
                    e[j] = g = e[j] - hh * f;
                    for (int k = 0; k <= j; ++k) Z(j, k) -= (f * e[k] + g * Z(i, k));
                }
            }
        } else {
            e[i] = Z(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; ++i) {
        int l = i - 1;
        if (d[i] != 0.0) {
            for (int j = 0; j <= l; ++j) {
                double g = 0.0;
                for (int k = 0; k <= l; ++k) g += Z(i, k) * Z(k, j);
                for (int k = 0; k <= l; ++k) Z(k, j) -= g * Z(k, i);
            }
        }
        d[i] = Z(i, i);
        Z(i, i) = 1.0;
        for (int j = 0; j <= l; ++j) Z(j, i) = Z(i, j) = 0.0;
    }
    /* tql2 */
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    int rc = 0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 1e-300 + 2.220446049250313e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 60) { rc = -1; break; }
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i], b = c * e[i];
                    e[i + 1] = (r = hypot(f, g));
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {
                        f = Z(k, i + 1);
                        Z(k, i + 1) = s * Z(k, i) + c * f;
                        Z(k, i) = c * Z(k, i) - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
        if (rc) break;
    }
    /* sort ascending */
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        double p = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < p) p = d[k = j];
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int j = 0; j < n; ++j) {
                double t = Z(j, i);
                Z(j, i) = Z(j, k);
                Z(j, k) = t;
            }
        }
    }
#undef Z
    free(e);
    return rc;
}

/* Tridiagonal solve (T - lam I) y = b by Gaussian elimination with row
 * interchanges (the factorisation of LAPACK's DGTTRF/DGTTS2, restated);
 * pivots smaller than `tiny` are replaced by +-tiny, as inverse iteration
 * perturbs an exactly singular shift (DSTEIN/DLAGTF). */
static void tri_shift_solve(int n, const double *d, const double *e, double lam, double tiny,
                            double *dd, double *du, double *du2, double *dl, int *piv, double *b)
{
    for (int i = 0; i < n; ++i) dd[i] = d[i] - lam;
    for (int i = 0; i < n - 1; ++i) {
        du[i] = e[i];
        dl[i] = e[i];
        du2[i] = 0.0;
    }
    for (int i = 0; i < n - 1; ++i) {
        if (fabs(dd[i]) >= fabs(dl[i])) {
            if (fabs(dd[i]) < tiny) dd[i] = dd[i] < 0.0 ? -tiny : tiny;
            double fact = dl[i] / dd[i];
            dl[i] = fact;
            dd[i + 1] -= fact * du[i];
            piv[i] = 0;
        } else {
            double fact = dd[i] / dl[i];
            dd[i] = dl[i];
            dl[i] = fact;
            double temp = du[i];
            du[i] = dd[i + 1];
            dd[i + 1] = temp - fact * dd[i + 1];
            if (i < n - 2) {
                du2[i] = du[i + 1];
                du[i + 1] = -fact * du[i + 1];
            }
            piv[i] = 1;
        }
    }
    if (fabs(dd[n - 1]) < tiny) dd[n - 1] = dd[n - 1] < 0.0 ? -tiny : tiny;
    for (int i = 0; i < n - 1; ++i) {
        if (!piv[i]) {
            b[i + 1] -= dl[i] * b[i];
        } else {
            double temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - dl[i] * b[i];
        }
    }
    b[n - 1] /= dd[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / dd[n - 2];
    for (int i = n - 3; i >= 0; --i) b[i] = (b[i] - du[i] * b[i + 1] - du2[i] * b[i + 2]) / dd[i];
}

/* Number of eigenvalues of the tridiagonal (d, e2 = e^2) below x: the
 * Sturm sequence count of DSTEBZ (negative pivots of T - x I). */
static int sturm_count(int n, const double *d, const double *e2, double x, double pivmin)
{
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int cnt = q < 0.0;
    for (int i = 1; i < n; ++i) {
        q = d[i] - x - e2[i - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

/* Lowest nev eigenpairs of a dense symmetric matrix by the steps of
 * LAPACK's DSPEVX with RANGE = 'I', IL = 1, IU = nev (the call of
 * rci90/maneig.f90:220 and lib9290/iniest2.f90): Householder reduction to
 * tridiagonal form (DSPTRD), bisection on Sturm counts to 2 ulp (DSTEBZ),
 * inverse iteration with reorthogonalisation inside clusters of
 * eigenvalues closer than 1e-3 |T| (DSTEIN), and back-transformation by
 * the reflectors (DOPMTR).  The O(n^3) reduction runs on nthreads OpenMP
 * threads.  a: n x n column-major holding both triangles (destroyed);
 * w[nev] ascending; z: n x nev column-major. */
int orc_spevx_lowest(int n, double *a, int nev, double *w, double *z, int nthreads)
{
    if (n <= 0 || nev <= 0) return 0;
    if (nev > n) nev = n;
    if (nthreads < 1) nthreads = 1;
#define A(i, j) a[(size_t)(j) * n + (i)]
    double *d = malloc(sizeof(double) * n), *e = calloc(n, sizeof(double)), *tau = calloc(n, sizeof(double));
    double *p = malloc(sizeof(double) * n), *wv = malloc(sizeof(double) * n);
    /* reduction: step k annihilates A(k+2:n, k) by H_k = I - tau v v^T,
     * v = (1, A(k+2:n, k)) acting on rows/columns k+1..n-1 */
    for (int k = 0; k < n - 1; ++k) {
        const int m = n - k - 1;
        double *x = &A(k + 1, k);
        d[k] = A(k, k);
        if (m == 1) {
            e[k] = x[0];
            tau[k] = 0.0;
            continue;
        }
        double xnorm = 0.0;
        for (int i = 1; i < m; ++i) xnorm = hypot(xnorm, x[i]);
        const double alpha = x[0];
        if (xnorm == 0.0) {
            e[k] = alpha;
            tau[k] = 0.0;
            continue;
        }
        const double beta = -copysign(hypot(alpha, xnorm), alpha);
        const double t = (beta - alpha) / beta, sc = 1.0 / (alpha - beta);
        for (int i = 1; i < m; ++i) x[i] *= sc;
        e[k] = beta;
        tau[k] = t;
        x[0] = 1.0;  /* v in place while the trailing block is updated */
        /* p = tau A22 v (column i of A22 is contiguous and equals row i) */
#pragma omp parallel for num_threads(nthreads) schedule(static) if (m > 256)
        for (int i = 0; i < m; ++i) {
            const double *c = &A(k + 1, k + 1 + i);
            double sum = 0.0;
            for (int r = 0; r < m; ++r) sum += c[r] * x[r];
            p[i] = t * sum;
        }
        double pv = 0.0;
        for (int i = 0; i < m; ++i) pv += p[i] * x[i];
        const double al = -0.5 * t * pv;
        for (int i = 0; i < m; ++i) wv[i] = p[i] + al * x[i];
        /* A22 -= v w^T + w v^T, both triangles */
#pragma omp parallel for num_threads(nthreads) schedule(static) if (m > 256)
        for (int j = 0; j < m; ++j) {
            double *c = &A(k + 1, k + 1 + j);
            const double vj = x[j], wj = wv[j];
            for (int r = 0; r < m; ++r) c[r] -= x[r] * wj + wv[r] * vj;
        }
        x[0] = beta;
    }
    d[n - 1] = A(n - 1, n - 1);
    /* bisection for the nev lowest eigenvalues */
    const double eps = 2.220446049250313e-16, safmin = 2.2250738585072014e-308;
    double *e2 = malloc(sizeof(double) * n);
    double gl = d[0], gu = d[0], emax2 = 0.0;
    for (int i = 0; i < n; ++i) {
        const double off = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        if (d[i] - off < gl) gl = d[i] - off;
        if (d[i] + off > gu) gu = d[i] + off;
        if (i < n - 1) {
            e2[i] = e[i] * e[i];
            if (e2[i] > emax2) emax2 = e2[i];
        }
    }
    const double pivmin = safmin * (emax2 > 1.0 ? emax2 : 1.0);
    const double tnorm = fabs(gl) > fabs(gu) ? fabs(gl) : fabs(gu);
    gl -= 2.0 * eps * tnorm + 2.0 * pivmin;
    gu += 2.0 * eps * tnorm + 2.0 * pivmin;
    for (int j = 0; j < nev; ++j) {
        double lo = j > 0 ? w[j - 1] - 4.0 * eps * tnorm : gl, hi = gu;
        if (lo < gl) lo = gl;
        for (int it = 0; it < 400; ++it) {
            const double big = fabs(lo) > fabs(hi) ? fabs(lo) : fabs(hi);
            if (hi - lo <= 2.0 * eps * big + 2.0 * pivmin) break;
            const double mid = 0.5 * (lo + hi);
            if (sturm_count(n, d, e2, mid, pivmin) > j) hi = mid;
            else lo = mid;
        }
        w[j] = 0.5 * (lo + hi);
    }
    /* inverse iteration */
    double *dd = malloc(sizeof(double) * n), *du = malloc(sizeof(double) * n), *du2 = malloc(sizeof(double) * n);
    double *dl = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
    int *piv = malloc(sizeof(int) * n);
    const double tiny = eps * (tnorm > 0.0 ? tnorm : 1.0), ortol = 1.0e-3 * tnorm;
    uint64_t seed = 0x9E3779B97F4A7C15ull;
    for (int j = 0; j < nev; ++j) {
        int j0 = j;  /* first member of j's cluster */
        while (j0 > 0 && w[j] - w[j0 - 1] <= ortol) --j0;
        for (int i = 0; i < n; ++i) {
            seed = seed * 6364136223846793005ull + 1442695040888963407ull;
            y[i] = (double)(seed >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        }
        for (int it = 0; it < 4; ++it) {
            tri_shift_solve(n, d, e, w[j], tiny, dd, du, du2, dl, piv, y);
            for (int q = j0; q < j; ++q) {  /* against the cluster's earlier vectors */
                const double *zq = z + (size_t)q * n;
                double c = 0.0;
                for (int i = 0; i < n; ++i) c += zq[i] * y[i];
                for (int i = 0; i < n; ++i) y[i] -= c * zq[i];
            }
            double nrm = 0.0;
            for (int i = 0; i < n; ++i) nrm = hypot(nrm, y[i]);
            for (int i = 0; i < n; ++i) y[i] /= nrm;
        }
        memcpy(z + (size_t)j * n, y, sizeof(double) * n);
    }
    /* back-transformation: z = H_0 H_1 ... H_{n-2} y */
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nev > 1)
    for (int j = 0; j < nev; ++j) {
        double *zj = z + (size_t)j * n;
        for (int k = n - 2; k >= 0; --k) {
            if (tau[k] == 0.0) continue;
            const double *vk = &A(k + 2, k);
            double sum = zj[k + 1];
            for (int i = 0; i < n - k - 2; ++i) sum += vk[i] * zj[k + 2 + i];
            sum *= tau[k];
            zj[k + 1] -= sum;
            for (int i = 0; i < n - k - 2; ++i) zj[k + 2 + i] -= sum * vk[i];
        }
    }
#undef A
    free(d); free(e); free(tau); free(p); free(wv); free(e2);
    free(dd); free(du); free(du2); free(dl); free(y); free(piv);
    return 0;
}

/* Lowest ifind eigenpairs of a packed-upper symmetric matrix (the DSPEVX
 * 'V','I','U' contract used at dvdson.f90:724-726). */
static int packed_lowest(int k, const double *sp, int ifind, double *eig, double *vec)
{
    double *a = malloc(sizeof(double) * (size_t)k * k), *w = malloc(sizeof(double) * k);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i <= j; ++i) {
            double v = sp[i + (size_t)j * (j + 1) / 2];
            a[(size_t)j * k + i] = v;
            a[(size_t)i * k + j] = v;
        }
    int rc = orc_syev(k, a, w);
    for (int t = 0; t < ifind; ++t) {
        eig[t] = w[t];
        memcpy(vec + (size_t)t * k, a + (size_t)t * k, sizeof(double) * k);
    }
    free(a);
    free(w);
    return rc;
}

/* ------------------------------------------------------------ DVDSON */

static double ddot_(int n, const double *x, const double *y)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}

/* dvdson.f90:1417-1508 */
static void mgs_nrm(int n, int kp, int nnew, double *scra, double *b)
{
    for (int pass = 0; pass < 2; ++pass) {
        size_t newstart = (size_t)kp * n;
        size_t kcur = 0;
        for (int k = 0; k < kp; ++k) {
            size_t jcur = newstart;
            for (int j = 0; j < nnew; ++j) scra[j] = ddot_(n, b + newstart + (size_t)j * n, b + kcur);
            for (int j = 0; j < nnew; ++j) {
                for (int mm = 0; mm < n; ++mm) b[jcur + mm] = b[jcur + mm] - scra[j] * b[kcur + mm];
                jcur += n;
            }
            kcur += n;
        }
        for (int k = 0; k < nnew; ++k) {
            size_t jcur = kcur + n;
            double dnm = sqrt(ddot_(n, b + kcur, b + kcur));
            double inv = 1.0 / dnm;
            for (int mm = 0; mm < n; ++mm) b[kcur + mm] *= inv;
            for (int j = k + 1; j < nnew; ++j) scra[j - k - 1] = ddot_(n, b + kcur + (size_t)(j - k) * n, b + kcur);
            for (int j = k + 1; j < nnew; ++j) {
                for (int mm = 0; mm < n; ++mm) b[jcur + mm] = b[jcur + mm] - scra[j - k - 1] * b[kcur + mm];
                jcur += n;
            }
            kcur += n;
        }
    }
}

/* dvdson.f90:484-555 */
static void adds(int n, int kpass, int nncv, const double *basis, const double *ab, double *s)
{
    size_t idstart = (size_t)kpass * n;
    size_t isstart = (size_t)kpass * (kpass + 1) / 2;
    for (int iv = 1; iv <= nncv; ++iv) {
        for (int ibv = 1; ibv <= kpass + iv; ++ibv)
            s[isstart + ibv - 1] = ddot_(n, basis + (size_t)(ibv - 1) * n, ab + idstart);
        isstart += kpass + iv;
        idstart += n;
    }
}

/* dvdson.f90:953-998: B(:,1:m) = B(:,1:k) C(k,m), row by row */
static void multbc(int n, int k, int m, const double *c, double *temp, double *b)
{
    for (int irow = 0; irow < n; ++irow) {
        for (int j = 0; j < m; ++j) {
            double s = 0.0;
            for (int i = 0; i < k; ++i) s += c[(size_t)j * k + i] * b[irow + (size_t)i * n];
            temp[j] = s;
        }
        for (int j = 0; j < m; ++j) b[irow + (size_t)j * n] = temp[j];
    }
}

typedef struct {
    int n;
    const int64_t *row_end;
    const int32_t *col;
    const double *val;
    int nthreads;
} csr_op;

static void apply_op(const csr_op *op, int nb, const double *x, double *y)
{
    orc_spicmv2_mt(op->n, nb, op->row_end, op->col, op->val, x, y, op->nthreads);
}

int64_t orc_gdvd_worksize(int n, int lim, int nume)
{
    (void)nume;
    /* dvdson.f90:405-417 layout with NOC = 0 */
    return (int64_t)n * lim * 2 + (int64_t)lim * (lim + 1) + (int64_t)lim * lim + 11 * (int64_t)lim + 8;
}

/* libdvd90/gdvd.f90 + dvdson.f90 (ILOW = 1 .. IHIGH, NOC = 0, HIEND false).
 * work: basis (n*lim) | eigval (lim) | ab (n*lim) | s | temps | svec | scra1
 * (8 lim) | oldval (lim+1).  Initial estimates in basis[0..n*niv). */
int orc_gdvd(int n, const int64_t *row_end, const int32_t *col, const double *val,
             int lim, double *diag, int ihigh, int niv, int mblock,
             double crite, double critc, double critr, int maxiter,
             double *work, int nthreads, int *nloops_out, int *nmv_out)
{
    csr_op op = {n, row_end, col, val, nthreads};
    int ierr = 0, nume = ihigh, neig = ihigh;
    /* dvdson.f90:319-398 input checks */
    if (lim > n) ierr += 1;
    if (lim <= 0) ierr += 2;
    if (ihigh > n) ierr += 8;
    if (neig <= 0) ierr += 16;
    if (neig > lim) ierr += 32;
    if (nume > lim || (nume == lim && nume != n)) ierr += 128;
    if (mblock < 1 || mblock > neig) ierr += 256;
    if (niv > lim) ierr += 4096;
    if (niv == 0) return -1000;
    if (ierr) return ierr;
    int *iselec = malloc(sizeof(int) * (lim + 1));
    for (int i = 0; i < neig; ++i) iselec[i] = i + 1;

    double *basis = work;
    double *eigval = basis + (size_t)n * lim;
    double *ab = eigval + lim;
    double *s = ab + (size_t)n * lim;
    double *temps = s + (size_t)lim * (lim + 1) / 2;
    double *svec = temps + (size_t)lim * (lim + 1) / 2;
    double *scra1 = svec + (size_t)lim * lim;
    double *oldval = scra1 + 8 * (size_t)lim;
    int *incv = malloc(sizeof(int) * (lim + 1));
    int *icv = malloc(sizeof(int) * (nume + 2));
    int *ind = malloc(sizeof(int) * (lim + 1));
    double *rowlast = malloc(sizeof(double) * (lim + 1));
    double *tmpv = malloc(sizeof(double) * (lim + 1));
    int nmv = 0;

    /* INITDVD (dvdson.f90:848-948) with NUME+1 vectors */
    {
        int nume1 = nume + 1;
        int ist = 1, iend = niv, nnew = niv;
        for (;;) {
            apply_op(&op, nnew, basis + (size_t)(ist - 1) * n, ab + (size_t)(ist - 1) * n);
            nmv += nnew;
            if (iend < nume1) {
                nnew = nume1 - iend < iend - ist + 1 ? nume1 - iend : iend - ist + 1;
                memcpy(basis + (size_t)iend * n, ab + (size_t)(ist - 1) * n, sizeof(double) * (size_t)n * nnew);
                mgs_nrm(n, iend, nnew, scra1, basis);
                ist = iend + 1;
                iend = iend + nnew;
                continue;
            }
            break;
        }
        niv = iend;
        adds(n, 0, niv, basis, ab, s);
    }

    /* DVDRVR (dvdson.f90:690-836) */
    int nloops = 1;
    for (int i = 0; i < nume; ++i) { eigval[i] = 1.0e30; icv[i] = 0; }
    int kpass = niv, nncv = kpass;
    int rest = lim / 2 > 2 * nume ? lim / 2 : 2 * nume;
    if (rest > lim - 3) rest = lim - 3;
    int done = 0;
    for (;;) {
        int ifind = kpass < rest ? kpass : rest;
        memcpy(oldval, eigval, sizeof(double) * ifind);
        memcpy(temps, s, sizeof(double) * (size_t)kpass * (kpass + 1) / 2);
        if (packed_lowest(kpass, temps, ifind, eigval, svec)) { ierr = -1; goto out; }
        /* TSTSEL (dvdson.f90:1260-1413) */
        {
            int nnce = 0;
            for (int i = 0; i < neig; ++i) {
                int ival = iselec[i];
                if (fabs(F1(oldval, ival) - F1(eigval, ival)) < crite) continue;
                ++nnce;
            }
            if (nnce == 0) {
                done = 1;
            } else {
                int icnt = 0;
                for (int i = 0; i < neig; ++i) {
                    int icur = kpass * iselec[i];
                    double tmax = fabs(F1(svec, icur));
                    for (int l = 1; l <= nncv - 1; ++l) {
                        double t = fabs(F1(svec, icur - l));
                        tmax = tmax > t ? tmax : t;
                    }
                    if (tmax < critc) {
                        F1(icv, iselec[i]) = 1;
                    } else {
                        F1(icv, iselec[i]) = 0;
                        ++icnt;
                        F1(ind, icnt) = iselec[i];
                        F1(rowlast, icnt) = tmax;
                    }
                }
                nncv = icnt;
                if (nncv == 0) done = 1;
                for (int i = 1; i <= nncv; ++i) {
                    /* IDAMAX over rowlast(i..nncv): first maximal |.| */
                    int indx = 1;
                    double best = fabs(F1(rowlast, i));
                    for (int q = 2; q <= nncv - i + 1; ++q)
                        if (fabs(F1(rowlast, i + q - 1)) > best) { best = fabs(F1(rowlast, i + q - 1)); indx = q; }
                    F1(incv, i) = F1(ind, indx + i - 1);
                    double temp = F1(rowlast, indx + i - 1);
                    F1(rowlast, indx + i - 1) = F1(rowlast, i);
                    F1(rowlast, i) = temp;
                    int itemp = F1(ind, indx + i - 1);
                    F1(ind, indx + i - 1) = F1(ind, i);
                    F1(ind, i) = itemp;
                }
            }
        }
        if (done || kpass >= n) break;
        if (kpass == lim) {
            /* OVFLOW (dvdson.f90:1179-1256) with REST */
            multbc(n, kpass, rest, svec, scra1, basis);
            multbc(n, kpass, rest, svec, scra1, ab);
            memset(s, 0, sizeof(double) * (size_t)rest * (rest + 1) / 2);
            memset(svec, 0, sizeof(double) * (size_t)rest * rest);
            int ind0 = 0, icur = 0;
            for (int i = 1; i <= rest; ++i) {
                s[ind0 + i - 1] = F1(eigval, i);
                svec[icur + i - 1] = 1.0;
                icur += rest;
                ind0 += i;
            }
            kpass = rest;
        }
        /* NEWVEC (dvdson.f90:1082-1175) */
        {
            double *epsil = scra1 + lim;
            size_t icur = (size_t)kpass * n;
            int nadded = 0, icvc = 0;
            int limadd = lim < mblock + kpass ? lim : mblock + kpass;
            for (int i = 1; i <= nncv; ++i) {
                int indx = F1(incv, i);
                const double *sv = svec + (size_t)(indx - 1) * kpass;
                double *bx = basis + icur, *rx = ab + icur;
                for (int q = 0; q < n; ++q) {
                    double sb = 0.0, sa = 0.0;
                    for (int c = 0; c < kpass; ++c) {
                        sb += basis[q + (size_t)c * n] * sv[c];
                        sa += ab[q + (size_t)c * n] * sv[c];
                    }
                    bx[q] = sb;
                    rx[q] = sa;
                }
                double lam = F1(eigval, indx);
                for (int q = 0; q < n; ++q) rx[q] = rx[q] + (-lam) * bx[q];
                F1(scra1, indx) = sqrt(ddot_(n, rx, rx));
                if (F1(scra1, indx) < critr) {
                    ++icvc;
                    F1(icv, indx) = 1;
                    if (icvc < nncv) continue;
                    done = 1;
                    break;
                }
                double gap;
                if (indx == 1) {
                    gap = fabs(F1(eigval, 2) - F1(eigval, 1));
                } else {
                    double gaplow = fabs(F1(eigval, indx) - F1(eigval, indx - 1));
                    double gapup = fabs(F1(eigval, indx + 1) - F1(eigval, indx));
                    gap = gaplow < gapup ? gaplow : gapup;
                }
                double dl = fabs(F1(oldval, indx) - F1(eigval, indx));
                double res = F1(scra1, indx), rhseps;
                if (gap > res) rhseps = sqrt(dl * res);
                else rhseps = dl < sqrt(dl) ? dl : sqrt(dl);
                F1(epsil, indx) = 0.0;
                for (int q = 0; q < n; ++q) rx[q] = rx[q] + (-rhseps) * bx[q];
                ++nadded;
                F1(incv, nadded) = indx;
                if (nadded + kpass == limadd) break;
                icur += n;
            }
            if (done) break;
            nncv = nadded;
        }
        /* GDVD preconditioner (gdvd.f90:93-134), results into basis(kpass+1..) */
        {
            double *epsil = scra1 + lim;
            size_t icur = 0;
            for (int i = 1; i <= nncv; ++i) {
                int indx = F1(incv, i);
                /* DVDRVR shifts EIGVAL by -EPSIL before and restores after
                 * (dvdson.f90:762-785); EPSIL is zero (dvdson.f90:1157). */
                F1(eigval, indx) = F1(eigval, indx) - F1(epsil, indx);
                double value = F1(eigval, indx);
                const double *w1 = ab + (size_t)kpass * n + icur;
                double *w2 = basis + (size_t)kpass * n + icur;
                for (int j = 0; j < n; ++j) diag[j] = diag[j] - value;
                for (int j = 0; j < n; ++j) {
                    if (fabs(diag[j]) > 1.0e-06) w2[j] = w1[j] / diag[j];
                    else w2[j] = w1[j] * 1.0e06;
                }
                for (int j = 0; j < n; ++j) diag[j] = diag[j] + value;
                F1(eigval, indx) = F1(eigval, indx) + F1(epsil, indx);
                icur += n;
            }
        }
        mgs_nrm(n, kpass, nncv, scra1 + lim, basis);
        apply_op(&op, nncv, basis + (size_t)kpass * n, ab + (size_t)kpass * n);
        nmv += nncv;
        adds(n, kpass, nncv, basis, ab, s);
        kpass += nncv;
        nloops += 1;
        if (nloops <= maxiter) continue;
        ierr += 2048;
        nloops -= 1;
        kpass -= nncv;
        break;
    }
    /* dvdson.f90:819-836 */
    for (int i = 1; i <= nume; ++i) F1(oldval, i) = fabs(F1(oldval, i) - F1(eigval, i));
    multbc(n, kpass, nume, svec, scra1, basis);
    multbc(n, kpass, nume, svec, scra1, ab);
    for (int i = 1; i <= nume; ++i) {
        double *tmp = ab + (size_t)nume * n;
        memcpy(tmp, ab + (size_t)(i - 1) * n, sizeof(double) * n);
        double lam = F1(eigval, i);
        for (int q = 0; q < n; ++q) tmp[q] = tmp[q] + (-lam) * basis[(size_t)(i - 1) * n + q];
        F1(scra1, i) = sqrt(ddot_(n, tmp, tmp));
    }
    /* dvdson.f90:465-474: results after the vectors */
    memcpy(tmpv, eigval, sizeof(double) * nume);
    memcpy(basis + (size_t)n * nume, tmpv, sizeof(double) * nume);
    memcpy(tmpv, oldval, sizeof(double) * nume);
    memcpy(basis + (size_t)(n + 1) * nume, tmpv, sizeof(double) * nume);
    memcpy(tmpv, scra1, sizeof(double) * nume);
    memcpy(basis + (size_t)(n + 2) * nume, tmpv, sizeof(double) * nume);
out:
    *nloops_out = nloops;
    *nmv_out = nmv;
    free(iselec); free(incv); free(icv); free(ind); free(rowlast); free(tmpv);
    return ierr;
}

/* lib9290/iniest2.f90:55-95: the lowest niv eigenvectors of the leading
 * min(nmax, ncf) block (DSPEVX), zero-padded to ncf rows */
static int iniest2_mt(int nmax, int ncf, int niv, const int64_t *row_end, const int32_t *col,
                      const double *val, double *basis, int nthreads)
{
    int ns = nmax < ncf ? nmax : ncf;
    double *a = calloc((size_t)ns * ns, sizeof(double)), *w = malloc(sizeof(double) * ns);
    for (int j = 1; j <= ns; ++j)
        for (int64_t ir = row_end[j - 1] + 1; ir <= row_end[j]; ++ir) {
            int i = F1(col, ir);
            double v = F1(val, ir);
            a[(size_t)(j - 1) * ns + (i - 1)] = v;
            a[(size_t)(i - 1) * ns + (j - 1)] = v;
        }
    double *zv = malloc(sizeof(double) * (size_t)ns * niv);
    int rc = orc_spevx_lowest(ns, a, niv, w, zv, nthreads);
    memset(basis, 0, sizeof(double) * (size_t)ncf * niv);
    for (int j = 0; j < niv; ++j) memcpy(basis + (size_t)j * ncf, zv + (size_t)j * ns, sizeof(double) * ns);
    free(zv);
    free(a);
    free(w);
    return rc;
}

/* lib9290/iniest2.f90:55-95 (on the OpenMP default thread count) */
int orc_iniest2(int nmax, int ncf, int niv, const int64_t *row_end, const int32_t *col,
                const double *val, double *basis)
{
    return iniest2_mt(nmax, ncf, niv, row_end, col, val, basis, omp_get_max_threads());
}

/* rci90/maneig.f90:157-242 (dense, NCF <= IOLPCK = 2000) and :262-555
 * (in-core sparse Davidson), followed by the sign convention of :532-555. */
int orc_maneig(int ncf, const int64_t *row_end, const int32_t *col, const double *val,
               double eav, int nvec, double *eval, double *evec, int nthreads,
               int *nloops, int *nmv)
{
    *nloops = 0;
    *nmv = 0;
    int64_t nnz = row_end[ncf];
    double *sval = malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    memcpy(sval, val, sizeof(double) * (size_t)nnz);
    double *diag = malloc(sizeof(double) * ncf);
    for (int i = 1; i <= ncf; ++i) {
        F1(sval, row_end[i]) = F1(sval, row_end[i]) - eav;
        F1(diag, i) = F1(sval, row_end[i]);
    }
    int rc = 0;
    if (ncf == 1) {
        eval[0] = 0.0;
        evec[0] = 1.0;
    } else if (ncf <= 2000) {
        double *a = calloc((size_t)ncf * ncf, sizeof(double)), *w = malloc(sizeof(double) * ncf);
        for (int j = 1; j <= ncf; ++j)
            for (int64_t ir = row_end[j - 1] + 1; ir <= row_end[j]; ++ir) {
                int i = F1(col, ir);
                a[(size_t)(j - 1) * ncf + (i - 1)] = F1(sval, ir);
                a[(size_t)(i - 1) * ncf + (j - 1)] = F1(sval, ir);
            }
        rc = orc_spevx_lowest(ncf, a, nvec, w, evec, nthreads);
        for (int j = 0; j < nvec; ++j) eval[j] = w[j];
        free(a);
        free(w);
    } else {
        int lim = ncf < 2 * nvec + 60 ? ncf : 2 * nvec + 60;
        int maxitr = nvec * 200 > ncf / 10 ? nvec * 200 : ncf / 10;
        /* storage rule (maneig.f90:262-266, 416-420): NSTORE = NELMNT +
         * NELMNT/2 + (NCF+1)/2 above NINCOR = 2^28 keeps H on disk, and GDVD
         * then runs with MBLOCK = NVEX; in core MBLOCK = 1 */
        int64_t nstore = nnz + nnz / 2 + (ncf + 1) / 2;
        int mblock = nstore > ORC_NINCOR ? nvec : 1;
        double *work = calloc((size_t)orc_gdvd_worksize(ncf, lim, nvec), sizeof(double));
        iniest2_mt(2000, ncf, nvec, row_end, col, sval, work, nthreads);
        rc = orc_gdvd(ncf, row_end, col, sval, lim, diag, nvec, nvec, mblock, 1.0e-17, 1.0e-9, 1.0e-9,
                      maxitr, work, nthreads, nloops, nmv);
        for (int j = 0; j < nvec; ++j) eval[j] = work[(size_t)ncf * nvec + j];
        memcpy(evec, work, sizeof(double) * (size_t)ncf * nvec);
        free(work);
    }
    for (int j = 0; j < nvec; ++j) {
        double *v = evec + (size_t)j * ncf, amax = 0.0;
        int ia = 0;
        for (int i = 0; i < ncf; ++i) {
            double wa = fabs(v[i]);
            if (wa <= amax) continue;
            amax = wa;
            ia = i;
        }
        if (v[ia] >= 0.0) continue;
        for (int i = 0; i < ncf; ++i) v[i] = -v[i];
    }
    free(sval);
    free(diag);
    return rc;
}

/* rmcdhf90/maneig.f90:78-200: MANEIG of the SCF program, the second caller
 * of GDVD(SPICMV2).  Levels iccmin[0..nlev) (1-based, each <= NVEX) of the
 * block; NVECT = NCMAXBLK; NVEX = NIV = min(NVECT, NCF); LIM = min(NCF,
 * 2*NVECT+40) (:104); MAXITR = min(200*NVECT, NCF) (:112); N1000 = 4000.
 * Cold start (warm == 0, DVDFIRST) or NCF <= 4000: INIEST2(4000) (:115-116),
 * which for NCF <= 4000 is the whole answer.  Warm start: the previous
 * iteration's vectors evec[j] (and eval[j]) placed at column iccmin[j]
 * (:118-124).  The given vectors must be orthonormal: INITDVD takes its
 * start block as it is.  The reference leaves the remaining columns of WORK
 * as they were allocated; here column c then holds the unit vector at the
 * c-th smallest diagonal element (first index on ties), orthonormalised
 * against the given columns and the filled ones before it.  NCF > 4000: GDVD with
 * MBLOCK = 1, CRITE 1e-17, CRITC = CRITR = 1e-9 (:139-162).  Output: eval
 * and evec of the levels iccmin (:171-176), sign-fixed as :179-196.  The
 * CSR is the unshifted H; its diagonal is shifted by -eav here. */
int orc_maneig_mcdhf(int ncf, const int64_t *row_end, const int32_t *col, const double *val,
                     double eav, int nvect, int nlev, const int32_t *iccmin, int warm,
                     double *eval, double *evec, int nthreads, int *nloops, int *nmv)
{
    *nloops = 0;
    *nmv = 0;
    if (ncf == 1) {
        for (int j = 0; j < nlev; ++j) {
            eval[j] = 0.0;
            evec[j] = 1.0;
        }
        return 0;
    }
    int64_t nnz = row_end[ncf];
    double *sval = malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    memcpy(sval, val, sizeof(double) * (size_t)nnz);
    double *diag = malloc(sizeof(double) * ncf);
    for (int i = 1; i <= ncf; ++i) {
        F1(sval, row_end[i]) = F1(sval, row_end[i]) - eav;
        F1(diag, i) = F1(sval, row_end[i]);
    }
    const int nvex = nvect < ncf ? nvect : ncf;
    const int lim = ncf < 2 * nvect + 40 ? ncf : 2 * nvect + 40;
    const int maxitr = nvect * 200 < ncf ? nvect * 200 : ncf;
    double *work = calloc((size_t)orc_gdvd_worksize(ncf, lim, nvex), sizeof(double));
    int rc = 0;
    if (!warm || ncf <= 4000) {
        const int ns = ncf < 4000 ? ncf : 4000;
        double *a = calloc((size_t)ns * ns, sizeof(double)), *w = malloc(sizeof(double) * ns);
        for (int j = 1; j <= ns; ++j)
            for (int64_t ir = row_end[j - 1] + 1; ir <= row_end[j]; ++ir) {
                int i = F1(col, ir);
                a[(size_t)(j - 1) * ns + (i - 1)] = F1(sval, ir);
                a[(size_t)(i - 1) * ns + (j - 1)] = F1(sval, ir);
            }
        double *zv = malloc(sizeof(double) * (size_t)ns * nvex);
        rc = orc_spevx_lowest(ns, a, nvex, w, zv, nthreads);
        for (int j = 0; j < nvex; ++j) {
            memcpy(work + (size_t)j * ncf, zv + (size_t)j * ns, sizeof(double) * ns);
            work[(size_t)ncf * nvex + j] = w[j];
        }
        free(zv);
        free(a);
        free(w);
    } else {
        /* columns not given by the caller: unit vectors at the smallest
         * diagonal elements, in ascending order */
        int *given = calloc(nvex, sizeof(int));
        for (int j = 0; j < nlev; ++j) {
            const int c = iccmin[j] - 1;
            memcpy(work + (size_t)c * ncf, evec + (size_t)j * ncf, sizeof(double) * ncf);
            work[(size_t)ncf * nvex + c] = eval[j];
            given[c] = 1;
        }
        char *used = calloc(ncf, 1);
        for (int c = 0; c < nvex; ++c) {
            int best = -1;
            for (int i = 0; i < ncf; ++i)
                if (!used[i] && (best < 0 || diag[i] < diag[best])) best = i;
            used[best] = 1;
            if (!given[c]) work[(size_t)c * ncf + best] = 1.0;
        }
        /* INITDVD takes its start vectors as orthonormal (dvdson.f90:
         * 886-905 orthonormalises only the vectors it adds), so a filled
         * column is made orthogonal to every given column and to the
         * filled columns before it: two classical Gram-Schmidt passes */
        for (int c = 0; c < nvex; ++c) {
            if (given[c]) continue;
            double *v = work + (size_t)c * ncf;
            for (int pass = 0; pass < 2; ++pass)
                for (int q = 0; q < nvex; ++q) {
                    if (q == c || (!given[q] && q > c)) continue;
                    const double *u = work + (size_t)q * ncf;
                    double dot = 0.0;
                    for (int i = 0; i < ncf; ++i) dot += u[i] * v[i];
                    for (int i = 0; i < ncf; ++i) v[i] -= dot * u[i];
                }
            double nrm = 0.0;
            for (int i = 0; i < ncf; ++i) nrm += v[i] * v[i];
            nrm = sqrt(nrm);
            for (int i = 0; i < ncf; ++i) v[i] /= nrm;
        }
        free(used);
        free(given);
    }
    if (rc == 0 && ncf > 4000)
        rc = orc_gdvd(ncf, row_end, col, sval, lim, diag, nvex, nvex, 1, 1.0e-17, 1.0e-9, 1.0e-9, maxitr,
                      work, nthreads, nloops, nmv);
    for (int j = 0; j < nlev; ++j) {
        const int c = iccmin[j] - 1;
        eval[j] = work[(size_t)ncf * nvex + c];
        memcpy(evec + (size_t)j * ncf, work + (size_t)c * ncf, sizeof(double) * ncf);
    }
    for (int j = 0; j < nlev; ++j) {
        double *v = evec + (size_t)j * ncf, amax = 0.0;
        int ia = 0;
        for (int i = 0; i < ncf; ++i) {
            double wa = fabs(v[i]);
            if (wa <= amax) continue;
            amax = wa;
            ia = i;
        }
        if (v[ia] >= 0.0) continue;
        for (int i = 0; i < ncf; ++i) v[i] = -v[i];
    }
    free(work);
    free(sval);
    free(diag);
    return rc;
}
