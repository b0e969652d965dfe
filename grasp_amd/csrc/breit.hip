// breit.hip -- transverse Breit radial integrals of rci90.
//
// Replaces GENINTBREIT1/2 (rci90/genintbreit1.f90:45-104, genintbreit2.f90)
// and the memoised BRINT3..6 (rci90/brint3.f90:30-107) with their
// BRINTF -> BESSEL / BRRA -> RKINT, SKINT -> ZKF, QUAD chain
// (brintf.f90:44-96, bessel.f90:46-241, brra.f90:41-124, rkint.f90:43-81,
// skint.f90:44-109, zkf.f90:33-83).
//
// The reference evaluates integrals one at a time through module state:
// BESSEL keeps four Bessel-function slots KEEP(IK, IW) and two photon
// frequencies WIJ(IW), and an integral reads whatever the slots hold after
// its own BESSEL calls -- in one case (a type-3 integral with a = b and
// c = d) a slot it did not set.  The host therefore replays BRINTF's BESSEL
// calls in the reference's order on slot codes alone (integers), and each
// integral is handed to the device with the codes its slots hold and its
// frequencies.  The device then evaluates all integrals of a table, or all
// new lazy integrals of a block, at once:
//   * k_bessel: one wavefront per distinct (a, b, k) code -- j_k(w r) and
//     n_k(w r) on the grid as BESSEL forms them (series while it converges,
//     then the closed form from the first point where it does not);
//   * k_breit: one wavefront per integral -- BRRA's sum of RKINT/SKINT
//     terms, each an outward Newton-Cotes march of ZKF (four chains, one
//     lane each, in the reference's order and association) and a QUAD.
// Arithmetic follows the reference operation by operation, FP contraction
// off; sin/cos of the device library may round differently from the
// host's in the last bit, so frequency-dependent integrals agree with the
// reference to rounding, the rest bit for bit.
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <vector>

#include "gci_internal.h"
#include "radial.h"

struct gci_breit;

namespace gci {

namespace {

constexpr int KEYORB = 215;  // parameter_def KEYORB, BESSEL's packing
constexpr int BR_LDS = NMAXG + 16;

// One BRINTF evaluation: the integral and what its Bessel slots hold.
struct BrEval {
    int type, ia, ib, ic, id, k;  // 1-based orbitals, BRINTF's argument order
    int slot[4];                  // Bessel table rows of (IK,IW) = (1,1), (2,1), (1,2), (2,2); -1: zero
    double wij[2];
};

struct BesselRow {
    int a, b, k;
    double w;
};

struct BrConst {
    RkConst kc;
    int n;
    const double *r, *rp, *rpor;
    const double *P, *Q;
    int ld;
    const int *mf;
    const double *bj, *bn;  // Bessel table: row-major (N + 1) values per row
};

// ---------------------------------------------------------------- BESSEL

// bessel.f90:103-239 for one (a, b, k) row: the series in (w r)^2 while it
// converges (4 terms), then, from the first point where it does not
// (JCHAN), the closed form in sin/cos(w r).
__global__ __launch_bounds__(64) void k_bessel(int n, const double *__restrict__ r, const BesselRow *__restrict__ rows,
                                               double epsi, double *__restrict__ bj, double *__restrict__ bn) {
    const BesselRow row = rows[blockIdx.x];
    const int lane = threadIdx.x;
    double *oj = bj + (size_t)blockIdx.x * (n + 1), *on = bn + (size_t)blockIdx.x * (n + 1);
    const double w = row.w;
    if (w < epsi * epsi) {
        for (int j = lane; j <= n; j += 64) {
            oj[j] = 0.0;
            on[j] = 0.0;
        }
        return;
    }
    const int nn = row.k;
    if (lane == 0) {
        oj[0] = 0.0;
        on[0] = 0.0;
    }
    int jchan = n + 1;
    for (int base = 2; base <= n; base += 64) {
        const int j = base + lane;
        bool fail = false;
        double s1 = 0.0, s2 = 0.0;
        if (j <= n) {
            const double rw = r[j - 1] * w;
            const double wa = -0.5 * (rw * rw);
            double xb1 = 1.0, xb2 = 1.0;
            fail = true;
            for (int i = 1; i <= 4; ++i) {
                xb1 = xb1 * wa / (double)(i * (2 * (nn + i) + 1));
                xb2 = xb2 * wa / (double)(i * (2 * (i - nn) - 1));
                s1 = s1 + xb1;
                s2 = s2 + xb2;
                if (fabs(xb1) >= fabs(s1) * epsi || fabs(xb2) >= fabs(s2) * epsi) continue;
                fail = false;
                break;
            }
        }
        // the first point whose series does not converge (JCHAN); the
        // series values are stored below it only, the closed form from it on
        const unsigned long long m = __ballot(fail);
        const int first = m ? base + __ffsll((long long)m) - 1 : n + 1;
        if (j <= n && j < first) {
            oj[j - 1] = s1;
            on[j - 1] = s2;
        }
        if (m) {
            jchan = first;
            break;
        }
    }
    if (jchan > n) return;
    double dfnm = 1.0, dfn = 1.0;
    if (nn != 0) {
        for (int i = 3; i <= 2 * nn - 1; i += 2) dfnm = dfnm * (double)i;
        dfn = dfnm * (double)(2 * nn + 1);
    }
    dfnm = 1.0 / dfnm;
    double ssn, scn;
    int iswap;
    switch (nn % 4) {  // MOD(NN, 4): negative for NN = -1, the default case
        case 1: ssn = -1.0; scn = 1.0; iswap = 1; break;
        case 2: ssn = -1.0; scn = -1.0; iswap = 0; break;
        case 3: ssn = 1.0; scn = -1.0; iswap = 1; break;
        default: ssn = 1.0; scn = 1.0; iswap = 0; break;
    }
    for (int j = jchan + lane; j <= n; j += 64) {
        const double wa = w * r[j - 1];
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
        s1 = s1 * dfn / powi_d(wa, nn) - 1.0;
        s2 = s2 * powi_d(wa, nn + 1) * dfnm - 1.0;
        oj[j - 1] = s1;
        on[j - 1] = s2;
    }
}

// ------------------------------------------------------- BRRA and friends

// LDS of one integral's wavefront (1-based views: x[i - 1] = X(i))
struct BrLds {
    double ta[BR_LDS], tb[BR_LDS], rac[BR_LDS], rbd[BR_LDS], tkeep[BR_LDS], temp[BR_LDS];
    double grp[NMAXG / 4 + 8];
};

__device__ inline void wsync() { __syncthreads(); }

// ZKF (zkf.f90:33-83): TB (ZK) from TA(2:MTP); R**K for K /= 0, also K < 0
__device__ void zkf(const BrConst &c, BrLds &s, int k, int mtp) {
    const int lane = threadIdx.x, N = c.n;
    const RkConst &kc = c.kc;
    for (int i = 1 + lane; i <= mtp + 4; i += 64) {
        double v = 0.0;
        if (i >= 2 && i <= mtp) {
            const double rhop = c.rp[i - 1] * s.ta[i - 1];
            v = (k == 0) ? rhop : powi_d(c.r[i - 1], k) * rhop;
        }
        s.temp[i - 1] = v;
    }
    wsync();
    if (lane == 0) s.tb[0] = 0.0;
    if (lane >= 1 && lane <= 3) {
        const int ii = lane + 1;
        double sum = 0.0;
        for (int kk = 2; kk <= 5; ++kk) sum = sum + kc.cnc5c[ii - 2][kk - 2] * s.temp[kk - 1];
        s.tb[ii - 1] = sum;
    }
    wsync();
    if (lane < 4) {
        // the chain ii = lane + 5, lane + 9, ... in the reference's order
        for (int ii = lane + 5; ii <= mtp + 3; ii += 4)
            s.tb[ii - 1] = s.tb[ii - 5] + kc.c1 * (s.temp[ii - 5] + s.temp[ii - 1]) +
                           kc.c2 * (s.temp[ii - 4] + s.temp[ii - 2]) + kc.c3 * s.temp[ii - 3];
    }
    wsync();
    const double zklim = s.tb[mtp + 2];
    for (int i = 2 + lane; i <= N; i += 64) {
        if (k == 0) {
            if (i >= mtp + 4) s.tb[i - 1] = zklim;
        } else {
            const double rt = powi_d(c.r[i - 1], k);
            s.tb[i - 1] = (i <= mtp + 3) ? s.tb[i - 1] / rt : zklim / rt;
        }
    }
    wsync();
}

__device__ inline double quad_ta(const BrConst &c, BrLds &s, int mtp) {
    // QUAD zeroes TA(MTP+1:MTP+3) (quad.f90:80-82)
    if (threadIdx.x < 3) s.ta[mtp + threadIdx.x] = 0.0;
    wsync();
    const double v = wave_quad(s.ta, c.r, c.rp, mtp, c.kc, s.grp);
    wsync();
    return v;
}

__device__ inline const double *slot_j(const BrConst &c, const BrEval &e, int ik, int iw) {
    const int row = e.slot[(ik - 1) + 2 * (iw - 1)];
    return row < 0 ? nullptr : c.bj + (size_t)row * (c.n + 1);
}
__device__ inline const double *slot_n(const BrConst &c, const BrEval &e, int ik, int iw) {
    const int row = e.slot[(ik - 1) + 2 * (iw - 1)];
    return row < 0 ? nullptr : c.bn + (size_t)row * (c.n + 1);
}
__device__ inline double at(const double *a, int i) { return a ? a[i - 1] : 0.0; }

// RKINT (rkint.f90:43-81); rac/rbd are LDS arrays of BRRA
__device__ double rkint(const BrConst &c, const BrEval &e, BrLds &s, const double *rac, int ia, int ic,
                        const double *rbd, int ib, int id, int k, int iw) {
    const int lane = threadIdx.x;
    const int mxrbd = min(c.mf[ib - 1], c.mf[id - 1]);
    const int mxrac = min(c.mf[ia - 1], c.mf[ic - 1]);
    const double *bj = iw ? slot_j(c, e, 1, iw) : nullptr, *bn = iw ? slot_n(c, e, 1, iw) : nullptr;
    for (int i = 1 + lane; i <= mxrac; i += 64) s.ta[i - 1] = iw == 0 ? rac[i - 1] : rac[i - 1] * (1.0 + at(bj, i));
    wsync();
    zkf(c, s, k, mxrac);
    const int mtp = min(mxrac, mxrbd);
    for (int i = 1 + lane; i <= mtp; i += 64) {
        double v = 0.0;
        if (i >= 2) v = iw == 0 ? rbd[i - 1] * s.tb[i - 1] * c.rpor[i - 1]
                                : rbd[i - 1] * (1.0 + at(bn, i)) * s.tb[i - 1] * c.rpor[i - 1];
        s.ta[i - 1] = v;
    }
    wsync();
    return quad_ta(c, s, mtp);
}

// SKINT (skint.f90:44-109)
__device__ double skint(const BrConst &c, const BrEval &e, BrLds &s, const double *rac, int ia, int ic,
                        const double *rbd, int ib, int id, int k, int iw) {
    const int lane = threadIdx.x;
    const double w = e.wij[iw - 1];
    const double wk = (double)(k + k + 1);
    const int mxrac = min(c.mf[ia - 1], c.mf[ic - 1]);
    const int mxrbd = min(c.mf[ib - 1], c.mf[id - 1]);
    for (int i = 1 + lane; i <= mxrac; i += 64) s.ta[i - 1] = rac[i - 1];
    wsync();
    int mtp = mxrac;
    zkf(c, s, k - 1, mtp);
    if (fabs(w) < 1.0e-10) {
        for (int i = 1 + lane; i <= mtp; i += 64) s.tkeep[i - 1] = s.tb[i - 1];
        wsync();
        zkf(c, s, k + 1, mtp);
        mtp = min(mtp, mxrbd);
        for (int i = 1 + lane; i <= mtp; i += 64)
            s.ta[i - 1] = i == 1 ? 0.0 : rbd[i - 1] * c.rpor[i - 1] * (s.tb[i - 1] - s.tkeep[i - 1]);
        wsync();
        const double valu = quad_ta(c, s, mtp);
        return wk * valu * 0.5;
    }
    const double *bj1 = slot_j(c, e, 1, iw), *bn2 = slot_n(c, e, 2, iw);
    const double *bj2 = slot_j(c, e, 2, iw), *bn1 = slot_n(c, e, 1, iw);
    for (int i = 1 + lane; i <= mtp; i += 64) {
        s.tkeep[i - 1] = s.tb[i - 1];
        s.ta[i - 1] = -s.ta[i - 1] * at(bj1, i);
    }
    wsync();
    zkf(c, s, k - 1, mtp);
    mtp = min(mtp, mxrbd);
    for (int i = 1 + lane; i <= mtp; i += 64) {
        double v = 0.0;
        if (i >= 2) {
            const double r = c.r[i - 1];
            v = ((1.0 + at(bn2, i)) * s.tb[i - 1] - s.tkeep[i - 1] * at(bn2, i)) * rbd[i - 1] / (r * r) *
                c.rpor[i - 1];
        }
        s.ta[i - 1] = v;
    }
    wsync();
    double valu = quad_ta(c, s, mtp);
    double res = (wk / w) * (wk / w) * valu;
    for (int i = 1 + lane; i <= mxrbd; i += 64) s.ta[i - 1] = rbd[i - 1] * (1.0 + at(bj2, i));
    wsync();
    zkf(c, s, k + 1, mxrbd);
    mtp = min(mxrbd, mxrac);
    for (int i = 1 + lane; i <= mtp; i += 64)
        s.ta[i - 1] = i == 1 ? 0.0 : rac[i - 1] * (1.0 + at(bn1, i)) * s.tb[i - 1] * c.r[i - 1] * c.rp[i - 1];
    wsync();
    valu = quad_ta(c, s, mtp);
    return res - valu * w * w / (double)((2 * k + 3) * (2 * k - 1));
}

// BRRA (brra.f90:41-124) in its own argument order (IA, IC, IB, ID)
__device__ double brra(const BrConst &c, const BrEval &e, BrLds &s, int itype, int ia, int ic, int ib, int id, int k) {
    const int lane = threadIdx.x;
    const int mxrac = min(c.mf[ia - 1], c.mf[ic - 1]), mxrbd = min(c.mf[ib - 1], c.mf[id - 1]);
    const double *Pa = c.P + (size_t)(ia - 1) * c.ld, *Qc = c.Q + (size_t)(ic - 1) * c.ld;
    const double *Pb = c.P + (size_t)(ib - 1) * c.ld, *Qd = c.Q + (size_t)(id - 1) * c.ld;
    for (int i = 1 + lane; i <= mxrac; i += 64) s.rac[i - 1] = Pa[i - 1] * Qc[i - 1];
    for (int i = 1 + lane; i <= mxrbd; i += 64) s.rbd[i - 1] = Pb[i - 1] * Qd[i - 1];
    wsync();
    const double *rac = s.rac, *rbd = s.rbd;
    auto rk = [&](bool swap, int iw) {
        return swap ? rkint(c, e, s, rbd, ib, id, rac, ia, ic, k, iw) : rkint(c, e, s, rac, ia, ic, rbd, ib, id, k, iw);
    };
    switch (itype) {
        case 1:
            if (ia == ib && ic == id) return 2.0 * rk(false, 1);
            if (ia == id && ic == ib) return rk(false, 1) + rk(true, 1);
            return (rk(false, 1) + rk(false, 2) + rk(true, 1) + rk(true, 2)) * 0.5;
        case 2:
            if ((ia == ib && ic == id) || (ia == id && ic == ib)) return skint(c, e, s, rac, ia, ic, rbd, ib, id, k, 1);
            return (skint(c, e, s, rac, ia, ic, rbd, ib, id, k, 1) + skint(c, e, s, rac, ia, ic, rbd, ib, id, k, 2)) *
                   0.5;
        case 3: {
            if (ia == ic) {
                const double *Pd = c.P + (size_t)(id - 1) * c.ld, *Qb = c.Q + (size_t)(ib - 1) * c.ld;
                for (int i = 1 + lane; i <= mxrbd; i += 64) s.rbd[i - 1] = s.rbd[i - 1] + Pd[i - 1] * Qb[i - 1];
                wsync();
                return (rk(false, 0) + rk(true, 0) + rk(false, 2) + rk(true, 2)) * 0.5;
            }
            const double *Pc = c.P + (size_t)(ic - 1) * c.ld, *Qa = c.Q + (size_t)(ia - 1) * c.ld;
            for (int i = 1 + lane; i <= mxrac; i += 64) s.rac[i - 1] = s.rac[i - 1] + Pc[i - 1] * Qa[i - 1];
            wsync();
            return (rk(false, 1) + rk(true, 1) + rk(false, 0) + rk(true, 0)) * 0.5;
        }
        case 4: return rk(false, 0) + rk(true, 0);
        case 5:
            if (ia == id && ic == ib) return rk(false, 1) + rk(true, 1);
            return 2.0 * rk(false, 1);
        default: return skint(c, e, s, rac, ia, ic, rbd, ib, id, k, 1);
    }
}

__global__ __launch_bounds__(64) void k_breit(BrConst c, const BrEval *__restrict__ evals, int64_t n,
                                              double *__restrict__ out) {
    __shared__ BrLds s;
    const int64_t q = blockIdx.x;
    if (q >= n) return;
    const BrEval e = evals[q];
    // BRINTF passes (IA, IB, IC, ID) to BRRA's (IA, IC, IB, ID)
    const double v = brra(c, e, s, e.type, e.ia, e.ib, e.ic, e.id, e.k);
    if (threadIdx.x == 0) out[q] = v;
}

// --------------------------------------------------- host-side replay

bool triangbreit(int type, const int *nkl, const int *nkj, int a, int b, int c, int d, int k) {
    const int la = nkl[a - 1], lb = nkl[b - 1], lc = nkl[c - 1], ld = nkl[d - 1];
    const int ja = nkj[a - 1], jb = nkj[b - 1], jc = nkj[c - 1], jd = nkj[d - 1];
    if (type == 1) {  // rci90/triangbreit1.f90
        for (int l = k - 1; l <= k + 1; ++l)
            if ((k + la + lb) % 2 == 1 && std::abs(ja - jb) <= 2 * l && std::abs(ja + jb) >= 2 * l &&
                (k + lc + ld) % 2 == 1 && std::abs(jc - jd) <= 2 * l && std::abs(jc + jd) >= 2 * l)
                return true;
        return false;
    }
    const int l = k;  // rci90/triangbreit2.f90
    return (l + la + lb) % 2 == 0 && std::abs(ja - jb) <= 2 * l && std::abs(ja + jb) >= 2 * l &&
           (l + lc + ld) % 2 == 0 && std::abs(jc - jd) <= 2 * l && std::abs(jc + jd) >= 2 * l;
}

}  // namespace

}  // namespace gci

using namespace gci;

struct gci_breit {
    // grid and orbitals on the device
    int n = 0, nw = 0, ld = 0;
    RkConst kc{};
    DBuf<double> r, rp, rpor, P, Q;
    DBuf<int> mf;
    std::vector<int> mf_h, nkl_h, nkj_h;
    std::vector<double> e_h;
    double c = 0.0, wfact = 0.0, epsi = 0.0;
    // BESSEL's bookkeeping (codes only) and the rows of the Bessel table
    int keep[2][2] = {{0, 0}, {0, 0}};
    double wij[2] = {0.0, 0.0};
    std::map<int, int> row_of;  // code -> row
    std::vector<BesselRow> rows;
    DBuf<double> bj, bn;
    int rows_on_device = 0;
    // GENINTBREIT1/2 tables
    int j2max = 0;
    std::vector<int32_t> kst[2], keys[2];
    std::vector<double> vals[2];
    // BRINT3..6: key -> value, first computed value kept
    std::map<int64_t, double> memo[4];

    double freq(int a, int b) const { return wfact * std::fabs(e_h[a - 1] - e_h[b - 1]) / c; }

    // BESSEL(IA, IB, IK, IW, K)'s effect on KEEP and WIJ (bessel.f90:78-101)
    void bessel(int ia, int ib, int ik, int iw, int k) {
        const int icode = std::max(ia, ib) + KEYORB * (std::min(ia, ib) + KEYORB * k);
        if (icode == keep[ik - 1][iw - 1]) return;
        wij[iw - 1] = freq(ia, ib);
        keep[ik - 1][iw - 1] = icode;
        if (!row_of.count(icode)) {
            row_of[icode] = (int)rows.size();
            rows.push_back(BesselRow{std::max(ia, ib), std::min(ia, ib), k, freq(ia, ib)});
        }
    }

    // BRINTF's BESSEL calls (brintf.f90:44-92), then the evaluation record
    BrEval brintf(int type, int ia, int ib, int ic, int id, int k) {
        const int key2 = KEYORB * KEYORB;
        switch (type) {
            case 1:
                if (ia != ic || ib != id || ia != id || ic == ib) bessel(ic, id, 1, 2, k);
                bessel(ia, ib, 1, 1, k);
                break;
            case 5: bessel(ia, ib, 1, 1, k); break;
            case 3:
                if (ia != ib) bessel(ia, ib, 1, 1, k);
                if (ic != id) bessel(ic, id, 1, 2, k);
                break;
            case 2:
                if (ia != ic || ib != id || ia != id || ic != ib) {
                    int icod = std::max(ic, id) + KEYORB * std::min(ic, id);
                    const int icod1 = icod + key2 * (k - 1), icod2 = icod + key2 * (k + 1);
                    icod = std::max(ia, ib) + KEYORB * std::min(ia, ib);
                    const int icod3 = icod + key2 * (k - 1), icod4 = icod + key2 * (k + 1);
                    if (icod1 == keep[0][1] && icod2 == keep[1][1] && icod3 == keep[0][0] && icod4 == keep[1][0])
                        break;
                    if (icod1 == keep[0][0] && icod2 == keep[1][0] && icod3 == keep[0][1] && icod4 == keep[1][1])
                        break;
                    bessel(ic, id, 1, 2, k - 1);
                    bessel(ic, id, 2, 2, k + 1);
                }
                bessel(ia, ib, 1, 1, k - 1);
                bessel(ia, ib, 2, 1, k + 1);
                break;
            case 6:
                bessel(ia, ib, 1, 1, k - 1);
                bessel(ia, ib, 2, 1, k + 1);
                break;
            default: break;
        }
        BrEval e{type, ia, ib, ic, id, k, {-1, -1, -1, -1}, {wij[0], wij[1]}};
        for (int iw = 0; iw < 2; ++iw)
            for (int ik = 0; ik < 2; ++ik) e.slot[ik + 2 * iw] = keep[ik][iw] ? row_of.at(keep[ik][iw]) : -1;
        return e;
    }

    // evaluates a batch of BRINTF records on the device
    std::vector<double> evaluate(const std::vector<BrEval> &ev) {
        std::vector<double> out(ev.size());
        if (ev.empty()) return out;
        if ((int)rows.size() > rows_on_device) {
            // the Bessel table grows: rows computed so far are kept
            const size_t rl = (size_t)n + 1;
            DBuf<double> nj(rows.size() * rl), nn(rows.size() * rl);
            if (rows_on_device) {
                GCI_HIP(hipMemcpy(nj.p, bj.p, rows_on_device * rl * sizeof(double), hipMemcpyDeviceToDevice));
                GCI_HIP(hipMemcpy(nn.p, bn.p, rows_on_device * rl * sizeof(double), hipMemcpyDeviceToDevice));
            }
            const int add = (int)rows.size() - rows_on_device;
            DBuf<BesselRow> drows(add);
            GCI_HIP(hipMemcpy(drows.p, rows.data() + rows_on_device, add * sizeof(BesselRow), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_bessel, dim3(add), dim3(64), 0, 0, n, r.p, drows.p, epsi, nj.p + rows_on_device * rl,
                               nn.p + rows_on_device * rl);
            GCI_HIP(hipGetLastError());
            GCI_HIP(hipDeviceSynchronize());
            bj = std::move(nj);
            bn = std::move(nn);
            rows_on_device = (int)rows.size();
        }
        BrConst cst{kc, n, r.p, rp.p, rpor.p, P.p, Q.p, ld, mf.p, bj.p, bn.p};
        const int64_t chunk = 1 << 20;
        DBuf<BrEval> dev(std::min<int64_t>(chunk, (int64_t)ev.size()));
        DBuf<double> dout(std::min<int64_t>(chunk, (int64_t)ev.size()));
        for (int64_t a = 0; a < (int64_t)ev.size(); a += chunk) {
            const int64_t m = std::min<int64_t>(chunk, (int64_t)ev.size() - a);
            GCI_HIP(hipMemcpy(dev.p, ev.data() + a, m * sizeof(BrEval), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_breit, dim3((unsigned)m), dim3(64), 0, 0, cst, dev.p, m, dout.p);
            GCI_HIP(hipGetLastError());
            GCI_HIP(hipMemcpy(out.data() + a, dout.p, m * sizeof(double), hipMemcpyDeviceToHost));
        }
        return out;
    }

    // GENINTBREIT1 / GENINTBREIT2 (genintbreit1.f90:45-104)
    void genint(int type) {
        const int t = type - 1, key = nw + 1;
        kst[t].assign(j2max + 2, 0);
        kst[t][0] = 1;
        std::vector<BrEval> ev;
        keys[t].clear();
        for (int k = 0; k <= j2max; ++k) {
            for (int ia = 1; ia <= nw; ++ia)
                for (int ib = 1; ib <= nw; ++ib)
                    for (int ic = 1; ic <= nw; ++ic)
                        for (int id = 1; id <= nw; ++id) {
                            if (!triangbreit(type, nkl_h.data(), nkj_h.data(), ia, ib, ic, id, k)) continue;
                            keys[t].push_back(((ia * key + ib) * key + ic) * key + id);
                            ev.push_back(brintf(type, ia, ib, ic, id, k));
                        }
            kst[t][k + 1] = (int32_t)keys[t].size() + 1;
        }
        vals[t] = evaluate(ev);
    }

    // BRINT1/2 (brint1.f90:27-67): false where the reference STOPs
    bool table(int type, int ia, int ib, int ic, int id, int k, double &out) const {
        const int t = type - 1, key = nw + 1;
        if (k < 0 || k > j2max) return false;
        const int index = ((ia * key + ib) * key + ic) * key + id;
        int jl = kst[t][k], ju = kst[t][k + 1] - 1;
        if (ju < jl || index < keys[t][jl - 1] || index > keys[t][ju - 1]) return false;
        while (ju - jl > 1) {
            const int jm = (ju + jl) / 2;
            if (keys[t][jm - 1] > index) ju = jm;
            else jl = jm;
        }
        int loc;
        if (index == keys[t][ju - 1]) loc = ju;
        else if (index == keys[t][jl - 1]) loc = jl;
        else return false;
        out = vals[t][loc - 1];
        return true;
    }
};

namespace gci {

// The Breit terms of a block's pairs in SETHAM's order
// (setham_gg.f90:298-357): the value of every term with |COEFF| > 1e-12
// (NaN marks a skipped term), types 1/2 from the tables, 3..6 from the memo
// or evaluated on first use -- all new ones of the block in one batch, the
// BESSEL replay in list order.  Pairs skipped by LFORDR are not visited.
void breit_term_values(gci_breit *b, int64_t npairs, const int32_t *pic, const int32_t *pir, int lfordr, int iccut,
                       const gci_breit_lists *bl, std::vector<double> &tv) {
    constexpr double CUTOFF = 1.0e-12;
    tv.assign((size_t)bl->nterms, std::nan(""));
    std::vector<BrEval> fresh;
    std::vector<std::pair<int, int64_t>> fresh_key;  // (type slot, memo key)
    std::vector<std::pair<int64_t, size_t>> pending;  // term -> index in fresh
    std::map<int64_t, size_t> fresh_of[4];
    const int key = b->nw + 1;
    for (int64_t p = 0; p < npairs; ++p) {
        if (lfordr && pir[p] > iccut && pir[p] != pic[p]) continue;
        for (int64_t q = bl->bt_beg[p]; q < bl->bt_end[p]; ++q) {
            if (!(std::fabs(bl->coeff[q]) > CUTOFF)) continue;
            const int32_t *l = bl->label + 6 * q;
            const int type = std::abs(l[5]);
            if (type <= 2) {
                double v;
                if (!b->table(type, l[0], l[1], l[2], l[3], l[4], v))
                    throw Error{GCI_ERR_NOTFOUND, "Breit integral of type " + std::to_string(type) + " not found"};
                tv[q] = v;
                continue;
            }
            const int t = type - 3;
            const int64_t mkey = ((((int64_t)l[4] * key + l[0]) * key + l[1]) * key + l[2]) * key + l[3];
            auto it = b->memo[t].find(mkey);
            if (it != b->memo[t].end()) {
                tv[q] = it->second;
                continue;
            }
            auto jt = fresh_of[t].find(mkey);
            if (jt == fresh_of[t].end()) {
                jt = fresh_of[t].emplace(mkey, fresh.size()).first;
                fresh.push_back(b->brintf(type, l[0], l[1], l[2], l[3], l[4]));
                fresh_key.emplace_back(t, mkey);
            }
            pending.emplace_back(q, jt->second);
        }
    }
    const std::vector<double> v = b->evaluate(fresh);
    for (size_t i = 0; i < fresh.size(); ++i) b->memo[fresh_key[i].first][fresh_key[i].second] = v[i];
    for (auto &pq : pending) tv[pq.first] = v[pq.second];
}

}  // namespace gci

extern "C" {

int gci_breit_create(const gci_grid *grid, const gci_orbitals *orb, const double *e, double c, double wfact,
                     gci_breit **out) {
    return guarded([&] {
        GCI_REQUIRE(grid && orb && e && out, "gci_breit_create: null argument");
        GCI_REQUIRE(grid->n >= 14 && grid->n + 10 <= NMAXG, "gci_breit_create: n must be in 14..590");
        GCI_REQUIRE(orb->nw >= 1 && orb->nw <= 127 && orb->ld >= grid->n, "gci_breit_create: bad nw/ld");
        GCI_REQUIRE(c > 0.0, "gci_breit_create: c must be positive");
        for (int j = 0; j < orb->nw; ++j)
            GCI_REQUIRE(orb->mf[j] >= 2 && orb->mf[j] <= grid->n - 4, "gci_breit_create: MF out of range (2..N-4)");
        auto b = std::make_unique<gci_breit>();
        b->n = grid->n;
        b->nw = orb->nw;
        b->ld = orb->ld;
        b->kc = make_const(grid);
        b->c = c;
        b->wfact = wfact;
        b->epsi = std::sqrt(0.1 * grid->accy);  // bessel.f90:76
        const size_t n10 = (size_t)grid->n + 10;
        b->r.alloc(n10);
        b->rp.alloc(n10);
        b->rpor.alloc(n10);
        GCI_HIP(hipMemcpy(b->r.p, grid->r, n10 * sizeof(double), hipMemcpyHostToDevice));
        GCI_HIP(hipMemcpy(b->rp.p, grid->rp, n10 * sizeof(double), hipMemcpyHostToDevice));
        GCI_HIP(hipMemcpy(b->rpor.p, grid->rpor, n10 * sizeof(double), hipMemcpyHostToDevice));
        const size_t po = (size_t)orb->nw * orb->ld;
        b->P.alloc(po);
        b->Q.alloc(po);
        GCI_HIP(hipMemcpy(b->P.p, orb->P, po * sizeof(double), hipMemcpyHostToDevice));
        GCI_HIP(hipMemcpy(b->Q.p, orb->Q, po * sizeof(double), hipMemcpyHostToDevice));
        b->mf.alloc(orb->nw);
        GCI_HIP(hipMemcpy(b->mf.p, orb->mf, orb->nw * sizeof(int), hipMemcpyHostToDevice));
        b->mf_h.assign(orb->mf, orb->mf + orb->nw);
        b->nkl_h.assign(orb->nkl, orb->nkl + orb->nw);
        b->nkj_h.assign(orb->nkj, orb->nkj + orb->nw);
        b->e_h.assign(e, e + orb->nw);
        b->j2max = *std::max_element(b->nkj_h.begin(), b->nkj_h.end());
        GCI_REQUIRE(b->j2max <= 20, "genintbreit: KMAX too small");
        // rci90/rci92.f90:198-202: type 1, then type 2
        b->genint(1);
        b->genint(2);
        *out = b.release();
    });
}

int gci_breit_table(const gci_breit *b, int type, int32_t *j2max, int32_t *kstart, int32_t *keys, double *vals,
                    int64_t cap, int64_t *n) {
    return guarded([&] {
        GCI_REQUIRE(b && n && (type == 1 || type == 2), "gci_breit_table: bad argument");
        const int t = type - 1;
        *n = (int64_t)b->keys[t].size();
        if (j2max) *j2max = b->j2max;
        if (kstart) std::copy(b->kst[t].begin(), b->kst[t].end(), kstart);
        if (!keys && !vals) return;
        GCI_REQUIRE(cap >= *n, "gci_breit_table: capacity too small");
        if (keys) std::copy(b->keys[t].begin(), b->keys[t].end(), keys);
        if (vals) std::copy(b->vals[t].begin(), b->vals[t].end(), vals);
    });
}

int gci_breit_integral(gci_breit *b, int type, int ia, int ib, int ic, int id, int k, double *out) {
    return guarded([&] {
        GCI_REQUIRE(b && out && type >= 1 && type <= 6, "gci_breit_integral: bad argument");
        GCI_REQUIRE(ia >= 1 && ia <= b->nw && ib >= 1 && ib <= b->nw && ic >= 1 && ic <= b->nw && id >= 1 &&
                        id <= b->nw && k >= 0 && k <= 20,
                    "gci_breit_integral: label out of range");
        const int32_t pic = 1, pir = 1;
        const int64_t beg = 0, end = 1;
        const int32_t lab[6] = {ia, ib, ic, id, k, type};
        const double one = 1.0;
        gci_breit_lists bl{&beg, &end, lab, &one, 1};
        std::vector<double> tv;
        breit_term_values(b, 1, &pic, &pir, 0, 0, &bl, tv);
        *out = tv[0];
    });
}

int gci_breit_free(gci_breit *b) {
    return guarded([&] { delete b; });
}

}  // extern "C"
