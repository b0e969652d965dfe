// rk.hip -- R^k Slater-integral tables and one-body Dirac integrals.
//
// Replaces GENINTRK (rci90/genintrk.f90:40-86) with its SLATER/YZK/QUAD
// calls (lib9290/slater.f90, yzk.f90, quad.f90), and RINTI/DPBDT
// (lib9290/rinti.f90, dpbdt.f90) as IABINT memoises them.
//
// The host enumerates the table exactly as GENINTRK does (same loop order,
// key packing and KSTART).  Entries sharing a Hartree function Y^k(b,d) are
// grouped; one workgroup per (k, b, d):
//   * wave 0 builds Y^k on the exponential grid in FP64: the Newton-Cotes
//     marches of YZK are four interleaved linear recurrences, one lane per
//     chain in the reference's order (outward for Z^k, inward for W^k),
//     with YZK's K = 0 asymptotic correction applied per chain;
//   * the densities P_b P_d + Q_b Q_d and Y^k live in LDS;
//   * every wave then integrates entries (a, c) against Y^k with QUAD's
//     power-law head and strided 5-point Newton-Cotes sum, reduced across
//     the wave.
// Elementwise arithmetic follows the reference operation by operation with
// floating-point contraction disabled, and the recurrences and QUAD's sum
// run in the reference's order, so a table entry differs from the
// reference's only where the device log() (QUAD's power-law exponent)
// rounds differently from the host's.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

#include "gci_internal.h"
#include "radial.h"

#pragma clang fp contract(off)

namespace gci {

namespace {

struct YDesc {
    int k, b, d;          // 0-based orbitals
    int e0, e1;           // entries [e0, e1) in the grouped order
};

__global__ __launch_bounds__(256) void k_rk_table(RkConst kc, const YDesc *__restrict__ ys,
                                                  const int *__restrict__ ent_ac, const int *__restrict__ ent_out,
                                                  const double *__restrict__ P, const double *__restrict__ Q, int ld,
                                                  const int *__restrict__ mf, const double *__restrict__ gr,
                                                  const double *__restrict__ grp, const double *__restrict__ grpor,
                                                  double *__restrict__ vals, double *__restrict__ yk_out,
                                                  int *__restrict__ mtp_out) {
    __shared__ double r[NMAXG], rp[NMAXG], rpor[NMAXG];
    __shared__ double yk[NMAXG];
    __shared__ double temp[NMAXG + 8];   // r^k RHOP, the outward march's integrand
    __shared__ double tempw[NMAXG + 8];  // RHOP / r^(k+1), the inward march's
    __shared__ double zk[NMAXG + 8];
    __shared__ double wk[NMAXG + 8];
    __shared__ double tas[4][NMAXG + 8];
    __shared__ double grps[4][NMAXG / 4 + 8];  // QUAD's group terms, per wave
    __shared__ int ymtp;
    const YDesc y = ys[blockIdx.x];
    const int n = kc.n;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    for (int i = t; i < n + 10 && i < NMAXG; i += blockDim.x) {
        r[i] = gr[i];
        rp[i] = grp[i];
        rpor[i] = grpor[i];
    }
    __syncthreads();
    const int K = y.k;
    const double *Pb = P + (size_t)y.b * ld, *Pd = P + (size_t)y.d * ld;
    const double *Qb = Q + (size_t)y.b * ld, *Qd = Q + (size_t)y.d * ld;
    const int mtp = min(mf[y.b], mf[y.d]);
    // ---- YZK (yzk.f90:99-223); 1-based index ii lives at [ii-1]
    // TEMP = r^k * RP (Pb Pd + Qb Qd) on 2..MTP and, for the inward march,
    // RHOP / r^(k+1) (yzk.f90:197-202); zero on 1 and MTP+1..N+4
    for (int ii = 1 + t; ii <= n + 4; ii += blockDim.x) {
        double v = 0.0, vw = 0.0;
        if (ii >= 2 && ii <= mtp) {
            const double rhop = rp[ii - 1] * (Pb[ii - 1] * Pd[ii - 1] + Qb[ii - 1] * Qd[ii - 1]);
            v = (K == 0) ? rhop : ipow_d(r[ii - 1], K) * rhop;
            vw = (K == 0) ? rhop * (1.0 / r[ii - 1]) : rhop * (1.0 / (ipow_d(r[ii - 1], K) * r[ii - 1]));
        }
        temp[ii - 1] = v;
        tempw[ii - 1] = vw;
    }
    __syncthreads();
    // The outward (Z^k) and inward (W^k) marches are independent: wave 0
    // runs the first, wave 1 the second, at the same time.
    if (wave == 0 && lane < 4) {
        // Lane c runs the chain of ii = c+1, c+5, c+9, ... in YZK's order:
        // its start ZK(c+1) (ZK(1) = 0, ZK(2..4) by the semi-open rule,
        // yzk.f90:128-135), then the closed march ZK(ii) = ZK(ii-4) + RTMP
        // for ii = 5..MTP+3 (yzk.f90:138-144).  Sequential chains give the
        // reference's rounding: near the origin Y^k can be pure rounding
        // noise (a density whose inward moment cancels), and QUAD's
        // power-law head then starts where that noise says, which fixes the
        // Newton-Cotes panel alignment of the whole integral.
        const int c = lane;
        if (c == 0) {
            zk[0] = 0.0;
        } else {
            double sum = 0.0;
            for (int kk = 2; kk <= 5; ++kk) sum = sum + kc.cnc5c[c - 1][kk - 2] * temp[kk - 1];
            zk[c] = sum;
        }
        for (int ii = c + 5; ii <= mtp + 3; ii += 4) {
            const double rtmp = kc.c1 * (temp[ii - 5] + temp[ii - 1]) + kc.c2 * (temp[ii - 4] + temp[ii - 2]) +
                                kc.c3 * temp[ii - 3];
            zk[ii - 1] = zk[ii - 5] + rtmp;
        }
    }
    // inward march WK(ii) = WK(ii+4) + (...), ii = MTP..2, WK(MTP+1..) = 0
    // (yzk.f90:205-209): lane c of wave 1 runs the chain ii = MTP-c,
    // MTP-c-4, ... in the reference's order and association
    if (wave == 1 && lane < 4) {
        const int c = lane;
        wk[mtp + 3 - c] = 0.0;  // WK(MTP+4-c): the chain's first WK(ii+4), so a lane touches its own chain only
        for (int ii = mtp - c; ii >= 2; ii -= 4) {
            double w = wk[ii + 3] + kc.c1 * (tempw[ii - 1] + tempw[ii + 3]);
            w = w + kc.c2 * (tempw[ii] + tempw[ii + 2]);
            w = w + kc.c3 * tempw[ii + 1];
            wk[ii - 1] = w;
        }
    }
    __syncthreads();
    // asymptotic value (yzk.f90:154-173) and tabulation (yzk.f90:177-187)
    __shared__ double zklim_s;
    __shared__ double dif_s[4];
    if (t == 0) {
        if (K == 0) {
            const double zklim = (y.b == y.d) ? 1.0 : 0.0;
            for (int kk = mtp + 3; kk >= mtp; --kk) {
                const double dif = zk[kk - 1] - zklim;
                dif_s[(kk - 1) % 4] = (fabs(dif) <= kc.accy) ? 0.0 : dif;
            }
            zklim_s = zklim;
        } else {
            zklim_s = zk[mtp + 3 - 1];
            for (int c = 0; c < 4; ++c) dif_s[c] = 0.0;
        }
    }
    __syncthreads();
    for (int ii = 2 + t; ii <= n; ii += blockDim.x) {
        double z;
        if (ii <= mtp + 3) {
            z = zk[ii - 1];
            if (K == 0) {
                const double dif = dif_s[(ii - 1) % 4];
                if (dif != 0.0) z = z - dif;
            } else {
                z = z * (1.0 / ipow_d(r[ii - 1], K));
            }
        } else {
            z = (K == 0) ? zklim_s : zklim_s * (1.0 / ipow_d(r[ii - 1], K));
        }
        zk[ii - 1] = z;
    }
    __syncthreads();
    // YK = ZK + WK r^(k+1) (yzk.f90:214-223)
    for (int ii = 1 + t; ii <= n; ii += blockDim.x) {
        double w = 0.0;
        if (ii >= 2 && ii <= mtp)
            w = wk[ii - 1] * ((K == 0) ? r[ii - 1] : ipow_d(r[ii - 1], K) * r[ii - 1]);
        yk[ii - 1] = (ii == 1) ? 0.0 : zk[ii - 1] + w;
    }
    if (t == 0) ymtp = mtp;
    __syncthreads();
    // the Hartree function itself, TB(1:N) and MTP as YZK leaves them
    if (yk_out) {
        for (int ii = t; ii < n; ii += blockDim.x) yk_out[(size_t)blockIdx.x * n + ii] = yk[ii];
        if (t == 0) mtp_out[blockIdx.x] = mtp;
    }
    // ---- SLATER (slater.f90:278-294) for every (a, c) sharing this Y^k
    double *ta = tas[wave];
    for (int e = y.e0 + wave; e < y.e1; e += blockDim.x / 64) {
        const int a = ent_ac[2 * e], c = ent_ac[2 * e + 1];
        int m;
        if (K == 0 && y.b == y.d) m = min(mf[a], mf[c]);
        else m = min(min(ymtp, mf[a]), mf[c]);
        const double *Pa = P + (size_t)a * ld, *Pc = P + (size_t)c * ld;
        const double *Qa = Q + (size_t)a * ld, *Qc = Q + (size_t)c * ld;
        for (int i = 1 + lane; i <= m + 3; i += 64) {
            double v = 0.0;
            if (i >= 2 && i <= m) v = (Pa[i - 1] * Pc[i - 1] + Qa[i - 1] * Qc[i - 1]) * rpor[i - 1] * yk[i - 1];
            ta[i - 1] = v;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const double res = wave_quad(ta, r, rp, m, kc, grps[wave]);
        if (lane == 0) vals[ent_out[e]] = res;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// ---- RINTI (rinti.f90:49-103) with DPBDT (dpbdt.f90), one wave per pair
__global__ __launch_bounds__(64) void k_rinti(RkConst kc, const int *__restrict__ pairs, int npairs,
                                              const double *__restrict__ a13, double h, double cl,
                                              const double *__restrict__ zz, const double *__restrict__ P,
                                              const double *__restrict__ Q, int ld, const int *__restrict__ mf,
                                              const int *__restrict__ nak, const double *__restrict__ gr,
                                              const double *__restrict__ grp, const double *__restrict__ grpor,
                                              double *__restrict__ out, int nw) {
    __shared__ double r[NMAXG], rp[NMAXG], rpor[NMAXG];
    __shared__ double dp[NMAXG], dq[NMAXG], ta[NMAXG + 8];
    __shared__ double qgrp[NMAXG / 4 + 8];  // QUAD's group terms
    __shared__ double A[13][13];
    const int pi = blockIdx.x;
    if (pi >= npairs) return;
    const int j = pairs[2 * pi], k = pairs[2 * pi + 1];  // 0-based, RINTI(J, K)
    const int lane = threadIdx.x;
    const int n = kc.n;
    for (int i = lane; i < n + 10; i += 64) {
        r[i] = gr[i];
        rp[i] = grp[i];
        rpor[i] = grpor[i];
    }
    for (int i = lane; i < 169; i += 64) A[i / 13][i % 13] = a13[i];
    __syncthreads();
    const double *Pj = P + (size_t)j * ld, *Qj = Q + (size_t)j * ld;
    const double *Pk = P + (size_t)k * ld, *Qk = Q + (size_t)k * ld;
    // DPBDT(K): 13-point derivative
    for (int i = 1 + lane; i <= n; i += 64) {
        double hp = 0.0, hq = 0.0;
        if (i <= 6 || i >= n - 5) {
            const int row = (i <= 6) ? i : i - n + 13;
            const int base = (i <= 6) ? 0 : n - 13;
            for (int c = 1; c <= 13; ++c) {
                hp += A[row - 1][c - 1] * Pk[base + c - 1];
                hq += A[row - 1][c - 1] * Qk[base + c - 1];
            }
        } else {
            hp = A[6][0] * (Pk[i - 7] - Pk[i + 5]) + A[6][1] * (Pk[i - 6] - Pk[i + 4]) +
                 A[6][2] * (Pk[i - 5] - Pk[i + 3]) + A[6][3] * (Pk[i - 4] - Pk[i + 2]) +
                 A[6][4] * (Pk[i - 3] - Pk[i + 1]) + A[6][5] * (Pk[i - 2] - Pk[i]);
            hq = A[6][0] * (Qk[i - 7] - Qk[i + 5]) + A[6][1] * (Qk[i - 6] - Qk[i + 4]) +
                 A[6][2] * (Qk[i - 5] - Qk[i + 3]) + A[6][3] * (Qk[i - 4] - Qk[i + 2]) +
                 A[6][4] * (Qk[i - 3] - Qk[i + 1]) + A[6][5] * (Qk[i - 2] - Qk[i]);
        }
        dp[i - 1] = hp;
        dq[i - 1] = hq;
    }
    __syncthreads();
    const int mtp = max(mf[j], mf[k]);
    double piece[4];
    for (int term = 0; term < 4; ++term) {
        for (int i = 1 + lane; i <= mtp + 3; i += 64) {
            double v = 0.0;
            if (i >= 2 && i <= mtp) {
                if (term == 0) v = Qj[i - 1] * dp[i - 1] - Pj[i - 1] * dq[i - 1];
                else if (term == 1) v = rp[i - 1] * Qj[i - 1] * Qk[i - 1];
                else if (term == 2) v = rpor[i - 1] * (Pj[i - 1] * Qk[i - 1] + Qj[i - 1] * Pk[i - 1]);
                else v = rpor[i - 1] * zz[i - 1] * (Pj[i - 1] * Pk[i - 1] + Qj[i - 1] * Qk[i - 1]);
            }
            ta[i - 1] = v;
        }
        __syncthreads();
        piece[term] = wave_quad(ta, r, rp, mtp, kc, qgrp);
        __syncthreads();
    }
    const double p1 = cl * piece[0] / h;
    const double p2 = -2.0 * cl * cl * piece[1];
    const double p3 = piece[2] * cl * (double)nak[k];
    const double p4 = -piece[3];
    if (lane == 0) out[(size_t)j * nw + k] = p1 + p2 + p3 + p4;
}

bool triangrk(int la, int k, int lb) {
    if ((k + la + lb) % 2 != 0) return false;
    if (std::abs(la - lb) > k) return false;
    if (la + lb < k) return false;
    return true;
}

template <class T>
DBuf<T> to_dev(const T *h, size_t n) {
    DBuf<T> d(n);
    if (n) GCI_HIP(hipMemcpy(d.p, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

void check_inputs(const gci_grid *g, const gci_orbitals *o) {
    GCI_REQUIRE(g && o, "null grid/orbitals");
    GCI_REQUIRE(g->n >= 14 && g->n + 10 <= NMAXG, "grid: n must be in 14..590");
    GCI_REQUIRE(g->r && g->rp && g->rpor, "grid: null arrays");
    GCI_REQUIRE(o->nw >= 1 && o->nw <= 127 && o->ld >= g->n, "orbitals: bad nw/ld");
    GCI_REQUIRE(o->P && o->Q && o->mf && o->nak && o->nkl && o->nkj, "orbitals: null arrays");
    for (int j = 0; j < o->nw; ++j)
        GCI_REQUIRE(o->mf[j] >= 2 && o->mf[j] <= g->n - 4, "orbitals: MF out of range (2..N-4)");
}

}  // namespace

}  // namespace gci

using namespace gci;

extern "C" {

int gci_rk_table(const gci_grid *grid, const gci_orbitals *orb, int32_t *j2max_out, int32_t *kstart, int32_t *keys,
                 double *vals, int64_t cap, int64_t *n_out) {
    return guarded([&] {
        check_inputs(grid, orb);
        GCI_REQUIRE(n_out, "gci_rk_table: null count");
        const int nw = orb->nw, key = nw + 1;
        int j2max = orb->nkj[0];
        for (int i = 1; i < nw; ++i) j2max = std::max(j2max, (int)orb->nkj[i]);
        GCI_REQUIRE(j2max <= 20, "genintrk: KMAX too small");
        if (j2max_out) *j2max_out = j2max;
        // enumeration (genintrk.f90:56-86)
        std::vector<int32_t> k_of, lab;
        std::vector<int32_t> kst(j2max + 2);
        kst[0] = 1;
        for (int k = 0; k <= j2max; ++k) {
            for (int ia = 1; ia <= nw; ++ia)
                for (int ib = ia; ib <= nw; ++ib)
                    for (int ic = ia; ic <= nw; ++ic) {
                        if (!triangrk(orb->nkl[ia - 1], k, orb->nkl[ic - 1])) continue;
                        for (int id = ib; id <= nw; ++id) {
                            if (!triangrk(orb->nkl[ib - 1], k, orb->nkl[id - 1])) continue;
                            k_of.push_back(k);
                            lab.insert(lab.end(), {ia, ib, ic, id});
                        }
                    }
            kst[k + 1] = (int32_t)k_of.size() + 1;
        }
        const int64_t n = (int64_t)k_of.size();
        *n_out = n;
        if (kstart) std::copy(kst.begin(), kst.end(), kstart);
        if (!keys && !vals) return;
        GCI_REQUIRE(cap >= n, "gci_rk_table: capacity too small");
        if (keys)
            for (int64_t e = 0; e < n; ++e)
                keys[e] = ((lab[4 * e] * key + lab[4 * e + 1]) * key + lab[4 * e + 2]) * key + lab[4 * e + 3];
        if (!vals || n == 0) return;
        // group entries by Y^k(b, d)
        std::map<std::tuple<int, int, int>, std::vector<int>> groups;
        for (int64_t e = 0; e < n; ++e) groups[{k_of[e], lab[4 * e + 1] - 1, lab[4 * e + 3] - 1}].push_back((int)e);
        std::vector<YDesc> ys;
        std::vector<int> ac, outi;
        for (auto &kv : groups) {
            YDesc y;
            y.k = std::get<0>(kv.first);
            y.b = std::get<1>(kv.first);
            y.d = std::get<2>(kv.first);
            y.e0 = (int)outi.size();
            for (int e : kv.second) {
                ac.push_back(lab[4 * e] - 1);
                ac.push_back(lab[4 * e + 2] - 1);
                outi.push_back(e);
            }
            y.e1 = (int)outi.size();
            ys.push_back(y);
        }
        RkConst kc = make_const(grid);
        const int n10 = grid->n + 10;
        auto dP = to_dev(orb->P, (size_t)nw * orb->ld);
        auto dQ = to_dev(orb->Q, (size_t)nw * orb->ld);
        auto dmf = to_dev(orb->mf, nw);
        auto dr = to_dev(grid->r, n10), drp = to_dev(grid->rp, n10), drpor = to_dev(grid->rpor, n10);
        auto dys = to_dev(ys.data(), ys.size());
        auto dac = to_dev(ac.data(), ac.size());
        auto dout = to_dev(outi.data(), outi.size());
        DBuf<double> dv(n);
        hipLaunchKernelGGL(k_rk_table, dim3((unsigned)ys.size()), dim3(256), 0, 0, kc, dys.p, dac.p, dout.p, dP.p,
                           dQ.p, orb->ld, dmf.p, dr.p, drp.p, drpor.p, dv.p, nullptr, nullptr);
        GCI_HIP(hipGetLastError());
        GCI_HIP(hipMemcpy(vals, dv.p, n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int gci_yzk(const gci_grid *grid, const gci_orbitals *orb, int npairs, const int32_t *k, const int32_t *b,
            const int32_t *d, double *yk, int32_t *mtp) {
    return guarded([&] {
        check_inputs(grid, orb);
        GCI_REQUIRE(npairs >= 0, "gci_yzk: negative count");
        if (npairs == 0) return;
        GCI_REQUIRE(k && b && d && yk && mtp, "gci_yzk: null argument");
        const int nw = orb->nw;
        std::vector<YDesc> ys(npairs);
        for (int q = 0; q < npairs; ++q) {
            GCI_REQUIRE(k[q] >= 0 && k[q] <= 20, "gci_yzk: k out of range (0..20)");
            GCI_REQUIRE(b[q] >= 1 && b[q] <= nw && d[q] >= 1 && d[q] <= nw, "gci_yzk: orbital index out of range");
            ys[q] = YDesc{k[q], b[q] - 1, d[q] - 1, 0, 0};  // no (a, c) entries: the function only
        }
        RkConst kc = make_const(grid);
        const int n = grid->n, n10 = n + 10;
        auto dP = to_dev(orb->P, (size_t)nw * orb->ld);
        auto dQ = to_dev(orb->Q, (size_t)nw * orb->ld);
        auto dmf = to_dev(orb->mf, nw);
        auto dr = to_dev(grid->r, n10), drp = to_dev(grid->rp, n10), drpor = to_dev(grid->rpor, n10);
        auto dys = to_dev(ys.data(), ys.size());
        DBuf<double> dyk((size_t)npairs * n);
        DBuf<int> dm(npairs);
        hipLaunchKernelGGL(k_rk_table, dim3((unsigned)npairs), dim3(256), 0, 0, kc, dys.p, nullptr, nullptr, dP.p,
                           dQ.p, orb->ld, dmf.p, dr.p, drp.p, drpor.p, nullptr, dyk.p, dm.p);
        GCI_HIP(hipGetLastError());
        GCI_HIP(hipMemcpy(yk, dyk.p, dyk.bytes(), hipMemcpyDeviceToHost));
        GCI_HIP(hipMemcpy(mtp, dm.p, dm.bytes(), hipMemcpyDeviceToHost));
    });
}

int gci_one_body(const gci_grid *grid, const gci_orbitals *orb, const double *zz, double c, double *iab) {
    return guarded([&] {
        check_inputs(grid, orb);
        GCI_REQUIRE(zz && iab, "gci_one_body: null argument");
        const int nw = orb->nw;
        std::vector<int> pairs;
        for (int a = 0; a < nw; ++a)
            for (int b = a; b < nw; ++b)
                if (orb->nak[a] == orb->nak[b]) pairs.insert(pairs.end(), {a, b});
        std::fill(iab, iab + (size_t)nw * nw, 0.0);
        const int np = (int)pairs.size() / 2;
        if (!np) return;
        // Lagrange table A(13,13) = B13 / 12! (setqic.f90:291-296)
        double a13[169];
        for (int i = 1; i <= 13; ++i)
            for (int j = 1; j <= 13; ++j) {
                double b;
                if (i == j) {
                    double s = 0.0;
                    for (int m = 1; m <= 13; ++m)
                        if (m != j) s += 479001600.0 / (double)(j - m);
                    b = s;
                } else {
                    double num = 1.0, den = 1.0;
                    for (int m = 1; m <= 13; ++m) {
                        if (m == j) continue;
                        den *= (double)(j - m);
                        if (m != i) num *= (double)(i - m);
                    }
                    b = 479001600.0 * num / den;
                }
                a13[(i - 1) * 13 + (j - 1)] = (double)std::llround(b) * (1.0 / 479001600.0);
            }
        RkConst kc = make_const(grid);
        const int n10 = grid->n + 10;
        auto dP = to_dev(orb->P, (size_t)nw * orb->ld);
        auto dQ = to_dev(orb->Q, (size_t)nw * orb->ld);
        auto dmf = to_dev(orb->mf, nw);
        auto dnak = to_dev(orb->nak, nw);
        auto dr = to_dev(grid->r, n10), drp = to_dev(grid->rp, n10), drpor = to_dev(grid->rpor, n10);
        auto dzz = to_dev(zz, grid->n);
        auto da = to_dev(a13, 169);
        auto dpairs = to_dev(pairs.data(), pairs.size());
        DBuf<double> dout((size_t)nw * nw);
        GCI_HIP(hipMemset(dout.p, 0, dout.bytes()));
        hipLaunchKernelGGL(k_rinti, dim3(np), dim3(64), 0, 0, kc, dpairs.p, np, da.p, grid->h, c, dzz.p, dP.p, dQ.p,
                           orb->ld, dmf.p, dnak.p, dr.p, drp.p, drpor.p, dout.p, nw);
        GCI_HIP(hipGetLastError());
        GCI_HIP(hipMemcpy(iab, dout.p, dout.bytes(), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
