// radial.h -- pieces shared by the radial-integral kernels (rk.hip,
// breit.hip): the grid constants of SETQIC (lib9290/setqic.f90:249-258),
// Fortran REAL**INTEGER, and QUAD (lib9290/quad.f90) as one wavefront.
#pragma once

#include <cmath>

#include "gci_internal.h"

// the reference's arithmetic operation by operation: no fused multiply-adds
#pragma clang fp contract(off)

namespace gci {

constexpr int NMAXG = 600;  // grid points N + 10 (parameter_def NNN1)

struct RkConst {
    int n;
    double accy;
    double c1, c2, c3, c4;
    double cnc5c[3][4];  // [j-2][i-2]
};

__device__ inline double ipow_d(double a, int k) {
    double res = 1.0;
    for (;;) {
        if (k & 1) res *= a;
        k >>= 1;
        if (!k) break;
        a *= a;
    }
    return res;
}

// Fortran REAL**INTEGER for any sign of k (repeated squaring, then a
// reciprocal for k < 0)
__device__ inline double powi_d(double a, int k) {
    const bool recip = k < 0;
    double res = 1.0;
    for (;;) {
        if (k & 1) res *= a;
        k /= 2;
        if (!k) break;
        a *= a;
    }
    return recip ? 1.0 / res : res;
}

// QUAD (quad.f90:388-437) for TA in LDS (1-based view ta[i-1]), one wave;
// grp: LDS scratch of (mtp + 3) / 4 doubles.  TA(mtp+1..mtp+3) must be zero.
__device__ inline double wave_quad(const double *ta, const double *r, const double *rp, int mtp, const RkConst &k,
                           double *grp) {
    const int lane = threadIdx.x & 63;
    // first admissible start i in [2, mtp-1]
    int found = 1 << 30;
    for (int base = 2; base <= mtp - 1; base += 64) {
        const int i = base + lane;
        bool ok = false;
        if (i <= mtp - 1) {
            const double tai = ta[i - 1];
            if (!(fabs(tai) <= 0.0)) {
                const double quott = ta[i] / tai;
                ok = quott > 0.0;
            }
        }
        const unsigned long long m = __ballot(ok);
        if (m) {
            found = base + __ffsll((long long)m) - 1;
            break;
        }
    }
    if (found == (1 << 30)) return 0.0;
    const int i = found, ip1 = i + 1;
    // closed Newton-Cotes sum over groups a = ip1, ip1+4, ... <= mtp: the
    // group terms in parallel, their sum sequential in group order as
    // QUAD's SUM runs (quad.f90:84-85)
    const int ngrp = (mtp - ip1) / 4 + 1;
    for (int q = lane; q < ngrp; q += 64) {
        const int a = ip1 + 4 * q;
        grp[q] = k.c2 * (ta[a - 1] + ta[a + 1]) + k.c3 * ta[a] + k.c4 * ta[a + 2];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double s = 0.0;
    if (lane == 0)
        for (int q = 0; q < ngrp; ++q) s += grp[q];
    s = __shfl(s, 0, 64);
    const double tai = ta[i - 1], taip1 = ta[ip1 - 1];
    const double frip1 = taip1 / rp[ip1 - 1];
    double fri = tai / rp[i - 1];
    const double ratio = frip1 / fri;
    const double rip1 = r[ip1 - 1], ri = r[i - 1];
    const double sigma = log(ratio) / log(rip1 / ri);
    fri = ri * fri;
    double result = fri / (sigma + 1.0);
    result = result + k.c1 * tai;
    result = result + s;
    if ((mtp - i) % 4 == 0) result = result - k.c1 * ta[mtp - 1];
    return result;
}

inline RkConst make_const(const gci_grid *g) {
    RkConst kc;
    kc.n = g->n;
    kc.accy = g->accy;
    const double factor = g->h / 720.0;
    // Newton-Cotes numerators (setqic.f90:249-258); derived from the
    // Lagrange weights of the 5-point rule, see oracle/graspci_oracle.c
    static const double c5[3][5] = {{251, 646, -264, 106, -19}, {232, 992, 192, 32, -8}, {243, 918, 648, 378, -27}};
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 4; ++i) kc.cnc5c[j][i] = factor * c5[j][i + 1];
    kc.c1 = factor * 224.0;
    kc.c2 = factor * 1024.0;
    kc.c3 = factor * 384.0;
    kc.c4 = kc.c1 + kc.c1;
    return kc;
}


}  // namespace gci
