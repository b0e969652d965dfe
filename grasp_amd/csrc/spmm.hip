// spmm.hip -- symmetric sparse matrix x block-of-vectors for CI Hamiltonians.
//
// Replaces SPICMV2/DMERGE (lib9290/spicmv2.f90:40-56, dmerge.f90:36-40):
// y = H x with H symmetric and only the strictly-lower off-diagonal part
// stored (CSR, ascending columns) plus an FP64 diagonal.  Each stored entry
// contributes twice: a gather into its own row (dot product, kept in
// registers and reduced across the lanes of the row group) and a symmetric
// scatter into the row named by its column (hardware FP64 atomic add).
//
// Row groups of G lanes (G in {4, 8, 16, 32, 64}) stride over the entries of
// one row; G is chosen per matrix by a short timing sweep (spmm_group).
// Column-major x, y with leading dimensions; M vectors per pass.
#include "gci_internal.h"

#include <algorithm>
#include <chrono>

namespace gci {

namespace {

__global__ void k_zero(int64_t n, int m, double *__restrict__ y, int ldy) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t total = n * m;
    for (; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t v = i / n, r = i - v * n;
        y[r + v * (int64_t)ldy] = 0.0;
    }
}

template <int MODE>
struct Val;
template <>
struct Val<GCI_OP_FP64> {
    const double *__restrict__ v;
    __device__ double a(int64_t e) const { return v[e]; }
    __device__ static double xin(double x) { return x; }
};
template <>
struct Val<GCI_OP_SPLIT_BULK> {
    const float *__restrict__ hi;
    __device__ double a(int64_t e) const { return (double)hi[e]; }
    // fp32-rounded operand; the product of two fp32 values is exact in FP64
    __device__ static double xin(double x) { return (double)(float)x; }
};
template <>
struct Val<GCI_OP_SPLIT_FULL> {
    const float *__restrict__ hi;
    const float *__restrict__ lo;
    __device__ double a(int64_t e) const { return (double)hi[e] + (double)lo[e]; }
    __device__ static double xin(double x) { return x; }
};

template <int G, int M, int MODE>
__global__ __launch_bounds__(256) void k_spmm_sym(int64_t n, const int64_t *__restrict__ rp,
                                                  const int32_t *__restrict__ col, Val<MODE> val,
                                                  const double *__restrict__ diag,
                                                  const double *__restrict__ x, int ldx,
                                                  double *__restrict__ y, int ldy) {
    const int lane = threadIdx.x % G;
    const int64_t group = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
    const int64_t ngroups = (int64_t)gridDim.x * blockDim.x / G;
    for (int64_t r = group; r < n; r += ngroups) {
        double xr[M], acc[M];
#pragma unroll
        for (int v = 0; v < M; ++v) {
            xr[v] = x[r + (int64_t)v * ldx];
            acc[v] = 0.0;
        }
        const int64_t beg = rp[r], end = rp[r + 1];
        for (int64_t e = beg + lane; e < end; e += G) {
            const int64_t c = col[e];
            const double a = val.a(e);
#pragma unroll
            for (int v = 0; v < M; ++v) {
                acc[v] = fma(a, Val<MODE>::xin(x[c + (int64_t)v * ldx]), acc[v]);
                unsafeAtomicAdd(&y[c + (int64_t)v * ldy], a * Val<MODE>::xin(xr[v]));
            }
        }
#pragma unroll
        for (int v = 0; v < M; ++v) {
#pragma unroll
            for (int off = G / 2; off > 0; off >>= 1) acc[v] += __shfl_xor(acc[v], off, G);
        }
        if (lane == 0) {
            const double d = diag[r];
#pragma unroll
            for (int v = 0; v < M; ++v) unsafeAtomicAdd(&y[r + (int64_t)v * ldy], fma(d, xr[v], acc[v]));
        }
    }
}

template <int G, int M, int MODE>
void launch_g(const HMat &h, const double *x, int ldx, double *y, int ldy, hipStream_t s) {
    const int block = 256;
    int64_t groups_needed = h.n;
    int64_t blocks = (groups_needed * G + block - 1) / block;
    blocks = std::min<int64_t>(blocks, 256 * 16);
    if (blocks < 1) blocks = 1;
    Val<MODE> vv;
    if constexpr (MODE == GCI_OP_FP64) {
        vv.v = h.val.p;
    } else if constexpr (MODE == GCI_OP_SPLIT_BULK) {
        vv.hi = h.hi.p;
    } else {
        vv.hi = h.hi.p;
        vv.lo = h.lo.p;
    }
    hipLaunchKernelGGL((k_spmm_sym<G, M, MODE>), dim3((unsigned)blocks), dim3(block), 0, s, h.n, h.rp.p,
                       h.col.p, vv, h.diag.p, x, ldx, y, ldy);
}

template <int M, int MODE>
void launch_m(const HMat &h, int G, const double *x, int ldx, double *y, int ldy, hipStream_t s) {
    switch (G) {
        case 4: launch_g<4, M, MODE>(h, x, ldx, y, ldy, s); break;
        case 8: launch_g<8, M, MODE>(h, x, ldx, y, ldy, s); break;
        case 16: launch_g<16, M, MODE>(h, x, ldx, y, ldy, s); break;
        case 32: launch_g<32, M, MODE>(h, x, ldx, y, ldy, s); break;
        default: launch_g<64, M, MODE>(h, x, ldx, y, ldy, s); break;
    }
}

template <int MODE>
void launch_mode(const HMat &h, int G, int m, const double *x, int ldx, double *y, int ldy,
                 hipStream_t s) {
    int done = 0;
    while (done < m) {
        int left = m - done;
        const double *xs = x + (int64_t)done * ldx;
        double *ys = y + (int64_t)done * ldy;
        if (left >= 8) { launch_m<8, MODE>(h, G, xs, ldx, ys, ldy, s); done += 8; }
        else if (left >= 4) { launch_m<4, MODE>(h, G, xs, ldx, ys, ldy, s); done += 4; }
        else if (left >= 2) { launch_m<2, MODE>(h, G, xs, ldx, ys, ldy, s); done += 2; }
        else { launch_m<1, MODE>(h, G, xs, ldx, ys, ldy, s); done += 1; }
    }
}

int default_group(const HMat &h) {
    double avg = h.n ? (double)h.nnz / (double)h.n : 0.0;
    if (avg < 6) return 4;
    if (avg < 12) return 8;
    if (avg < 24) return 16;
    if (avg < 48) return 32;
    return 64;
}

}  // namespace

void spmm(const HMat &h, int m, const double *x, int ldx, double *y, int ldy, int mode,
          hipStream_t s) {
    GCI_REQUIRE(m >= 1, "spmm: m < 1");
    GCI_REQUIRE(ldx >= h.n && ldy >= h.n, "spmm: leading dimension < n");
    if (mode != GCI_OP_FP64)
        GCI_REQUIRE(h.precision == GCI_SPLIT32, "spmm: split operator on an FP64-only matrix");
    else
        GCI_REQUIRE(h.val.p != nullptr, "spmm: FP64 operator on a matrix without FP64 values");
    if (h.n == 0) return;
    const int G = h.group ? h.group : default_group(h);
    int64_t zb = std::min<int64_t>((h.n * m + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)zb), dim3(256), 0, s, h.n, m, y, ldy);
    switch (mode) {
        case GCI_OP_FP64: launch_mode<GCI_OP_FP64>(h, G, m, x, ldx, y, ldy, s); break;
        case GCI_OP_SPLIT_BULK: launch_mode<GCI_OP_SPLIT_BULK>(h, G, m, x, ldx, y, ldy, s); break;
        case GCI_OP_SPLIT_FULL: launch_mode<GCI_OP_SPLIT_FULL>(h, G, m, x, ldx, y, ldy, s); break;
        default: GCI_REQUIRE(false, "spmm: unknown mode");
    }
    GCI_HIP(hipGetLastError());
}

// Row-group autotuner: times each candidate G on the matrix itself (x = 1)
// and keeps the fastest.  Runs once per matrix and operator width.
int spmm_group(HMat &h, int m, int mode) {
    if (h.group) return h.group;
    if (h.n == 0 || h.nnz == 0) {
        h.group = 4;
        return h.group;
    }
    DBuf<double> x((size_t)h.n * m), y((size_t)h.n * m);
    std::vector<double> ones((size_t)h.n * m, 1.0);
    GCI_HIP(hipMemcpyAsync(x.p, ones.data(), x.bytes(), hipMemcpyHostToDevice, h.stream));
    hipEvent_t e0, e1;
    GCI_HIP(hipEventCreate(&e0));
    GCI_HIP(hipEventCreate(&e1));
    int best = default_group(h);
    float best_ms = 1e30f;
    for (int G : {4, 8, 16, 32, 64}) {
        h.group = G;
        spmm(h, m, x.p, (int)h.n, y.p, (int)h.n, mode, h.stream);  // warm
        GCI_HIP(hipEventRecord(e0, h.stream));
        for (int rep = 0; rep < 3; ++rep) spmm(h, m, x.p, (int)h.n, y.p, (int)h.n, mode, h.stream);
        GCI_HIP(hipEventRecord(e1, h.stream));
        GCI_HIP(hipEventSynchronize(e1));
        float ms = 0;
        GCI_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = G;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    h.group = best;
    return best;
}

}  // namespace gci
