// spmm.hip -- symmetric sparse matrix x block-of-vectors for CI Hamiltonians.
//
// Replaces SPICMV2/DMERGE (lib9290/spicmv2.f90:40-56, dmerge.f90:36-40):
// y = H x with H symmetric and only the strictly-lower off-diagonal part
// stored, plus an FP64 diagonal.  Every stored entry H_ij (i > j) acts twice:
// a gather into row i and the symmetric scatter into row j.
//
// Layout (panel-CSC, built once per matrix on the device): rows are cut into
// panels of R rows; inside a panel the entries are sorted by column, then
// row, so each column the panel touches is one contiguous segment.  Per entry
// the stream carries only the fp32/fp64 value and a 16-bit row offset.
//
// Kernel: one workgroup per work item (a slice of a panel's segments, sized
// by entry count for load balance).  The panel's rows of x are staged in LDS
// (fp32 for the bulk operator) and its row accumulators y_i live in LDS.
// Groups of M lanes (lane v = vector v, M = launch width) walk a segment:
// the gather y_i += H_ij x_j goes to LDS; the scatter sum s = sum_i H_ij x_i
// stays in a register and leaves as one M-wide atomic add into y_j -- one
// global atomic per segment instead of one per entry.  The panel's y rows
// are flushed with atomics at the end.  x and y pass through row-major
// staging so that both the staging loads and the atomics are line-sized.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "gci_internal.h"

namespace gci {

namespace {

// ------------------------------------------------------ panel-CSC build

__global__ void k_keys(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col, int R,
                       uint64_t *__restrict__ keys, int64_t *__restrict__ idx) {
    // one wave per row
    const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const uint64_t panel = (uint64_t)(row / R), lr = (uint64_t)(row % R);
    for (int64_t e = rp[row] + lane; e < rp[row + 1]; e += 64) {
        keys[e] = (panel << 40) | ((uint64_t)col[e] << 16) | lr;
        idx[e] = e;
    }
}

__global__ void k_seg_flags(int64_t nnz, const uint64_t *__restrict__ keys, int32_t *__restrict__ flag) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
        flag[e] = (e == 0 || (keys[e] >> 16) != (keys[e - 1] >> 16)) ? 1 : 0;
}

__global__ void k_seg_fill(int64_t nnz, const uint64_t *__restrict__ keys, const int32_t *__restrict__ flag,
                           const int64_t *__restrict__ segid, const int64_t *__restrict__ perm,
                           const double *__restrict__ val, const float *__restrict__ hi, const float *__restrict__ lo,
                           int32_t *__restrict__ seg_col, int64_t *__restrict__ seg_start,
                           uint16_t *__restrict__ lrow, double *__restrict__ pval, float *__restrict__ phi,
                           float *__restrict__ plo) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[e];
        if (flag[e]) {
            const int64_t s = segid[e];
            seg_col[s] = (int32_t)((k >> 16) & 0xFFFFFFull);
            seg_start[s] = e;
        }
        lrow[e] = (uint16_t)(k & 0xFFFFull);
        const int64_t src = perm[e];
        if (pval) pval[e] = val[src];
        if (phi) {
            phi[e] = hi[src];
            plo[e] = lo[src];
        }
    }
}

// ------------------------------------------------------------ kernels

template <int MODE>
struct XT;
template <>
struct XT<GCI_OP_FP64> {
    using T = double;
};
template <>
struct XT<GCI_OP_SPLIT_BULK> {
    using T = float;
};
template <>
struct XT<GCI_OP_SPLIT_FULL> {
    using T = double;
};

// Broadcast lane Q of each group of M lanes to the whole group: DPP quad
// permutes for M <= 4 (a VALU move, no LDS traffic), ds_bpermute for M = 8.
template <int M, int Q>
__device__ inline int bcast(int x) {
    if constexpr (M == 1) {
        return x;
    } else if constexpr (M == 2) {
        constexpr int ctrl = Q == 0 ? 0xA0 : 0xF5;  // quad_perm [0,0,2,2] / [1,1,3,3]
        return __builtin_amdgcn_mov_dpp(x, ctrl, 0xF, 0xF, false);
    } else if constexpr (M == 4) {
        constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);  // quad_perm [Q,Q,Q,Q]
        return __builtin_amdgcn_mov_dpp(x, ctrl, 0xF, 0xF, false);
    } else {
        return __shfl(x, (int)((threadIdx.x & 63) & ~7u) | Q, 64);
    }
}

// One stored entry as it travels through the group: its row offset and its
// value in the operator's precision (two 32-bit words for FP64 and hi+lo).
template <int MODE>
struct Ent {
    int lr;
    int w0, w1;
    __device__ inline void load(const SpmmJob &j, int64_t e) {
        lr = j.lrow[e];
        if constexpr (MODE == GCI_OP_FP64) {
            const double v = j.val[e];
            w0 = __double2loint(v);
            w1 = __double2hiint(v);
        } else if constexpr (MODE == GCI_OP_SPLIT_BULK) {
            w0 = __float_as_int(j.hi[e]);
            w1 = 0;
        } else {
            w0 = __float_as_int(j.hi[e]);
            w1 = __float_as_int(j.lo[e]);
        }
    }
    __device__ inline double value() const {
        if constexpr (MODE == GCI_OP_FP64) return __hiloint2double(w1, w0);
        else if constexpr (MODE == GCI_OP_SPLIT_BULK) return (double)__int_as_float(w0);
        else return (double)__int_as_float(w0) + (double)__int_as_float(w1);
    }
    template <int M, int Q>
    __device__ inline Ent from_lane() const {
        Ent r;
        r.lr = bcast<M, Q>(lr);
        r.w0 = bcast<M, Q>(w0);
        if constexpr (MODE != GCI_OP_SPLIT_BULK) r.w1 = bcast<M, Q>(w1);
        else r.w1 = 0;
        return r;
    }
};

template <int M, int MODE, int U, int Q>
__device__ inline void consume(const Ent<MODE> (&en)[U], int u, const typename XT<MODE>::T *xs, double *ys,
                               int v, double xc, double &acc) {
    const Ent<MODE> q = en[u].template from_lane<M, Q>();
    if (q.lr >= 0) {
        const double a = q.value();
        acc = fma(a, (double)xs[q.lr * M + v], acc);
        atomicAdd(&ys[q.lr * M + v], a * xc);
    }
    if constexpr (Q + 1 < M) consume<M, MODE, U, Q + 1>(en, u, xs, ys, v, xc, acc);
}

template <int M, int MODE>
__global__ __launch_bounds__(256) void k_pcsc(const SpmmItem *__restrict__ items, const SpmmJob *__restrict__ jobs) {
    using T = typename XT<MODE>::T;
    constexpr int U = 4;  // entry batches each lane keeps in flight
    extern __shared__ __align__(16) char smem[];
    const SpmmItem it = items[blockIdx.x];
    const SpmmJob jb = jobs[it.job];
    if (!jb.active || jb.mode != MODE) return;
    const int rows = it.rows;
    T *xs = reinterpret_cast<T *>(smem);
    double *ys = reinterpret_cast<double *>(smem + (((size_t)rows * M * sizeof(T) + 15) & ~(size_t)15));
    const T *gx = reinterpret_cast<const T *>(jb.xs);
    const int t = threadIdx.x;
    for (int i = t; i < rows * M; i += blockDim.x) {
        xs[i] = gx[it.row0 * M + i];
        ys[i] = 0.0;
    }
    __syncthreads();
    const int v = t % M;
    const int g = t / M, ng = blockDim.x / M;
    // segment metadata is fetched one segment ahead
    int64_t s = it.seg0 + g;
    int64_t cn = 0, e0n = 0, e1n = 0;
    if (s < it.seg1) {
        cn = jb.seg_col[s];
        e0n = jb.seg_start[s];
        e1n = jb.seg_start[s + 1];
    }
    for (; s < it.seg1; s += ng) {
        const int64_t c = cn, e0 = e0n, e1 = e1n;
        const int64_t sn = s + ng;
        if (sn < it.seg1) {
            cn = jb.seg_col[sn];
            e0n = jb.seg_start[sn];
            e1n = jb.seg_start[sn + 1];
        }
        const double xc = (double)gx[c * M + v];
        double acc = 0.0;
        for (int64_t e = e0; e < e1; e += U * M) {
            Ent<MODE> en[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t idx = e + u * M + v;  // lanes of the group load consecutive entries
                if (idx < e1) en[u].load(jb, idx);
                else en[u] = Ent<MODE>{-1, 0, 0};
            }
#pragma unroll
            for (int u = 0; u < U; ++u) consume<M, MODE, U, 0>(en, u, xs, ys, v, xc, acc);
        }
        if (v < jb.m) unsafeAtomicAdd(&jb.ys[c * M + v], acc);
    }
    __syncthreads();
    for (int i = t; i < rows * M; i += blockDim.x) {
        const double y = ys[i];
        if (y != 0.0) unsafeAtomicAdd(&jb.ys[it.row0 * M + i], y);
    }
}

// staging: xs[i*M + v] = x[i + v*ldx] (rounded to T), ys = 0
template <class T>
__global__ void k_pack(int64_t n, int m, int M, const double *__restrict__ x, int64_t ldx, T *__restrict__ xs,
                       double *__restrict__ ys) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n * M; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = q / M;
        const int v = (int)(q % M);
        xs[q] = v < m ? (T)x[i + v * ldx] : (T)0;
        ys[q] = 0.0;
    }
}

// y[i + v*ldy] = ys[i*M + v] + d_i x[i + v*ldx]  (diagonal in FP64)
__global__ void k_unpack(int64_t n, int m, int M, const double *__restrict__ ys, const double *__restrict__ diag,
                         const double *__restrict__ x, int64_t ldx, double *__restrict__ y, int64_t ldy) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n * m; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = q % n;
        const int v = (int)(q / n);
        y[i + v * ldy] = fma(diag[i], x[i + v * ldx], ys[i * M + v]);
    }
}

template <int M, int MODE>
void launch_items(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, int R, hipStream_t s) {
    using T = typename XT<MODE>::T;
    const size_t shm = (((size_t)R * M * sizeof(T) + 15) & ~(size_t)15) + (size_t)R * M * sizeof(double);
    static bool attr_set = false;  // one instantiation per (M, MODE)
    if (!attr_set) {
        GCI_HIP(hipFuncSetAttribute((const void *)k_pcsc<M, MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)shm));
        attr_set = true;
    }
    hipLaunchKernelGGL((k_pcsc<M, MODE>), dim3((unsigned)nitems), dim3(256), shm, s, d_items, d_jobs);
}

template <int MODE>
void launch_mode(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, int M, int R, hipStream_t s) {
    switch (M) {
        case 1: launch_items<1, MODE>(d_items, nitems, d_jobs, R, s); break;
        case 2: launch_items<2, MODE>(d_items, nitems, d_jobs, R, s); break;
        case 4: launch_items<4, MODE>(d_items, nitems, d_jobs, R, s); break;
        default: launch_items<8, MODE>(d_items, nitems, d_jobs, R, s); break;
    }
}

int g_R = 0;  // panel height all matrices use (set by ensure_pcsc)

}  // namespace

int spmm_width(int m) { return m <= 1 ? 1 : m <= 2 ? 2 : m <= 4 ? 4 : 8; }

void spmm_items(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, int M, int mode, hipStream_t s) {
    if (nitems <= 0) return;
    const int R = g_R > 0 ? g_R : 1024;
    switch (mode) {
        case GCI_OP_FP64: launch_mode<GCI_OP_FP64>(d_items, nitems, d_jobs, M, R, s); break;
        case GCI_OP_SPLIT_BULK: launch_mode<GCI_OP_SPLIT_BULK>(d_items, nitems, d_jobs, M, R, s); break;
        default: launch_mode<GCI_OP_SPLIT_FULL>(d_items, nitems, d_jobs, M, R, s); break;
    }
}

void ensure_pcsc(HMat &h, int R, int64_t target) {
    if (h.pc.R) return;
    GCI_REQUIRE(R >= 1 && R <= 65536, "pcsc: panel height out of range");
    GCI_REQUIRE(g_R == 0 || g_R == R, "pcsc: all matrices share one panel height");
    g_R = R;
    Pcsc &pc = h.pc;
    const int64_t nnz = h.nnz, n = h.n;
    hipStream_t s = h.stream;
    pc.nnz = nnz;
    if (nnz > 0) {
        DBuf<uint64_t> k0(nnz), k1(nnz);
        DBuf<int64_t> i0(nnz), i1(nnz);
        const int64_t blocks = (n * 64 + 255) / 256;
        hipLaunchKernelGGL(k_keys, dim3((unsigned)blocks), dim3(256), 0, s, n, h.rp.p, h.col.p, R, k0.p, i0.p);
        GCI_HIP(hipGetLastError());
        size_t tb = 0;
        GCI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0.p, k1.p, i0.p, i1.p, nnz, 0, 64, s));
        {
            DBuf<char> tmp(tb);
            GCI_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k0.p, k1.p, i0.p, i1.p, nnz, 0, 64, s));
        }
        k0.release();
        i0.release();
        DBuf<int32_t> flag(nnz);
        DBuf<int64_t> segid(nnz + 1);
        const int64_t eb = std::min<int64_t>((nnz + 255) / 256, 65536);
        hipLaunchKernelGGL(k_seg_flags, dim3((unsigned)eb), dim3(256), 0, s, nnz, k1.p, flag.p);
        tb = 0;
        GCI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag.p, segid.p, nnz, s));
        {
            DBuf<char> tmp(tb);
            GCI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flag.p, segid.p, nnz, s));
        }
        int64_t last = 0;
        int32_t lastf = 0;
        GCI_HIP(hipMemcpyAsync(&last, segid.p + nnz - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipMemcpyAsync(&lastf, flag.p + nnz - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipStreamSynchronize(s));
        pc.nseg = last + lastf;
        pc.seg_col.alloc(pc.nseg);
        pc.seg_start.alloc(pc.nseg + 1);
        pc.lrow.alloc(nnz);
        if (h.val.p) pc.val.alloc(nnz);
        if (h.precision == GCI_SPLIT32) {
            pc.hi.alloc(nnz);
            pc.lo.alloc(nnz);
        }
        hipLaunchKernelGGL(k_seg_fill, dim3((unsigned)eb), dim3(256), 0, s, nnz, k1.p, flag.p, segid.p, i1.p,
                           h.val.p, h.hi.p, h.lo.p, pc.seg_col.p, pc.seg_start.p, pc.lrow.p, pc.val.p, pc.hi.p,
                           pc.lo.p);
        GCI_HIP(hipGetLastError());
        GCI_HIP(hipMemcpyAsync(pc.seg_start.p + pc.nseg, &nnz, sizeof(int64_t), hipMemcpyHostToDevice, s));
        GCI_HIP(hipStreamSynchronize(s));
    }
    // work items: panels cut into slices of about `target` entries
    std::vector<int64_t> ss(pc.nseg + 1);
    std::vector<int32_t> sc(pc.nseg);
    if (pc.nseg) {
        GCI_HIP(hipMemcpy(ss.data(), pc.seg_start.p, (pc.nseg + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        GCI_HIP(hipMemcpy(sc.data(), pc.seg_col.p, pc.nseg * sizeof(int32_t), hipMemcpyDeviceToHost));
    } else {
        ss[0] = 0;
    }
    // segment -> panel: segments are sorted by (panel, column); recover the
    // panel of each segment from its first entry's row via the CSR
    std::vector<int64_t> rp(n + 1);
    GCI_HIP(hipMemcpy(rp.data(), h.rp.p, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
    const int64_t npan = (n + R - 1) / R;
    std::vector<int64_t> panel_nnz(npan, 0);
    for (int64_t p = 0; p < npan; ++p) panel_nnz[p] = rp[std::min(n, (p + 1) * R)] - rp[p * R];
    pc.items.clear();
    int64_t seg = 0;
    for (int64_t p = 0; p < npan; ++p) {
        const int64_t row0 = p * R;
        const int rows = (int)std::min<int64_t>(R, n - row0);
        const int64_t end_e = ss[seg] + panel_nnz[p];  // entries of this panel are contiguous
        int64_t s0 = seg;
        while (seg < pc.nseg && ss[seg] < end_e) {
            ++seg;
            if (ss[seg] - ss[s0] >= target || seg == pc.nseg || ss[seg] >= end_e) {
                pc.items.push_back(SpmmItem{s0, seg, row0, rows, 0});
                s0 = seg;
            }
        }
    }
    GCI_REQUIRE(seg == pc.nseg, "pcsc: segment/panel bookkeeping mismatch");
    pc.R = R;  // built
}

void spmm(HMat &h, int m, const double *x, int ldx, double *y, int ldy, int mode, hipStream_t s) {
    GCI_REQUIRE(m >= 1, "spmm: m < 1");
    GCI_REQUIRE(ldx >= h.n && ldy >= h.n, "spmm: leading dimension < n");
    if (mode != GCI_OP_FP64)
        GCI_REQUIRE(h.precision == GCI_SPLIT32, "spmm: split operator on an FP64-only matrix");
    else
        GCI_REQUIRE(h.val.p != nullptr, "spmm: FP64 operator on a matrix without FP64 values");
    if (h.n == 0) return;
    ensure_pcsc(h);
    for (int c0 = 0; c0 < m; c0 += 8) {
        const int mc = std::min(8, m - c0);
        const int M = spmm_width(mc);
        const size_t xb = mode == GCI_OP_SPLIT_BULK ? sizeof(float) : sizeof(double);
        DBuf<char> xs((size_t)h.n * M * xb);
        DBuf<double> ys((size_t)h.n * M);
        const double *xc = x + (int64_t)c0 * ldx;
        double *yc = y + (int64_t)c0 * ldy;
        const int64_t pb = std::min<int64_t>((h.n * M + 255) / 256, 65536);
        if (mode == GCI_OP_SPLIT_BULK)
            hipLaunchKernelGGL(k_pack<float>, dim3((unsigned)pb), dim3(256), 0, s, h.n, mc, M, xc, (int64_t)ldx,
                               (float *)xs.p, ys.p);
        else
            hipLaunchKernelGGL(k_pack<double>, dim3((unsigned)pb), dim3(256), 0, s, h.n, mc, M, xc, (int64_t)ldx,
                               (double *)xs.p, ys.p);
        if (!h.pc.items.empty()) {
            SpmmJob jb{h.pc.seg_col.p, h.pc.seg_start.p, h.pc.lrow.p, h.pc.val.p, h.pc.hi.p, h.pc.lo.p, xs.p, ys.p,
                       mc, mode, 1, 0};
            DBuf<SpmmJob> dj(1);
            DBuf<SpmmItem> di(h.pc.items.size());
            GCI_HIP(hipMemcpyAsync(dj.p, &jb, sizeof(SpmmJob), hipMemcpyHostToDevice, s));
            GCI_HIP(hipMemcpyAsync(di.p, h.pc.items.data(), h.pc.items.size() * sizeof(SpmmItem),
                                   hipMemcpyHostToDevice, s));
            spmm_items(di.p, (int64_t)h.pc.items.size(), dj.p, M, mode, s);
            GCI_HIP(hipGetLastError());
            const int64_t ub = std::min<int64_t>((h.n * mc + 255) / 256, 65536);
            hipLaunchKernelGGL(k_unpack, dim3((unsigned)ub), dim3(256), 0, s, h.n, mc, M, (const double *)ys.p,
                               (const double *)h.diag.p, xc, (int64_t)ldx, yc, (int64_t)ldy);
            GCI_HIP(hipStreamSynchronize(s));  // staging buffers are released on return
        } else {
            const int64_t ub = std::min<int64_t>((h.n * mc + 255) / 256, 65536);
            hipLaunchKernelGGL(k_unpack, dim3((unsigned)ub), dim3(256), 0, s, h.n, mc, M, (const double *)ys.p,
                               (const double *)h.diag.p, xc, (int64_t)ldx, yc, (int64_t)ldy);
            GCI_HIP(hipStreamSynchronize(s));
        }
        GCI_HIP(hipGetLastError());
    }
}

}  // namespace gci
