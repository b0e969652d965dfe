// spmm.hip -- symmetric sparse matrix x block-of-vectors for CI Hamiltonians.
//
// Replaces SPICMV2/DMERGE (lib9290/spicmv2.f90:40-56, dmerge.f90:36-40):
// y = H x with H symmetric and only the strictly-lower off-diagonal part
// stored, plus an FP64 diagonal.  Every stored entry H_ij (i > j) acts twice:
// a gather into row i and the symmetric scatter into row j.
//
// Layout (panel-CSC, built once per matrix on the device): rows are cut into
// panels of R rows; inside a panel the entries are sorted by column, then
// row, so each column the panel touches is one contiguous segment.  Segments
// are padded to a multiple of 4 entries (pad entries: value 0, row of the
// segment's last entry) so that every lane reads whole aligned quads.  Per
// entry the stream carries only the value (fp32 hi, hi+lo or fp64) and a
// 16-bit row offset; per segment its column and start.
//
// Kernel: one workgroup per work item (a slice of one panel's segments).
// The panel's rows of x are staged in LDS (fp32 for the bulk operator) and
// its row accumulators y_i (fp64) live in LDS.  Groups of L lanes walk
// segments; lane l of a group owns entries 4l..4l+3 of each 4L-entry batch
// and applies each entry to all M vectors: one LDS vector read of x_i, M
// FMAs into the lane's column partials, M LDS atomics into y_i.  At a
// segment's end the group reduces its partials (DPP) and issues one M-wide
// atomic add into y_j.  Loads of the next batch, the next segment's x_j and
// the metadata two segments ahead are issued before the current batch is
// consumed.  x and y pass through row-major staging buffers so that the
// x_j reads and the y atomics are line-sized.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>

#include "gci_internal.h"

namespace gci {

namespace {

#define GAS __attribute__((address_space(1)))

constexpr size_t LDS_MAX = 160 * 1024;  // LDS per workgroup on gfx950

typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

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

// segment s of entry e = (exclusive count of flags before e) + flag[e] - 1
__device__ inline int64_t seg_of(const int32_t *flag, const int64_t *segid, int64_t e) {
    return segid[e] + flag[e] - 1;
}

__global__ void k_seg_meta(int64_t nnz, const uint64_t *__restrict__ keys, const int32_t *__restrict__ flag,
                           const int64_t *__restrict__ segid, int32_t *__restrict__ seg_col,
                           int64_t *__restrict__ ustart, int32_t *__restrict__ seg_panel) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[e]) continue;
        const uint64_t k = keys[e];
        const int64_t s = segid[e];
        seg_col[s] = (int32_t)((k >> 16) & 0xFFFFFFull);
        seg_panel[s] = (int32_t)(k >> 40);
        ustart[s] = e;
    }
}

// padded length of every segment (multiple of 4); plen[nseg] = 0
__global__ void k_seg_plen(int64_t nseg, int64_t nnz, const int64_t *__restrict__ ustart, int64_t *__restrict__ plen) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s <= nseg; s += (int64_t)gridDim.x * blockDim.x) {
        if (s == nseg) {
            plen[s] = 0;
        } else {
            const int64_t len = (s + 1 < nseg ? ustart[s + 1] : nnz) - ustart[s];
            plen[s] = (len + 3) & ~(int64_t)3;
        }
    }
}

// Position of a segment's p-th entry in storage.  Quad q (entries 4q..4q+3)
// is stored rotated by q: slot t holds entry 4q + ((t + q) & 3).  The lane
// that loads quad q is lane q mod L of its group, and the kernel applies
// slot j of every lane's quad in the same instruction.  So one instruction's
// lanes of a group touch entries 4l + ((j + l) & 3) -- rows 5 apart, not 4 --
// which spreads their y rows over the LDS banks.
__host__ __device__ inline int64_t quad_slot(int64_t p) {
    const int64_t q = p >> 2, t = p & 3;
    return (q << 2) | ((t - q) & 3);
}

__global__ void k_seg_fill(int64_t nnz, int64_t nseg, const uint64_t *__restrict__ keys,
                           const int32_t *__restrict__ flag, const int64_t *__restrict__ segid,
                           const int64_t *__restrict__ ustart, const int64_t *__restrict__ pstart,
                           const int64_t *__restrict__ perm, const double *__restrict__ val,
                           const float *__restrict__ hi, const float *__restrict__ lo, uint16_t *__restrict__ lrow,
                           double *__restrict__ pval, float *__restrict__ phi, float *__restrict__ plo) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = seg_of(flag, segid, e);
        const int64_t p = e - ustart[s];
        const int64_t pe = pstart[s] + quad_slot(p);
        const uint16_t lr = (uint16_t)(keys[e] & 0xFFFFull);
        lrow[pe] = lr;
        const int64_t src = perm[e];
        if (pval) pval[pe] = val[src];
        if (phi) {
            phi[pe] = hi[src];
            plo[pe] = lo[src];
        }
        // the segment's last entry also writes the pad rows (values stay 0)
        const int64_t uend = s + 1 < nseg ? ustart[s + 1] : nnz;
        if (e == uend - 1)
            for (int64_t q = p + 1; q < pstart[s + 1] - pstart[s]; ++q) lrow[pstart[s] + quad_slot(q)] = lr;
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

// Reinterpret a pointer as a pointer to vector type V in the same address
// space (LDS/generic, or global).
template <class V, class T>
__device__ inline const V *vcast(const T *p) {
    return reinterpret_cast<const V *>(p);
}
template <class V, class T>
__device__ inline const GAS V *vcast(const GAS T *p) {
    return reinterpret_cast<const GAS V *>(p);
}

// M consecutive values of type T (M * sizeof(T) bytes, naturally aligned)
// converted to double; one vector access per 16 bytes.
template <int M, class T, class P>
__device__ inline void load_row(P p, double (&o)[M]) {
    if constexpr (std::is_same<T, float>::value) {
        if constexpr (M == 1) {
            o[0] = (double)p[0];
        } else if constexpr (M == 2) {
            const f2v w = vcast<f2v>(p)[0];
            o[0] = (double)w.x;
            o[1] = (double)w.y;
        } else {
#pragma unroll
            for (int q = 0; q < M / 4; ++q) {
                const f4v w = vcast<f4v>(p)[q];
                o[4 * q + 0] = (double)w.x;
                o[4 * q + 1] = (double)w.y;
                o[4 * q + 2] = (double)w.z;
                o[4 * q + 3] = (double)w.w;
            }
        }
    } else {
        if constexpr (M == 1) {
            o[0] = p[0];
        } else {
#pragma unroll
            for (int q = 0; q < M / 2; ++q) {
                const d2v w = vcast<d2v>(p)[q];
                o[2 * q + 0] = w.x;
                o[2 * q + 1] = w.y;
            }
        }
    }
}

// The matrix stream of one chunk as buffer resources (scalar registers):
// every per-lane access is a 32-bit byte offset, and an offset past the
// chunk reads zeros without touching memory (the hardware range check), so
// a lane with nothing to load needs no clamped address.
typedef int i2v __attribute__((ext_vector_type(2)));
typedef int i4v __attribute__((ext_vector_type(4)));
constexpr int RSRC_WORD3 = 0x00020000;  // raw (untyped) buffer, range-checked
constexpr uint32_t OFF_NONE = 0x7FFFFFF0u;  // an entry offset past any chunk

struct ChunkRsrc {
    __amdgpu_buffer_rsrc_t rows, hi, lo, val;
};

// The base and size are uniform over the workgroup; readfirstlane states it,
// so the descriptor lives in scalar registers.
__device__ inline __amdgpu_buffer_rsrc_t rsrc(const void *p, uint32_t bytes) {
    const uint64_t a = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    void *u = (void *)(((uint64_t)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(u, (short)0, (int)__builtin_amdgcn_readfirstlane(bytes), RSRC_WORD3);
}

// Four consecutive entries of one lane (entry offset `e` from the chunk
// start, a multiple of 4).  settle() marks the registers as read on every
// path (an empty asm statement with them as inputs), so the wait for their
// load is placed before the branch that may skip them and the wait state is
// the same on all paths.
template <int MODE>
struct Quad;

template <>
struct Quad<GCI_OP_SPLIT_BULK> {
    u2v r;
    f4v h;
    __device__ inline void load(const ChunkRsrc &p, uint32_t e) {
        r = __builtin_bit_cast(u2v, __builtin_amdgcn_raw_buffer_load_b64(p.rows, e * 2u, 0, 0));
        h = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(p.hi, e * 4u, 0, 0));
    }
    __device__ inline double value(int j) const { return (double)h[j]; }
    __device__ inline void settle() const { asm volatile("" ::"v"(r), "v"(h)); }
};

template <>
struct Quad<GCI_OP_SPLIT_FULL> {
    u2v r;
    f4v h, l;
    __device__ inline void load(const ChunkRsrc &p, uint32_t e) {
        r = __builtin_bit_cast(u2v, __builtin_amdgcn_raw_buffer_load_b64(p.rows, e * 2u, 0, 0));
        h = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(p.hi, e * 4u, 0, 0));
        l = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(p.lo, e * 4u, 0, 0));
    }
    __device__ inline double value(int j) const { return (double)h[j] + (double)l[j]; }
    __device__ inline void settle() const { asm volatile("" ::"v"(r), "v"(h), "v"(l)); }
};

template <>
struct Quad<GCI_OP_FP64> {
    u2v r;
    d2v a, b;
    __device__ inline void load(const ChunkRsrc &p, uint32_t e) {
        r = __builtin_bit_cast(u2v, __builtin_amdgcn_raw_buffer_load_b64(p.rows, e * 2u, 0, 0));
        a = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(p.val, e * 8u, 0, 0));
        b = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(p.val, e * 8u + 16u, 0, 0));
    }
    __device__ inline double value(int j) const { return j < 2 ? a[j] : b[j - 2]; }
    __device__ inline void settle() const { asm volatile("" ::"v"(r), "v"(a), "v"(b)); }
};

template <int CTRL>
__device__ inline double dpp_double(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over the lanes l, l + S, l + 2S, ... of a 16-lane DPP row (S = M,
// the row holding E = 16/S such lanes): row rotations by S, 2S, ... (DPP
// row_ror).  Every lane of the class gets the total; the pairing order is
// fixed, so the sum is the same on every run.
template <int S, int E>
__device__ inline double slot_sum(double x) {
    if constexpr (E >= 2) x += dpp_double<0x120 + S>(x);
    if constexpr (E >= 4) x += dpp_double<0x120 + 2 * S>(x);
    if constexpr (E >= 8) x += dpp_double<0x120 + 4 * S>(x);
    if constexpr (E >= 16) x += dpp_double<0x120 + 8 * S>(x);
    return x;
}

// One lane's share of a quad: the four entries (rows lr, values a) of its
// slot, each applied to the lane's vector v: y_i[v] += a x_c[v] through an
// LDS atomic, and the column partial acc += a x_i[v].
template <int M, int MODE>
__device__ inline void consume(const Quad<MODE> &q, const typename XT<MODE>::T *xs, double *ys, int v, double xc,
                               double &acc) {
    int lr[4];
    lr[0] = (int)(q.r.x & 0xFFFFu);
    lr[1] = (int)(q.r.x >> 16);
    lr[2] = (int)(q.r.y & 0xFFFFu);
    lr[3] = (int)(q.r.y >> 16);
    double xv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) xv[j] = (double)xs[lr[j] * M + v];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double a = q.value(j);
        acc = fma(a, xv[j], acc);
        __hip_atomic_fetch_add(ys + lr[j] * M + v, a * xc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// raw copy of one row of M values of type T, global -> LDS
template <int M, class T>
__device__ inline void copy_row(T *dst, const GAS T *src) {
    if constexpr (M * sizeof(T) >= 16) {
#pragma unroll
        for (int q = 0; q < (int)(M * sizeof(T) / 16); ++q)
            reinterpret_cast<f4v *>(dst)[q] = vcast<f4v>(src)[q];
    } else if constexpr (M * sizeof(T) == 8) {
        *reinterpret_cast<f2v *>(dst) = vcast<f2v>(src)[0];
    } else {
        *dst = *src;
    }
}

// LDS carve-up of one workgroup (offsets in bytes, 64-byte aligned parts):
// xs (R x M, T) | ys (R x M, fp64) | xcol (NS x M, fp64: a segment's x_j
// row while it runs, then its column sums) | sbeg (NS + 1, int) | scol (NS)
struct LdsMap {
    size_t ys, xcol, sbeg, scol, total;
};
template <int M, class T>
__host__ __device__ constexpr LdsMap lds_map(int R, int NS) {
    LdsMap m{};
    m.ys = ((size_t)R * M * sizeof(T) + 63) & ~(size_t)63;
    m.xcol = (m.ys + (size_t)R * M * sizeof(double) + 63) & ~(size_t)63;
    m.sbeg = m.xcol + (size_t)NS * M * sizeof(double);
    m.scol = m.sbeg + (((size_t)(NS + 1) * sizeof(int) + 15) & ~(size_t)15);
    m.total = m.scol + (size_t)NS * sizeof(int);
    return m;
}

// A group's walk over its segments in batches of E quads: segment i (chunk
// local), batch start k, segment end e1.
struct Cursor {
    int i, k, e1;
    template <int E, int G>
    __device__ inline bool advance(const int *sbeg, int nsc) {  // returns true at a segment end
        if (k + 4 * E < e1) {
            k += 4 * E;
            return false;
        }
        i += G;
        if (i < nsc) {
            k = sbeg[i];
            e1 = sbeg[i + 1];
        }
        return true;
    }
};

// One turn of the prefetch ring: D + 1 steps, step J consuming register set
// J while the load D batches ahead goes into set (J + D) mod (D + 1).  The
// sets are indexed by constants only, so they stay in registers.
template <int J, int D, class Step, class Q>
__device__ inline void ring_steps(Step &step, Q (&q)[D + 1], bool (&ok)[D + 1]) {
    step(q[J], ok[J], q[(J + D) % (D + 1)], ok[(J + D) % (D + 1)]);
    if constexpr (J < D) ring_steps<J + 1, D>(step, q, ok);
}

// The SpMM work-item kernel.  A group of L lanes (one DPP row or half of
// one) holds E = L/M entry slots x M vectors: lane v + M*e loads quad e of
// the group's batch (the M lanes of a slot load the same address) and
// applies its four entries to vector v.  So one LDS instruction covers E
// entries x M vectors, each entry's M lanes on one contiguous y row.
template <int M, int MODE, int L, int BS, int D>
__global__ __launch_bounds__(BS) void k_pcsc(const SpmmItem *__restrict__ items, const SpmmJob *__restrict__ jobs,
                                             int R, int NS) {
    using T = typename XT<MODE>::T;
    static_assert(L % M == 0 && L <= 16, "a group is E slots x M vectors inside one DPP row");
    constexpr int E = L / M;   // entry slots per group
    constexpr int G = BS / L;  // lane groups per workgroup
    extern __shared__ __align__(64) char smem[];
    const SpmmItem it = items[blockIdx.x];
    const SpmmJob jb = jobs[it.job];
    if (!jb.active || jb.mode != MODE) return;
    const int rows = it.rows;
    if (rows > R) return;  // never reached: items are cut for the shape's R
    const LdsMap lm = lds_map<M, T>(R, NS);
    T *xs = reinterpret_cast<T *>(smem);
    double *ys = reinterpret_cast<double *>(smem + lm.ys);
    double *xcol = reinterpret_cast<double *>(smem + lm.xcol);
    int *sbeg = reinterpret_cast<int *>(smem + lm.sbeg);
    int *scl = reinterpret_cast<int *>(smem + lm.scol);
    const GAS T *gx = (const GAS T *)jb.xs;
    const GAS int32_t *gcol = (const GAS int32_t *)jb.seg_col;
    const GAS int64_t *gst = (const GAS int64_t *)jb.seg_start;
    double *gy = jb.ys;
    const int t = threadIdx.x;
    const int g = t / L, r = t % L, e = r / M, v = r % M;
    // the panel's x rows and zeroed y rows stay for the whole item
    for (int rr = t; rr < rows; rr += BS) {
        copy_row<M, T>(xs + rr * M, gx + (it.row0 + rr) * M);
#pragma unroll
        for (int u = 0; u < M; ++u) ys[rr * M + u] = 0.0;
    }
    // the item's segments in chunks of at most NS (their tables fill the LDS)
    for (int64_t cs = it.seg0; cs < it.seg1; cs += NS) {
        const int nsc = (int)min((int64_t)NS, it.seg1 - cs);
        const int64_t base = gst[cs];  // first entry of the chunk (a multiple of 4)
        for (int i = t; i <= nsc; i += BS) {
            sbeg[i] = (int)(gst[cs + i] - base);
            if (i < nsc) {
                const int c = gcol[cs + i];
                scl[i] = c;
                double xr[M];
                load_row<M, T>(gx + (int64_t)c * M, xr);
#pragma unroll
                for (int u = 0; u < M; ++u) xcol[i * M + u] = xr[u];
            }
        }
        const uint32_t cn = (uint32_t)(gst[cs + nsc] - base);  // entries in the chunk
        ChunkRsrc P;
        P.rows = rsrc(jb.lrow + base, cn * 2u);
        P.hi = rsrc(jb.hi + base, cn * 4u);
        P.lo = rsrc(jb.lo + base, cn * 4u);
        P.val = rsrc(jb.val + base, cn * 8u);
        __syncthreads();

        // Main loop.  The load cursor runs D batches ahead of the consume
        // cursor through a ring of D + 1 register sets, so D batches stay in
        // flight while one is consumed.  Every step issues its load (out of
        // range, reading zeros, once the group is done), so all paths carry
        // the same loads and the same waits; the only global accesses in
        // the loop are these loads.
        Cursor cc{g, 0, 0}, lc{g, 0, 0};
        if (g < nsc) {
            cc.k = lc.k = sbeg[g];
            cc.e1 = lc.e1 = sbeg[g + 1];
        }
        double acc = 0.0, xc = g < nsc ? xcol[g * M + v] : 0.0;
        Quad<MODE> q[D + 1];
        bool ok[D + 1];
        auto issue = [&](Quad<MODE> &dst, bool &dok) {
            dok = lc.i < nsc && lc.k + 4 * e < lc.e1;
            dst.load(P, dok ? (uint32_t)(lc.k + 4 * e) : OFF_NONE);
            if (lc.i < nsc) lc.advance<E, G>(sbeg, nsc);
        };
#pragma unroll
        for (int d = 0; d < D; ++d) issue(q[d], ok[d]);
        auto step = [&](Quad<MODE> &cur, bool cur_ok, Quad<MODE> &nxt, bool &nxt_ok) {
            issue(nxt, nxt_ok);
            cur.settle();
            if (cur_ok) consume<M, MODE>(cur, xs, ys, v, xc, acc);
            if (cc.i < nsc) {
                const int ci = cc.i;
                if (cc.advance<E, G>(sbeg, nsc)) {
                    // segment done: its column sums replace its x_j row
                    acc = slot_sum<M, E>(acc);
                    if (e == 0) xcol[ci * M + v] = acc;
                    acc = 0.0;
                    if (cc.i < nsc) xc = xcol[cc.i * M + v];
                }
            }
        };
        while (cc.i < nsc) ring_steps<0, D>(step, q, ok);
#pragma unroll
        for (int d = 0; d <= D; ++d) q[d].settle();
        __syncthreads();
        // the chunk's column sums, as coalesced atomics
        for (int idx = t; idx < nsc * M; idx += BS) {
            const double y = xcol[idx];
            const int u = idx % M;
            if (y != 0.0 && u < jb.m) unsafeAtomicAdd(&gy[(int64_t)scl[idx / M] * M + u], y);
        }
        __syncthreads();  // the next chunk overwrites the tables
    }
    for (int idx = t; idx < rows * M; idx += BS) {
        const double y = ys[idx];
        if (y != 0.0 && idx % M < jb.m) unsafeAtomicAdd(&gy[it.row0 * M + idx], y);
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

// ------------------------------------------------------ launch shapes

SpmmShape g_shape;         // shape all launches use
bool g_shape_set = false;
uint64_t g_shape_version = 1;

template <int M, int MODE, int L, int BS, int D>
void launch_depth(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, const SpmmShape &sh,
                  hipStream_t s) {
    using T = typename XT<MODE>::T;
    const size_t shm = lds_map<M, T>(sh.R, sh.nseg).total;
    GCI_REQUIRE(shm <= LDS_MAX, "spmm: panel height and segments per chunk exceed the LDS");
    static size_t attr = 0;  // one instantiation per template tuple
    if (shm > attr) {
        GCI_HIP(hipFuncSetAttribute((const void *)k_pcsc<M, MODE, L, BS, D>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        attr = shm;
    }
    hipLaunchKernelGGL((k_pcsc<M, MODE, L, BS, D>), dim3((unsigned)nitems), dim3(BS), shm, s, d_items, d_jobs,
                       sh.R, sh.nseg);
}

template <int M, int MODE, int L, int BS>
void launch_shape(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, const SpmmShape &sh,
                  hipStream_t s) {
    switch (sh.depth) {
        case 3: launch_depth<M, MODE, L, BS, 3>(d_items, nitems, d_jobs, sh, s); break;
        case 4: launch_depth<M, MODE, L, BS, 4>(d_items, nitems, d_jobs, sh, s); break;
        default: launch_depth<M, MODE, L, BS, 6>(d_items, nitems, d_jobs, sh, s); break;
    }
}

template <int M, int MODE, int L>
void launch_bs(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, const SpmmShape &sh, hipStream_t s) {
    if (sh.block == 256) launch_shape<M, MODE, L, 256>(d_items, nitems, d_jobs, sh, s);
    else launch_shape<M, MODE, L, 512>(d_items, nitems, d_jobs, sh, s);
}

template <int M, int MODE>
void launch_l(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, const SpmmShape &sh, hipStream_t s) {
    if (sh.L == 8) launch_bs<M, MODE, 8>(d_items, nitems, d_jobs, sh, s);
    else launch_bs<M, MODE, 16>(d_items, nitems, d_jobs, sh, s);
}

template <int MODE>
void launch_mode(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, int M, const SpmmShape &sh,
                 hipStream_t s) {
    switch (M) {
        case 1: launch_l<1, MODE>(d_items, nitems, d_jobs, sh, s); break;
        case 2: launch_l<2, MODE>(d_items, nitems, d_jobs, sh, s); break;
        case 4: launch_l<4, MODE>(d_items, nitems, d_jobs, sh, s); break;
        default: launch_l<8, MODE>(d_items, nitems, d_jobs, sh, s); break;
    }
}

int env_int(const char *name, int dflt) {
    const char *v = std::getenv(name);
    return v && *v ? std::atoi(v) : dflt;
}

}  // namespace

bool valid_shape(const SpmmShape &sh) {
    // the instantiated launch shapes: L in {8, 16}, 256 or 512 threads, depth 3, 4 or 6
    return sh.R >= 64 && sh.R <= 4096 && (sh.R & (sh.R - 1)) == 0 && (sh.L == 8 || sh.L == 16) &&
           (sh.block == 256 || sh.block == 512) && sh.target >= 4096 && sh.nseg >= 64 && sh.nseg <= 4096 &&
           (sh.depth == 3 || sh.depth == 4 || sh.depth == 6) && lds_map<1, double>(sh.R, sh.nseg).total <= LDS_MAX;
}

SpmmShape spmm_shape() {
    if (!g_shape_set) {
        SpmmShape sh;
        sh.R = env_int("GCI_SPMM_R", sh.R);
        sh.L = env_int("GCI_SPMM_L", sh.L);
        sh.block = env_int("GCI_SPMM_BLOCK", sh.block);
        sh.target = env_int("GCI_SPMM_TARGET", (int)sh.target);
        sh.nseg = env_int("GCI_SPMM_NSEG", sh.nseg);
        sh.depth = env_int("GCI_SPMM_DEPTH", sh.depth);
        GCI_REQUIRE(valid_shape(sh), "spmm: invalid launch shape");
        g_shape = sh;
        g_shape_set = true;
    }
    return g_shape;
}

void set_spmm_shape(const SpmmShape &sh) {
    GCI_REQUIRE(valid_shape(sh), "spmm: invalid launch shape");
    g_shape = sh;
    g_shape_set = true;
    ++g_shape_version;
}

uint64_t spmm_shape_version() { return g_shape_version; }

int spmm_max_width(int mode) {
    const SpmmShape sh = spmm_shape();
    for (int w = 8; w > 1; w /= 2)
        if (spmm_lds_bytes(sh, w, mode) <= LDS_MAX) return w;
    return 1;
}

int spmm_width(int m) { return m <= 1 ? 1 : m <= 2 ? 2 : m <= 4 ? 4 : 8; }

void spmm_items(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, int M, int mode, hipStream_t s) {
    if (nitems <= 0) return;
    const SpmmShape sh = spmm_shape();
    switch (mode) {
        case GCI_OP_FP64: launch_mode<GCI_OP_FP64>(d_items, nitems, d_jobs, M, sh, s); break;
        case GCI_OP_SPLIT_BULK: launch_mode<GCI_OP_SPLIT_BULK>(d_items, nitems, d_jobs, M, sh, s); break;
        default: launch_mode<GCI_OP_SPLIT_FULL>(d_items, nitems, d_jobs, M, sh, s); break;
    }
}

void build_items(HMat &h, int64_t target) {
    Pcsc &pc = h.pc;
    pc.items.clear();
    if (pc.nseg == 0) return;
    const int R = pc.R;
    for (int64_t s0 = 0, s = 1; s <= pc.nseg; ++s) {
        if (s == pc.nseg || pc.seg_panel_h[s] != pc.seg_panel_h[s0] ||
            pc.seg_start_h[s] - pc.seg_start_h[s0] >= target) {
            const int64_t row0 = (int64_t)pc.seg_panel_h[s0] * R;
            pc.items.push_back(SpmmItem{s0, s, row0, (int32_t)std::min<int64_t>(R, h.n - row0), 0});
            s0 = s;
        }
    }
    pc.target = target;
}

void recut_items(HMat &h) {
    const SpmmShape sh = spmm_shape();
    GCI_REQUIRE(h.pc.R == sh.R, "spmm: panel height changed while the matrix is in use");
    if (h.pc.target != sh.target) {
        build_items(h, sh.target);
        h.pc.d_items.release();
    }
}

void ensure_items(HMat &h) {
    if (h.pc.R != 0 && h.pc.R != spmm_shape().R) {
        GCI_HIP(hipDeviceSynchronize());  // queued launches may still read the old layout
        h.pc = Pcsc();
    }
    ensure_pcsc(h);
    recut_items(h);
}

size_t spmm_lds_bytes(const SpmmShape &sh, int M, int mode) {
    if (mode == GCI_OP_SPLIT_BULK) {
        switch (M) {
            case 1: return lds_map<1, float>(sh.R, sh.nseg).total;
            case 2: return lds_map<2, float>(sh.R, sh.nseg).total;
            case 4: return lds_map<4, float>(sh.R, sh.nseg).total;
            default: return lds_map<8, float>(sh.R, sh.nseg).total;
        }
    }
    switch (M) {
        case 1: return lds_map<1, double>(sh.R, sh.nseg).total;
        case 2: return lds_map<2, double>(sh.R, sh.nseg).total;
        case 4: return lds_map<4, double>(sh.R, sh.nseg).total;
        default: return lds_map<8, double>(sh.R, sh.nseg).total;
    }
}

size_t spmm_lds_max() { return LDS_MAX; }

void ensure_pcsc(HMat &h) {
    if (h.pc.R) return;
    const SpmmShape sh = spmm_shape();
    const int R = sh.R;
    Pcsc &pc = h.pc;
    const int64_t nnz = h.nnz, n = h.n;
    hipStream_t s = h.stream;
    pc.nnz = nnz;
    pc.nseg = 0;
    pc.npad = 0;
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
        DBuf<int64_t> segid(nnz);
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
        const int64_t nseg = last + lastf;
        pc.nseg = nseg;
        pc.seg_col.alloc(nseg);
        DBuf<int64_t> ustart(nseg);
        DBuf<int32_t> spanel(nseg);
        hipLaunchKernelGGL(k_seg_meta, dim3((unsigned)eb), dim3(256), 0, s, nnz, k1.p, flag.p, segid.p,
                           pc.seg_col.p, ustart.p, spanel.p);
        GCI_HIP(hipGetLastError());
        DBuf<int64_t> plen(nseg + 1);
        const int64_t sb = std::min<int64_t>((nseg + 256) / 256, 65536);
        hipLaunchKernelGGL(k_seg_plen, dim3((unsigned)sb), dim3(256), 0, s, nseg, nnz, ustart.p, plen.p);
        GCI_HIP(hipGetLastError());
        pc.seg_start.alloc(nseg + 1);
        tb = 0;
        GCI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, plen.p, pc.seg_start.p, nseg + 1, s));
        {
            DBuf<char> tmp(tb);
            GCI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, plen.p, pc.seg_start.p, nseg + 1, s));
        }
        pc.seg_start_h.resize(nseg + 1);
        pc.seg_panel_h.resize(nseg);
        GCI_HIP(hipMemcpyAsync(pc.seg_start_h.data(), pc.seg_start.p, (nseg + 1) * sizeof(int64_t),
                               hipMemcpyDeviceToHost, s));
        GCI_HIP(hipMemcpyAsync(pc.seg_panel_h.data(), spanel.p, nseg * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipStreamSynchronize(s));
        const int64_t npad = pc.seg_start_h[nseg];
        pc.npad = npad;
        pc.lrow.alloc(npad);
        if (h.val.p) pc.val.alloc(npad);
        if (h.precision == GCI_SPLIT32) {
            pc.hi.alloc(npad);
            pc.lo.alloc(npad);
        }
        // pad entries keep value 0
        if (pc.val.p) GCI_HIP(hipMemsetAsync(pc.val.p, 0, npad * sizeof(double), s));
        if (pc.hi.p) {
            GCI_HIP(hipMemsetAsync(pc.hi.p, 0, npad * sizeof(float), s));
            GCI_HIP(hipMemsetAsync(pc.lo.p, 0, npad * sizeof(float), s));
        }
        hipLaunchKernelGGL(k_seg_fill, dim3((unsigned)eb), dim3(256), 0, s, nnz, nseg, k1.p, flag.p, segid.p,
                           ustart.p, pc.seg_start.p, i1.p, h.val.p, h.hi.p, h.lo.p, pc.lrow.p, pc.val.p, pc.hi.p,
                           pc.lo.p);
        GCI_HIP(hipGetLastError());
        GCI_HIP(hipStreamSynchronize(s));
    }
    pc.R = R;
    build_items(h, sh.target);
}

void spmm(HMat &h, int m, const double *x, int ldx, double *y, int ldy, int mode, hipStream_t s) {
    GCI_REQUIRE(m >= 1, "spmm: m < 1");
    GCI_REQUIRE(ldx >= h.n && ldy >= h.n, "spmm: leading dimension < n");
    if (mode != GCI_OP_FP64)
        GCI_REQUIRE(h.precision == GCI_SPLIT32, "spmm: split operator on an FP64-only matrix");
    else
        GCI_REQUIRE(h.val.p != nullptr, "spmm: FP64 operator on a matrix without FP64 values");
    if (h.n == 0) return;
    ensure_items(h);
    Pcsc &pc = h.pc;
    // staging buffers, the job table (one descriptor per operator mode and
    // column count) and the device copy of the work items stay with the
    // matrix; they are (re)written only when they grow, after the stream
    // has drained, so queued launches never see them change
    const int W = spmm_max_width(mode);  // widest launch whose LDS fits the shape
    const int Mmax = spmm_width(std::min(W, m));
    if (pc.stage_M < Mmax) {
        GCI_HIP(hipStreamSynchronize(s));
        pc.stage_x.alloc((size_t)h.n * Mmax * sizeof(double));
        pc.stage_y.alloc((size_t)h.n * Mmax);
        pc.stage_M = Mmax;
        std::vector<SpmmJob> jt(3 * 8);
        for (int md = 0; md < 3; ++md)
            for (int mc = 1; mc <= 8; ++mc)
                jt[md * 8 + mc - 1] = SpmmJob{pc.seg_col.p, pc.seg_start.p, pc.lrow.p, pc.val.p, pc.hi.p, pc.lo.p,
                                              pc.stage_x.p, pc.stage_y.p, mc, md, 1, 0};
        pc.d_jobs.alloc(jt.size());
        GCI_HIP(hipMemcpy(pc.d_jobs.p, jt.data(), jt.size() * sizeof(SpmmJob), hipMemcpyHostToDevice));
    }
    if (!pc.items.empty() && pc.d_items.n < pc.items.size()) {
        GCI_HIP(hipStreamSynchronize(s));
        pc.d_items.alloc(pc.items.size());
        GCI_HIP(hipMemcpy(pc.d_items.p, pc.items.data(), pc.items.size() * sizeof(SpmmItem),
                          hipMemcpyHostToDevice));
    }
    for (int c0 = 0; c0 < m; c0 += W) {
        const int mc = std::min(W, m - c0);
        const int M = spmm_width(mc);
        const double *xc = x + (int64_t)c0 * ldx;
        double *yc = y + (int64_t)c0 * ldy;
        const int64_t pb = std::min<int64_t>((h.n * M + 255) / 256, 65536);
        if (mode == GCI_OP_SPLIT_BULK)
            hipLaunchKernelGGL(k_pack<float>, dim3((unsigned)pb), dim3(256), 0, s, h.n, mc, M, xc, (int64_t)ldx,
                               (float *)pc.stage_x.p, pc.stage_y.p);
        else
            hipLaunchKernelGGL(k_pack<double>, dim3((unsigned)pb), dim3(256), 0, s, h.n, mc, M, xc, (int64_t)ldx,
                               (double *)pc.stage_x.p, pc.stage_y.p);
        GCI_HIP(hipGetLastError());
        // items carry job 0; the table offset selects the descriptor
        spmm_items(pc.d_items.p, (int64_t)pc.items.size(), pc.d_jobs.p + mode * 8 + mc - 1, M, mode, s);
        GCI_HIP(hipGetLastError());
        const int64_t ub = std::min<int64_t>((h.n * mc + 255) / 256, 65536);
        hipLaunchKernelGGL(k_unpack, dim3((unsigned)ub), dim3(256), 0, s, h.n, mc, M, (const double *)pc.stage_y.p,
                           (const double *)h.diag.p, xc, (int64_t)ldx, yc, (int64_t)ldy);
        GCI_HIP(hipGetLastError());
    }
}

}  // namespace gci
