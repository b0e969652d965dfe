// spmm.hip -- symmetric sparse matrix x block-of-vectors for CI Hamiltonians.
//
// Replaces SPICMV2/DMERGE (lib9290/spicmv2.f90:40-56, dmerge.f90:36-40):
// y = H x with H symmetric and only the strictly-lower off-diagonal part
// stored, plus an FP64 diagonal.  Every stored entry H_ij (i > j) acts twice:
// y_i += H_ij x_j, and the symmetric y_j += H_ij x_i.
//
// Layout (panel slices, built once per matrix on the device).  Rows are cut
// into panels of R rows; inside a panel the entries of one column form a
// segment, and segments are dealt in groups of 64 -- one per wavefront lane
// -- into slices.  A slice is stored step-major: step k holds entries
// 4k..4k+3 of each of its 64 segments, so a step is 256 entries that a
// wavefront reads with one 16-byte load per lane (fp32 values) and one small
// load of row codes.  A slice has as many steps as its longest segment
// needs; the other lanes are padded with zero-valued entries.  Per slice the
// layout keeps its 64 columns and first step.
//
// Full layouts (split-precision matrices, and their FP64 copy): both
// triangles, so every entry H_ij appears as row i of column j and as row j
// of column i.  Segments keep row order and code each row in two bits: a
// step of 1..3 rows from the lane's previous row, or 0 for an escape entry,
// whose value slot holds the next row as an integer and which contributes
// nothing (a gap above 3 rows takes one).  A lane's four codes of a step are
// one byte.  Per slice and lane the layout keeps the segment's first row.
// Columns are dealt into slices in column order.
//
// Lower-triangle layouts (FP64 matrices): the strictly-lower entries once,
// with 16-bit row codes (the LDS byte offset (row - panel start) x 8).  A
// panel's segments are sorted by length, longest first, and inside a segment
// the entries are ordered by (row mod 32, row), lane l's segment beginning
// at its first entry of row class >= l mod 32: lanes of one slice have
// segments of nearly equal length, so at every step their rows fall in
// nearly distinct classes and the 64 LDS addresses of one row gather or
// scatter instruction spread over the banks.
//
// Kernel: one workgroup per work item.  A full layout's item is at most NI
// whole slices of one panel: the panel's x rows (fp32 for the bulk operator,
// fp64 otherwise, row-major) and the slice tables are staged in LDS once,
// then each wavefront walks its share of the slices (cut at slice starts)
// with the next `depth` steps' loads in flight.  Lane l keeps its column's
// partial sums sum_i H_ij x_i in registers, reading x_i from LDS at its
// running row; at a slice end they leave as global atomics -- the column
// sums are y, the layout holding both triangles.  A lower-triangle item is a
// run of one panel's steps, walked in chunks of NS slices: besides the
// column partials every entry adds H_ij x_j into y_i through an LDS atomic,
// and the panel's rows leave at the item's end.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "gci_internal.h"

namespace gci {

namespace {

#define GAS __attribute__((address_space(1)))

constexpr size_t LDS_MAX = 160 * 1024;  // LDS per workgroup on gfx950
constexpr int SW = 64;                  // lanes (segments) per slice
constexpr int SE = SW * 4;              // entries per step
constexpr int ITEM_SLICES_MAX = 128;    // most slices in a full layout's work item

typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

// ------------------------------------------------------ panel-slice build

// Row code of the sort key: orders a segment's rows by (row mod 32, row).
// Valid for rows below 65536.
__host__ __device__ inline uint32_t row_code(uint32_t lr) { return ((lr & 31u) << 11) | (lr >> 5); }
__host__ __device__ inline uint32_t code_row(uint32_t c) { return ((c & 0x7FFu) << 5) | (c >> 11); }

// Sort keys of the laid-out entries: entry e = (i, j) of the lower triangle
// as row i of column j; with full, also as row j of column i (key nnz + e).
__global__ void k_keys(int64_t n, int64_t nnz, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                       int R, int full, uint64_t *__restrict__ keys, uint32_t *__restrict__ idx) {
    // one wave per row
    const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const uint64_t panel = (uint64_t)(row / R), lr = (uint64_t)(row % R);
    for (int64_t e = rp[row] + lane; e < rp[row + 1]; e += 64) {
        const uint64_t c = (uint64_t)col[e];
        keys[e] = (panel << 40) | (c << 16) | row_code((uint32_t)lr);
        idx[e] = (uint32_t)e;
        if (full) {
            keys[nnz + e] = ((c / R) << 40) | ((uint64_t)row << 16) | row_code((uint32_t)(c % R));
            idx[nnz + e] = (uint32_t)e;
        }
    }
}

__global__ void k_seg_flags(int64_t nnz, const uint64_t *__restrict__ keys, int32_t *__restrict__ flag) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
        flag[e] = (e == 0 || (keys[e] >> 16) != (keys[e - 1] >> 16)) ? 1 : 0;
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

// segment s of entry e = (exclusive count of flags before e) + flag[e] - 1
__device__ inline int64_t seg_of(const int32_t *flag, const int64_t *segid, int64_t e) {
    return segid[e] + flag[e] - 1;
}

constexpr int SLACK = 4;  // positions a slice's schedule may add to its longest segment
// positions the schedule of a slice whose longest segment has `len` entries
// can use (a multiple of 4 at most len + SLACK + 3)
__host__ __device__ inline int sched_cap(int len) { return len + SLACK + 3; }

// per segment: its length and the key that sorts a panel's segments longest
// first (ties by column)
__global__ void k_seg_order(int64_t nseg, int64_t nnz, const int64_t *__restrict__ ustart,
                            const int32_t *__restrict__ seg_panel, int32_t *__restrict__ ulen,
                            uint64_t *__restrict__ okey, int64_t *__restrict__ oval) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nseg; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t len = (s + 1 < nseg ? ustart[s + 1] : nnz) - ustart[s];
        ulen[s] = (int32_t)len;
        okey[s] = ((uint64_t)seg_panel[s] << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)len);
        oval[s] = s;
    }
}

// per panel p (and p = npanel): its first segment (segments are sorted by
// panel: a binary search) and its slice count, ceil(segments / 64)
__global__ void k_panel_slices(int64_t npanel, int64_t nseg, const int32_t *__restrict__ seg_panel,
                               int64_t *__restrict__ pfirst, int64_t *__restrict__ pslices) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p <= npanel;
         p += (int64_t)gridDim.x * blockDim.x) {
        auto first = [&](int64_t q) {  // first segment of panel q or later
            int64_t lo = 0, hi = nseg;
            while (lo < hi) {
                const int64_t mid = (lo + hi) / 2;
                if (seg_panel[mid] < q) lo = mid + 1;
                else hi = mid;
            }
            return lo;
        };
        const int64_t f = first(p);
        pfirst[p] = f;
        pslices[p] = p < npanel ? (first(p + 1) - f + SW - 1) / SW : 0;
    }
}

// Deals the sorted segments into slices: the segment of rank r inside its
// panel goes to slice (panel's first slice) + r / 64, lane r % 64.  Lane 0
// holds the slice's longest segment and sizes its schedule buffer.
__global__ void k_seg_place(int64_t nseg, const int64_t *__restrict__ order, const int32_t *__restrict__ seg_panel,
                            const int32_t *__restrict__ seg_col, const int32_t *__restrict__ ulen,
                            const int64_t *__restrict__ pfirst, const int64_t *__restrict__ pslice0,
                            int32_t *__restrict__ scol, int64_t *__restrict__ slane_seg,
                            int32_t *__restrict__ slice_panel, int64_t *__restrict__ seg_slice,
                            int32_t *__restrict__ seg_lane, int64_t *__restrict__ slice_pcap) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nseg; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = order[r];
        const int p = seg_panel[s];
        const int64_t q = r - pfirst[p];
        const int64_t sl = pslice0[p] + q / SW;
        const int lane = (int)(q % SW);
        scol[sl * SW + lane] = seg_col[s];
        slane_seg[sl * SW + lane] = s;
        seg_slice[s] = sl;
        seg_lane[s] = lane;
        if (lane == 0) {
            slice_panel[sl] = p;
            slice_pcap[sl] = (int64_t)SW * sched_cap(ulen[s]);
        }
    }
}

// Stored position of entry t (0..3) of lane `lane` in step `step`: row codes
// and fp32 values lane-major (4 per lane); fp64 values as two halves of the
// step, entries 0-1 of all lanes, then entries 2-3, so each half is one
// 16-byte load per lane.
__host__ __device__ inline int64_t slot_pos(int64_t step, int lane, int t) { return step * SE + lane * 4 + t; }
__host__ __device__ inline int64_t val_pos(int64_t step, int lane, int t) {
    return step * SE + (t >> 1) * (SE / 2) + lane * 2 + (t & 1);
}

// Bank schedule of one slice.  An fp64 LDS access of row r falls on the bank
// pair r mod 32 (the row class), and a 64-lane access is served as two
// 32-lane halves.  One step position p is one LDS instruction over the 64
// lanes; it is conflict-free when no two lanes of a half share a class.  Per
// position, every lane with entries left asks for the class it has most
// entries of; in each half a class goes to the asker with most entries left
// (ties to the lower lane), and lanes turned away ask for their next class
// still free, for up to NROUND rounds.  A lane still without a place whose
// entries left would run past the slice's target length (its longest
// segment + SLACK positions) takes its best class anyway, sharing a bank pair;
// the other lanes without a place, and lanes whose segment is done, take pad
// entries (value 0) of the classes left free.  So the slice grows by at most
// SLACK positions, and most instructions are conflict-free.  The schedule
// records each entry's position in its slice and the class of every pad; the
// stream is written from these once the slices' first steps are known.
constexpr int NCLS = 32;
constexpr int NROUND = 4;

__global__ __launch_bounds__(64) void k_sched(const int64_t *__restrict__ slane_seg, const int64_t *__restrict__ ustart,
                                              const int32_t *__restrict__ ulen, const uint64_t *__restrict__ keys,
                                              const int32_t *__restrict__ slice_panel, int R, int64_t n,
                                              const int64_t *__restrict__ padoff, int64_t *__restrict__ slice_steps,
                                              int32_t *__restrict__ qpos, uint8_t *__restrict__ padcls,
                                              int *__restrict__ overflow) {
    __shared__ uint16_t nxt[NCLS][SW];  // next entry (index in the segment) per class
    __shared__ uint32_t best[SW];       // per half and class: the largest asking key this position
    __shared__ uint32_t tmask[2];       // per half: classes taken this position
    const int64_t sl = blockIdx.x;
    const int lane = threadIdx.x, half = lane & 32, hl = lane & 31, hi2 = lane >> 5;
    const uint32_t below = (hl == 0) ? 0u : (0xFFFFFFFFu >> (32 - hl));  // lanes below this one in its half
    const int64_t s = slane_seg[sl * SW + lane];
    const int len = s >= 0 ? ulen[s] : 0;
    const int64_t e0 = s >= 0 ? ustart[s] : 0;
    // entries left per class, in registers (all accesses below are unrolled)
    int cnt[NCLS];
#pragma unroll
    for (int c = 0; c < NCLS; ++c) cnt[c] = 0;
    for (int i = 0; i < len; ++i) {
        const int ci = (int)((keys[e0 + i] & 0xFFFFull) >> 11);  // entries are sorted by class
#pragma unroll
        for (int c = 0; c < NCLS; ++c) cnt[c] += (c == ci);
    }
    {
        int acc = 0;
#pragma unroll
        for (int c = 0; c < NCLS; ++c) {
            nxt[c][lane] = (uint16_t)acc;
            acc += cnt[c];
        }
    }
    const int64_t prow = min((int64_t)R, n - (int64_t)slice_panel[sl] * R);
    const int ncls = (int)min((int64_t)NCLS, prow);  // classes present in the panel
    const uint32_t present = ncls >= 32 ? 0xFFFFFFFFu : ((1u << ncls) - 1u);
    int rem = len;
    int lmax = len;  // the slice's longest segment
    for (int o = 32; o > 0; o >>= 1) lmax = max(lmax, __shfl_xor(lmax, o));
    const int target = lmax + SLACK, cap = sched_cap(lmax);
    uint8_t *pads = padcls + padoff[sl] + (int64_t)lane * cap;  // this lane's pad classes
    // the class with entries left, not in `excl`, with most entries; ties to
    // the first in the lane's rotated order starting at class lane + p
    auto pick = [&](uint32_t excl, int p) {
        int bk = -1, bc = -1;
#pragma unroll
        for (int c = 0; c < NCLS; ++c) {
            const int k = ((excl >> c) & 1u) || cnt[c] == 0 ? -1 : (cnt[c] << 5) | (31 - ((c - lane - p) & 31));
            if (k > bk) {
                bk = k;
                bc = c;
            }
        }
        return bc;
    };
    int p = 0;
    // positions until every lane is done, rounded up to whole steps (at most cap)
    for (; p < cap && (__ballot(rem > 0) != 0 || (p & 3) != 0); ++p) {
        best[lane] = 0u;
        if (lane < 2) tmask[lane] = ~present;
        int choice = -1;
        int cand = rem > 0 ? pick(0u, p) : -1;
        __syncthreads();
        // Rounds: each asking lane offers its key (entries left, then lower
        // lane first) to its class in its half; the largest key takes it.  A
        // class asked for is always taken, so later rounds ask only for
        // classes not yet taken and the keys left in best[] never mislead.
        for (int r = 0; r < NROUND; ++r) {
            const bool asking = cand >= 0 && choice < 0;
            if (__ballot(asking) == 0) break;
            const uint32_t key = ((uint32_t)rem << 6) | (uint32_t)(63 - lane);
            if (asking) atomicMax(&best[half + cand], key);
            __syncthreads();
            if (asking && best[half + cand] == key) {
                choice = cand;
                atomicOr(&tmask[hi2], 1u << cand);
            }
            __syncthreads();
            if (cand >= 0 && choice < 0) cand = pick(tmask[hi2], p);
            __syncthreads();
        }
        // critical lanes: their best class, free if one is, else shared
        const uint32_t tm = tmask[hi2];
        if (choice < 0 && rem > 0 && rem >= target - p) {
            const int cf = pick(tm, p);
            choice = cf >= 0 ? cf : pick(0u, p);
        }
        __syncthreads();
        if (choice >= 0) atomicOr(&tmask[hi2], 1u << choice);
        __syncthreads();
        // pads: the lanes of a half without an entry, in lane order, take its
        // free classes in class order
        const uint64_t padm = __ballot(choice < 0);
        if (choice < 0) {
            uint32_t f = ~tmask[hi2];
            const int rank = __popc((uint32_t)(padm >> half) & below);
            int padc = (lane + p) % ncls;  // no class left: any class
            if (rank < __popc(f)) {
                for (int k = 0; k < rank; ++k) f &= f - 1;
                padc = __ffs((int)f) - 1;
            }
            pads[p] = (uint8_t)padc;
        } else {
            const int idx = nxt[choice][lane];
            nxt[choice][lane] = (uint16_t)(idx + 1);
#pragma unroll
            for (int c = 0; c < NCLS; ++c) cnt[c] -= (c == choice);
            --rem;
            qpos[e0 + idx] = p;
        }
        __syncthreads();  // best[] and tmask[] are reset for the next position
    }
    if (lane == 0) slice_steps[sl] = p / 4;
    if (rem > 0) atomicAdd(overflow, 1);  // never: critical lanes finish by the target length
}

// The stream's entries at their scheduled positions: row codes and values.
__global__ void k_sell_fill(int64_t nnz, const uint64_t *__restrict__ keys, const int32_t *__restrict__ flag,
                            const int64_t *__restrict__ segid, const int64_t *__restrict__ seg_slice,
                            const int32_t *__restrict__ seg_lane, const int32_t *__restrict__ qpos,
                            const int64_t *__restrict__ sstep, const uint32_t *__restrict__ perm,
                            const double *__restrict__ val, const float *__restrict__ hi,
                            const float *__restrict__ lo, uint16_t *__restrict__ lrow, double *__restrict__ pval,
                            float *__restrict__ phi, float *__restrict__ plo) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t sg = seg_of(flag, segid, e);
        const int q = qpos[e], lane = seg_lane[sg], t = q & 3;
        const int64_t step = sstep[seg_slice[sg]] + q / 4;
        lrow[slot_pos(step, lane, t)] = (uint16_t)(code_row((uint32_t)(keys[e] & 0xFFFFull)) * 8u);
        const int64_t src = (int64_t)perm[e];
        if (pval) pval[val_pos(step, lane, t)] = val[src];
        if (phi) {
            phi[slot_pos(step, lane, t)] = hi[src];
            plo[slot_pos(step, lane, t)] = lo[src];
        }
    }
}

// The pad entries (values stay 0): rows of the scheduled classes.
__global__ void k_sell_pads(int64_t nslice, const int64_t *__restrict__ sstep, const int64_t *__restrict__ padoff,
                            const uint8_t *__restrict__ padcls, uint16_t *__restrict__ lrow) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nslice * SW;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t sl = g / SW;
        const int lane = (int)(g % SW);
        const int cap = (int)((padoff[sl + 1] - padoff[sl]) / SW);
        const uint8_t *pads = padcls + padoff[sl] + (int64_t)lane * cap;
        const int npos = (int)(4 * (sstep[sl + 1] - sstep[sl]));
        for (int p = 0; p < npos; ++p)
            if (pads[p] != 0xFF) lrow[slot_pos(sstep[sl] + p / 4, lane, p & 3)] = (uint16_t)(pads[p] * 8u);
    }
}

// ------------------------------------------ full layouts: row deltas

// Sort keys of a full layout: every lower entry e = (i, j) as row i of
// column j and as row j of column i (key nnz + e), ordered by (panel,
// column, row).
__global__ void k_keys_full(int64_t n, int64_t nnz, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                            int R, uint64_t *__restrict__ keys, uint32_t *__restrict__ idx) {
    const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const uint64_t panel = (uint64_t)(row / R), lr = (uint64_t)(row % R);
    for (int64_t e = rp[row] + lane; e < rp[row + 1]; e += 64) {
        const uint64_t c = (uint64_t)col[e];
        keys[e] = (panel << 40) | (c << 16) | lr;
        idx[e] = (uint32_t)e;
        keys[nnz + e] = ((c / R) << 40) | ((uint64_t)row << 16) | (c % R);
        idx[nnz + e] = (uint32_t)e;
    }
}

constexpr int MAXDELTA = 3;  // largest row step a two-bit code carries

// Escape entries before sorted entry e: a gap of more than 3 rows from the
// previous entry of its segment takes one escape entry (code 0), whose value
// slot holds the entry's row - 1 as an integer.
__host__ __device__ inline int32_t escapes(const uint64_t *keys, const int32_t *flag, int64_t e) {
    if (flag[e]) return 0;
    const int32_t gap = (int32_t)(keys[e] & 0xFFFFull) - (int32_t)(keys[e - 1] & 0xFFFFull);
    return gap > MAXDELTA ? 1 : 0;
}

struct EscapesOf {
    const uint64_t *keys;
    const int32_t *flag;
    __host__ __device__ int32_t operator()(int64_t e) const { return escapes(keys, flag, e); }
};

// per segment: its length in stream positions (entries plus escape steps)
// and the key sorting a panel's segments longest first
__global__ void k_seg_order_full(int64_t nseg, int64_t nnz, const int64_t *__restrict__ ustart,
                                 const int32_t *__restrict__ seg_panel, const int32_t *__restrict__ escx,
                                 int64_t esc_total, int32_t *__restrict__ ulen, uint64_t *__restrict__ okey,
                                 int64_t *__restrict__ oval) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nseg; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t end = s + 1 < nseg ? ustart[s + 1] : nnz;
        const int64_t pads = (end < nnz ? (int64_t)escx[end] : esc_total) - (int64_t)escx[ustart[s]];
        const int64_t len = end - ustart[s] + pads;
        ulen[s] = (int32_t)len;
        okey[s] = ((uint64_t)seg_panel[s] << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)len);
        oval[s] = s;
    }
}

// per slice: its steps (the longest segment, in lane 0, in quads) and each
// lane's first row
__global__ void k_slice_full(int64_t nseg, const int64_t *__restrict__ ustart, const int32_t *__restrict__ ulen,
                             const int64_t *__restrict__ seg_slice, const int32_t *__restrict__ seg_lane,
                             const uint64_t *__restrict__ keys, int64_t *__restrict__ slice_steps,
                             uint16_t *__restrict__ sfirst) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nseg; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t sl = seg_slice[s];
        const int lane = seg_lane[s];
        sfirst[sl * SW + lane] = (uint16_t)(keys[ustart[s]] & 0xFFFFull);
        if (lane == 0) slice_steps[sl] = (ulen[s] + 3) / 4;
    }
}

// Row code of position t of a lane's quad in step `step`: two bits of the
// lane's code byte (step * 64 + lane), packed four bytes to a word.
__device__ inline void put_code(uint32_t *codes, int64_t step, int lane, int t, uint32_t c) {
    const int64_t b = step * SW + lane;
    atomicOr(codes + (b >> 2), c << (8 * (int)(b & 3) + 2 * t));
}

// The stream of a full layout: each entry at its position in its lane (the
// entries before it in its segment plus their escape entries), with its
// two-bit row code and value.  Code c = 1..3: the row is the lane's previous
// row + c (the segment's first entry: its first row + 1 - 1); code 0: an
// escape -- the lane's row becomes the integer in the value slot, and the
// entry contributes nothing.  A gap above 3 rows puts an escape holding
// row - 1 before the entry, which then has code 1.  Tail pads are escapes to
// row 0 (all-zero code bytes and values).
__global__ void k_fill_full(int64_t nnz, const uint64_t *__restrict__ keys, const int32_t *__restrict__ flag,
                            const int64_t *__restrict__ segid, const int64_t *__restrict__ ustart,
                            const int32_t *__restrict__ escx,
                            const int64_t *__restrict__ seg_slice, const int32_t *__restrict__ seg_lane,
                            const int64_t *__restrict__ sstep, const uint32_t *__restrict__ perm,
                            const double *__restrict__ val, const float *__restrict__ hi, const float *__restrict__ lo,
                            uint32_t *__restrict__ codes, double *__restrict__ pval, float *__restrict__ phi,
                            float *__restrict__ plo) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t sg = seg_of(flag, segid, e);
        const int64_t us = ustart[sg];
        const int64_t p = escapes(keys, flag, e);
        const int64_t pos = (e - us) + ((int64_t)escx[e] + p - (int64_t)escx[us]);
        const int lane = seg_lane[sg];
        const int64_t s0 = sstep[seg_slice[sg]];
        const uint32_t row = (uint32_t)(keys[e] & 0xFFFFull);
        const uint32_t code =
            (flag[e] || p) ? 1u : (uint32_t)(row - (uint32_t)(keys[e - 1] & 0xFFFFull));
        if (p) {  // the escape entry: code 0 (nothing to set), value slot = row - 1
            const int64_t qs = s0 + (pos - 1) / 4;
            const int qt = (int)((pos - 1) & 3);
            const uint32_t r1 = row - 1;
            // stored as integers: the row must survive as bits (a float or
            // double view of a small integer is a denormal)
            if (pval) reinterpret_cast<uint64_t *>(pval)[val_pos(qs, lane, qt)] = (uint64_t)r1;
            if (phi) reinterpret_cast<uint32_t *>(phi)[slot_pos(qs, lane, qt)] = r1;
        }
        const int64_t step = s0 + pos / 4;
        const int t = (int)(pos & 3);
        put_code(codes, step, lane, t, code);
        const int64_t src = (int64_t)perm[e];
        if (pval) pval[val_pos(step, lane, t)] = val[src];
        if (phi) {
            phi[slot_pos(step, lane, t)] = hi[src];
            plo[slot_pos(step, lane, t)] = lo[src];
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

constexpr int RSRC_WORD3 = 0x00020000;       // raw (untyped) buffer, range-checked

// A buffer resource (scalar registers) over `bytes` bytes at p.  Every lane
// access is then a 32-bit byte offset, and an offset past the end reads
// zeros without touching memory (the hardware range check).  The base and
// size are uniform over the wavefront; readfirstlane states it.
__device__ inline __amdgpu_buffer_rsrc_t rsrc(const void *p, uint32_t bytes) {
    const uint64_t a = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    void *u = (void *)(((uint64_t)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(u, (short)0, (int)__builtin_amdgcn_readfirstlane(bytes), RSRC_WORD3);
}

// a wavefront-uniform 64-bit value, stated as such (scalar registers)
__device__ inline int64_t uniform64(int64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

struct StreamRsrc {
    __amdgpu_buffer_rsrc_t rows, hi, lo, val;
};

// Cache policy of the stream loads (row codes and values): non-temporal --
// each byte is read once per launch, so the caches need not keep it (3%
// faster bulk launches than the default policy on the batched workload).
constexpr int NT = 2;

// One lane's quad of one step (entry offset `e` from the wavefront's first
// entry, a multiple of 4).  settle() marks the registers as read on every
// path (an empty asm statement with them as inputs), so the wait for their
// load is placed before any branch and is the same on all paths.
//
// Row part: DL = false (lower layouts), four 16-bit LDS byte offsets (one
// 8-byte load); DL = true (full layouts), four two-bit row codes (one
// byte load).
template <bool DL>
struct QRows;
template <>
struct QRows<false> {
    u2v r;
    __device__ inline void load(const StreamRsrc &p, uint32_t e) {
        r = __builtin_bit_cast(u2v, __builtin_amdgcn_raw_buffer_load_b64(p.rows, e * 2u, 0, NT));
    }
    __device__ inline void settle() const { asm volatile("" ::"v"(r)); }
};
template <>
struct QRows<true> {
    uint32_t r;
    __device__ inline void load(const StreamRsrc &p, uint32_t e) {
        r = __builtin_amdgcn_raw_buffer_load_b8(p.rows, e >> 2, 0, NT);  // the lane's code byte
    }
    __device__ inline void settle() const { asm volatile("" ::"v"(r)); }
};

template <int MODE>
struct QVals;

// The value slots are kept as raw bits: an escape entry's slot holds a row
// as an integer, which must not pass through a float register view.  An
// element is copied out of the vector before it is reinterpreted
// (__builtin_bit_cast applied to a vector subscript reads the vector's first
// element).
typedef unsigned u4v __attribute__((ext_vector_type(4)));
__device__ inline float as_float(uint32_t u) { return __uint_as_float(u); }

template <>
struct QVals<GCI_OP_SPLIT_BULK> {
    u4v h;
    __device__ inline void load(const StreamRsrc &p, uint32_t e) {
        h = __builtin_bit_cast(u4v, __builtin_amdgcn_raw_buffer_load_b128(p.hi, e * 4u, 0, NT));
    }
    __device__ inline double value(int j) const {
        const uint32_t u = h[j];
        return (double)as_float(u);
    }
    __device__ inline uint32_t bits(int j) const {
        const uint32_t u = h[j];
        return u;
    }
    __device__ inline void settle() const { asm volatile("" ::"v"(h)); }
};

template <>
struct QVals<GCI_OP_SPLIT_FULL> {
    u4v h;
    f4v l;
    __device__ inline uint32_t bits(int j) const {
        const uint32_t u = h[j];
        return u;
    }
    __device__ inline void load(const StreamRsrc &p, uint32_t e) {
        h = __builtin_bit_cast(u4v, __builtin_amdgcn_raw_buffer_load_b128(p.hi, e * 4u, 0, NT));
        l = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(p.lo, e * 4u, 0, NT));
    }
    __device__ inline double value(int j) const {
        const uint32_t u = h[j];
        const float lj = l[j];
        return (double)as_float(u) + (double)lj;
    }
    __device__ inline void settle() const { asm volatile("" ::"v"(h), "v"(l)); }
};

template <>
struct QVals<GCI_OP_FP64> {
    typedef unsigned long long q2v __attribute__((ext_vector_type(2)));
    q2v a, b;
    // e = step * 256 + lane * 4: the lane's halves sit at step * 256 + lane * 2
    // and step * 256 + 128 + lane * 2 (val_pos)
    __device__ inline void load(const StreamRsrc &p, uint32_t e) {
        const uint32_t h0 = (e & ~(uint32_t)(SE - 1)) + ((e & (SE - 1)) >> 1);
        a = __builtin_bit_cast(q2v, __builtin_amdgcn_raw_buffer_load_b128(p.val, h0 * 8u, 0, NT));
        b = __builtin_bit_cast(q2v, __builtin_amdgcn_raw_buffer_load_b128(p.val, (h0 + SE / 2) * 8u, 0, NT));
    }
    __device__ inline uint64_t raw(int j) const {
        const uint64_t u = j < 2 ? a[j] : b[j - 2];
        return u;
    }
    __device__ inline double value(int j) const { return __longlong_as_double((long long)raw(j)); }
    __device__ inline uint32_t bits(int j) const { return (uint32_t)raw(j); }
    __device__ inline void settle() const { asm volatile("" ::"v"(a), "v"(b)); }
};

template <int MODE, bool DL>
struct Quad {
    QRows<DL> rw;
    QVals<MODE> v;
    __device__ inline void load(const StreamRsrc &p, uint32_t e) {
        rw.load(p, e);
        v.load(p, e);
    }
    __device__ inline double value(int j) const { return v.value(j); }
    __device__ inline uint32_t bits(int j) const { return v.bits(j); }
    __device__ inline void settle() const {
        rw.settle();
        v.settle();
    }
};

// M consecutive values of type T from global memory, unconverted.
template <int M, class T>
__device__ inline void load_raw(const GAS T *p, T (&o)[M]) {
    if constexpr ((M & (M - 1)) != 0) {  // odd widths: rows are not 16-byte aligned
#pragma unroll
        for (int v = 0; v < M; ++v) o[v] = p[v];
    } else if constexpr (M * sizeof(T) >= 16) {
#pragma unroll
        for (int q = 0; q < (int)(M * sizeof(T) / 16); ++q) {
            const f4v w = reinterpret_cast<const GAS f4v *>(p)[q];
            __builtin_memcpy(reinterpret_cast<char *>(o) + 16 * q, &w, 16);
        }
    } else if constexpr (M * sizeof(T) == 8) {
        const f2v w = reinterpret_cast<const GAS f2v *>(p)[0];
        __builtin_memcpy(o, &w, 8);
    } else {
        o[0] = p[0];
    }
}

// LDS carve-ups of one workgroup (bytes).
//
// Lower-triangle walk (FP64): ys (M x R fp64, vector-major) at 0, then xs
// (M x R of T), and the tables of a chunk's (at most NS) slices: cs (column
// partial sums, NS x 64 rows of M fp64), xj (the x_j rows, NS x 64 rows of M
// T), cl (columns, NS x 64 ints) and ss (first steps, NS + 1 ints).
//
// Stream walk (full layouts): xs (R rows of M T, row-major) at 0, then the
// tables of the item's (at most NI) slices: cl (columns), fr (first rows -
// panel start, u16) and ss (first steps), and per wavefront a staging row of
// 64 x M fp64 (st) that turns a slice's column sums into atomics over
// consecutive addresses.
struct SellLds {
    size_t xs, cs, xj, cl, fr, ss, st, total;
};
constexpr size_t round16(size_t b) { return (b + 15) & ~(size_t)15; }
template <int M, class T>
__host__ __device__ constexpr SellLds sell_lds(int R, int NS) {
    SellLds m{};
    m.xs = (size_t)R * M * sizeof(double);
    m.cs = m.xs + round16((size_t)R * M * sizeof(T));
    m.xj = m.cs + (size_t)NS * SW * M * sizeof(double);
    m.cl = m.xj + round16((size_t)NS * SW * M * sizeof(T));
    m.ss = m.cl + (size_t)NS * SW * sizeof(int);
    m.total = m.ss + round16((size_t)(NS + 1) * sizeof(int));
    return m;
}
// Elements per x row of the stream walk's LDS panel: M, except five FP64
// values, which are padded to six so that a row is three aligned 16-byte
// reads instead of five 8-byte ones.
template <int M, class T>
__host__ __device__ constexpr int xrow() { return (M == 5 && sizeof(T) == 8) ? 6 : M; }

template <int M, class T>
__host__ __device__ constexpr SellLds stream_lds(int R, int NI, int W) {
    SellLds m{};
    m.xs = 0;
    m.cl = round16((size_t)R * xrow<M, T>() * sizeof(T));
    m.fr = m.cl + (size_t)NI * SW * sizeof(int);
    m.ss = m.fr + (size_t)NI * SW * sizeof(uint16_t);
    m.st = m.ss + round16((size_t)(NI + 1) * sizeof(int));
    m.total = m.st + (size_t)W * SW * M * sizeof(double);
    return m;
}

// M consecutive values of type T from LDS (a row of row-major x), in as
// few reads as the row's size and 16-byte alignment allow.
template <int M, class T>
__device__ inline void lds_row(const T *p, T (&o)[M]) {
    if constexpr (xrow<M, T>() != M) {  // padded row: whole 16-byte reads, the pad dropped
        constexpr int Q = xrow<M, T>() * (int)sizeof(T) / 16;
        T w[Q * 16 / sizeof(T)];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const f4v v = reinterpret_cast<const f4v *>(p)[q];
            __builtin_memcpy(reinterpret_cast<char *>(w) + 16 * q, &v, 16);
        }
#pragma unroll
        for (int v = 0; v < M; ++v) o[v] = w[v];
    } else if constexpr ((M & (M - 1)) == 0 && M * sizeof(T) >= 16) {
#pragma unroll
        for (int q = 0; q < (int)(M * sizeof(T) / 16); ++q) {
            const f4v w = reinterpret_cast<const f4v *>(p)[q];
            __builtin_memcpy(reinterpret_cast<char *>(o) + 16 * q, &w, 16);
        }
    } else if constexpr ((M & (M - 1)) == 0 && M * sizeof(T) == 8) {
        const f2v w = reinterpret_cast<const f2v *>(p)[0];
        __builtin_memcpy(o, &w, 8);
    } else {
#pragma unroll
        for (int v = 0; v < M; ++v) o[v] = p[v];
    }
}

// One lane's quad: its four entries (row codes off = lr * 8, values a) applied
// to the M vectors: the column partials acc[v] += a x_i[v] (x_i from LDS) and
// y_i[v] += a x_j[v] through LDS atomics.  All row reads are issued first.
// For a full layout (DL) the rows come from the lane's running row lr
// (row - panel start) plus the quad's four deltas.
template <int M, int MODE, int R, bool SCATTER, bool DL>
__device__ inline void consume(const Quad<MODE, DL> &q, const char *xs, char *ys, const double (&xj)[M],
                               double (&acc)[M], uint32_t &lr) {
    using T = typename XT<MODE>::T;
    uint32_t off[4];
    T xv[4][M];
    double av[4];
    if constexpr (DL) {
        // full layouts keep x row-major in LDS: a row's M values are one
        // wide read.  Code c = 1..3 steps the row by c; code 0 (an escape)
        // sets it to the value slot's integer and contributes nothing.
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t c = (q.rw.r >> (2 * j)) & 3u;
            lr = c ? lr + c : q.bits(j);
            av[j] = c ? q.value(j) : 0.0;
            lds_row<M, T>(reinterpret_cast<const T *>(xs) + (size_t)lr * xrow<M, T>(), xv[j]);
        }
    } else {
        off[0] = q.rw.r.x & 0xFFFFu;
        off[1] = q.rw.r.x >> 16;
        off[2] = q.rw.r.y & 0xFFFFu;
        off[3] = q.rw.r.y >> 16;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const char *xb = xs + (sizeof(T) == 4 ? off[j] >> 1 : off[j]);
#pragma unroll
            for (int v = 0; v < M; ++v) xv[j][v] = *reinterpret_cast<const T *>(xb + (size_t)v * R * sizeof(T));
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double a = DL ? av[j] : q.value(j);
#pragma unroll
        for (int v = 0; v < M; ++v) acc[v] = fma(a, (double)xv[j][v], acc[v]);
        if constexpr (SCATTER) {
#pragma unroll
            for (int v = 0; v < M; ++v)
                __hip_atomic_fetch_add(reinterpret_cast<double *>(ys + off[j] + (size_t)v * R * 8), a * xj[v],
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

// One turn of the prefetch ring: D + 1 steps, step J consuming register set
// J while the load D steps ahead goes into set (J + D) mod (D + 1).  The sets
// are indexed by constants only, so they stay in registers.
template <int J, int D, class Step, class Q>
__device__ inline void ring_steps(Step &step, Q (&q)[D + 1]) {
    step(q[J], q[(J + D) % (D + 1)]);
    if constexpr (J < D) ring_steps<J + 1, D>(step, q);
}

// The walk of a full layout (split-precision matrices, and FP64 ones laid
// out full): one workgroup per work item (whole slices of one panel).  The
// panel's x rows and the item's slice tables are staged in LDS once; then
// every wavefront walks its share of the item's slices (cut at slice starts,
// where each lane's running row is known) with one prefetch ring over the
// whole share, and at each slice end adds its column sums -- the
// matrix-vector product's rows, as the layout holds every entry in both
// orientations -- into y with global atomics.  No barrier after the
// staging, so a wavefront's ring never drains until its share ends.
template <int M, int MODE, int R, int BS, int D>
__device__ inline void sell_stream(const SpmmItem &it, const SpmmJob &jb, char *smem, int NI) {
    using T = typename XT<MODE>::T;
    constexpr int W = BS / 64;  // wavefronts per workgroup
    const SellLds lm = stream_lds<M, T>(R, NI, W);
    T *xs = reinterpret_cast<T *>(smem + lm.xs);
    int *cl = reinterpret_cast<int *>(smem + lm.cl);
    uint16_t *fr = reinterpret_cast<uint16_t *>(smem + lm.fr);
    int *ss = reinterpret_cast<int *>(smem + lm.ss);
    const GAS T *gx = (const GAS T *)jb.xs;
    GAS double *gy = (GAS double *)jb.ys;
    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
    const int rows = it.rows;
    const int nsl = (int)(it.slice1 - it.slice0);
    const int nst = (int)(it.step1 - it.step0);
    if (nsl > NI) return;  // never reached: items are cut for the shape
    // x rows, row-major: every load of a batch is issued before its stores
    // (indexes clamped to the panel, stores masked), so one wait covers it
    constexpr int XP = (R + BS - 1) / BS;
    constexpr int XB = XP < 4 ? XP : 4;
    for (int j0 = 0; j0 < XP; j0 += XB) {
        T xr[XB][M];
#pragma unroll
        for (int j = 0; j < XB; ++j) load_raw<M, T>(gx + (it.row0 + min(t + (j0 + j) * BS, rows - 1)) * M, xr[j]);
#pragma unroll
        for (int j = 0; j < XB; ++j) {
            const int rr = t + (j0 + j) * BS;
            if (rr < rows) {
#pragma unroll
                for (int v = 0; v < M; ++v) xs[rr * xrow<M, T>() + v] = xr[j][v];
            }
        }
    }
    // slice tables: columns, first rows, first steps (relative to the item),
    // again in batches whose loads are all in flight together
    {
        const int nt = nsl * SW;
        const GAS int *gc = (const GAS int *)jb.scol + it.slice0 * SW;
        const GAS uint16_t *gf = (const GAS uint16_t *)jb.sfirst + it.slice0 * SW;
        constexpr int TB = 8;
        for (int i0 = 0; i0 < nt; i0 += TB * BS) {
            int c[TB];
            uint16_t f[TB];
#pragma unroll
            for (int j = 0; j < TB; ++j) {
                const int src = min(i0 + t + j * BS, nt - 1);
                c[j] = gc[src];
                f[j] = gf[src];
            }
#pragma unroll
            for (int j = 0; j < TB; ++j) {
                const int i = i0 + t + j * BS;
                if (i < nt) {
                    cl[i] = c[j];
                    fr[i] = f[j];
                }
            }
        }
        static_assert(ITEM_SLICES_MAX < BS, "one first step per thread");
        if (t <= nsl) ss[t] = (int)(((const GAS int64_t *)jb.sstep)[it.slice0 + t] - it.step0);
    }
    __syncthreads();

    // the first slice start at or after step x (ss ascends strictly: every
    // slice has a step; ss[nsl] = nst)
    auto slice_start_from = [&](int x) {
        int cnt = 0;
        for (int base = 0;; base += 64) {
            const int i = base + lane;
            const uint64_t mk = __ballot(i <= nsl && ss[min(i, nsl)] < x);
            cnt += __popcll(mk);
            if (mk != ~0ull) break;
        }
        return __builtin_amdgcn_readfirstlane(ss[min(cnt, nsl)]);
    };
    // an even split of the steps, moved to slice starts
    const int a = slice_start_from((int)((int64_t)nst * w / W));
    const int b = (w == W - 1) ? nst : slice_start_from((int)((int64_t)nst * (w + 1) / W));
    if (a >= b) return;
    int si = 0;  // the slice starting at step a
    for (int base = 1;; base += 64) {
        const int i = base + lane;
        const uint64_t mk = __ballot(i <= nsl && ss[min(i, nsl)] <= a);
        si += __popcll(mk);
        if (mk != ~0ull) break;
    }
    const int64_t e0 = (it.step0 + a) * (int64_t)SE;
    const uint32_t ne = (uint32_t)(b - a) * SE;
    StreamRsrc P;
    P.rows = rsrc(jb.drow + e0 / 4, ne / 4);  // one code byte per lane and step
    P.hi = rsrc(jb.hi + e0, ne * 4u);
    P.lo = rsrc(jb.lo + e0, ne * 4u);
    P.val = rsrc(jb.val + e0, ne * 8u);
    const double xj[M] = {};
    double acc[M];
    uint32_t lr = 0;  // the lane's running row - panel start
    auto enter = [&]() {
#pragma unroll
        for (int v = 0; v < M; ++v) acc[v] = 0.0;
        lr = (uint32_t)fr[si * SW + lane] - 1u;  // the first entry's code is 1
    };
    // a slice's column sums leave through the wavefront's staging row: M
    // atomics, the j-th covering (lane, vector) pairs 64j .. 64j + 63, which
    // are consecutive doubles of y where the slice's columns are
    double *st = reinterpret_cast<double *>(smem + lm.st) + w * (SW * M);
    auto leave = [&]() {
#pragma unroll
        for (int v = 0; v < M; ++v) st[lane * M + v] = acc[v];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const int e = j * SW + lane, v = e % M;
            const double y = st[e];
            const int c = cl[si * SW + e / M];
            if (c >= 0 && v < jb.m && y != 0.0) __builtin_amdgcn_global_atomic_fadd_f64(gy + (int64_t)c * M + v, y);
        }
    };
    enter();
    int snext = __builtin_amdgcn_readfirstlane(ss[si + 1]);
    // The loads run D steps ahead of the consumer through a ring of D + 1
    // register sets; a load past the wavefront's steps reads zeros, so every
    // step issues the same loads.
    Quad<MODE, true> q[D + 1];
    int k = a;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        q[d].load(P, (uint32_t)(d * SW + lane) * 4u);
        __builtin_amdgcn_sched_barrier(0);  // issued in ring order, as the loop issues them
    }
    auto step = [&](Quad<MODE, true> &cur, Quad<MODE, true> &nxt) {
        nxt.load(P, (uint32_t)((k - a + D) * SW + lane) * 4u);
        cur.settle();
        if (k < b) {
            consume<M, MODE, R, false, true>(cur, reinterpret_cast<const char *>(xs), nullptr, xj, acc, lr);
            if (k + 1 == snext && k + 1 < b) {
                leave();
                ++si;
                enter();
                snext = __builtin_amdgcn_readfirstlane(ss[si + 1]);
            }
        }
        ++k;
    };
    while (k < b) ring_steps<0, D>(step, q);
#pragma unroll
    for (int d = 0; d <= D; ++d) q[d].settle();
    leave();
}

// The walk of a lower-triangle layout (FP64 values): the item in chunks of
// at most NS slices, whose tables (columns, x_j rows, column sums) fill the
// LDS beside the panel's x rows and row accumulators.  Every entry acts
// twice: its column partial in registers, and y_i += H_ij x_j through an LDS
// atomic; chunk ends flush the column sums, the item's end the rows.
template <int M, int R, int BS, int D>
__device__ inline void sell_lower(const SpmmItem &it, const SpmmJob &jb, char *smem, int NS) {
    constexpr int MODE = GCI_OP_FP64;
    using T = double;
    constexpr int W = BS / 64;  // wavefronts per workgroup
    const int rows = it.rows;
    const SellLds lm = sell_lds<M, T>(R, NS);
    double *ys = reinterpret_cast<double *>(smem);
    T *xs = reinterpret_cast<T *>(smem + lm.xs);
    double *cs = reinterpret_cast<double *>(smem + lm.cs);
    T *xjt = reinterpret_cast<T *>(smem + lm.xj);
    int *cl = reinterpret_cast<int *>(smem + lm.cl);
    int *ss = reinterpret_cast<int *>(smem + lm.ss);
    const GAS T *gx = (const GAS T *)jb.xs;
    GAS double *gy = (GAS double *)jb.ys;
    const int t = threadIdx.x;
    // wavefront-uniform values are read through readfirstlane, so the walk's
    // control flow is scalar
    const int w = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
    // the panel's x rows and zeroed y rows stay for the whole item
    for (int rr = t; rr < rows; rr += BS) {
        T xr[M];
        load_raw<M, T>(gx + (it.row0 + rr) * M, xr);
#pragma unroll
        for (int v = 0; v < M; ++v) {
            xs[v * R + rr] = xr[v];
            ys[v * R + rr] = 0.0;
        }
    }
    // the item's slices in chunks of at most NS, whose tables fill the LDS
    for (int64_t c0 = it.slice0; c0 < it.slice1; c0 += NS) {
        const int nsl = (int)min((int64_t)NS, it.slice1 - c0);
        const int64_t st0 = uniform64(max(it.step0, jb.sstep[c0]));
        const int nst = __builtin_amdgcn_readfirstlane((int)(min(it.step1, jb.sstep[c0 + nsl]) - st0));
        // tables: first steps (relative to st0), columns, x_j rows, zeroed sums
        for (int i = t; i <= nsl; i += BS) ss[i] = (int)(jb.sstep[c0 + i] - st0);
        for (int i = t; i < nsl * SW; i += BS) {
            const int c = jb.scol[c0 * SW + i];
            cl[i] = c;
            T xr[M];
            if (c >= 0) {
                load_raw<M, T>(gx + (int64_t)c * M, xr);
            } else {
#pragma unroll
                for (int v = 0; v < M; ++v) xr[v] = (T)0;
            }
#pragma unroll
            for (int v = 0; v < M; ++v) {
                xjt[i * M + v] = xr[v];
                cs[i * M + v] = 0.0;
            }
        }
        __syncthreads();

        // One wavefront's walk over an even share [a, b) of the chunk's
        // steps.  The only global accesses inside a walk are the stream
        // loads, so the waits for them are exact.
        [&]() {
            const int a = (int)((int64_t)nst * w / W), b = (int)((int64_t)nst * (w + 1) / W);
            if (a >= b) return;
            // the slice holding step a: the last i with ss[i] <= a (ss[nsl] >= nst)
            int si = 0;
            for (int base = 1;; base += 64) {
                const int i = base + lane;
                const uint64_t mk = __ballot(i <= nsl && ss[min(i, nsl)] <= a);
                si += __popcll(mk);
                if (mk != ~0ull) break;
            }
            const int64_t e0 = (st0 + a) * (int64_t)SE;
            const uint32_t ne = (uint32_t)(b - a) * SE;
            StreamRsrc P;
            P.rows = rsrc(jb.lrow + e0, ne * 2u);
            P.val = rsrc(jb.val + e0, ne * 8u);
            double xj[M], acc[M];
            uint32_t lr = 0;  // unused by lower layouts
            auto enter = [&]() {  // slice si: its lane's x_j
#pragma unroll
                for (int v = 0; v < M; ++v) {
                    xj[v] = (double)xjt[(si * SW + lane) * M + v];
                    acc[v] = 0.0;
                }
            };
            auto leave = [&]() {  // the column partials of slice si into its table
                double *c = cs + (si * SW + lane) * M;
#pragma unroll
                for (int v = 0; v < M; ++v)
                    __hip_atomic_fetch_add(c + v, acc[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            };
            enter();
            int snext = __builtin_amdgcn_readfirstlane(ss[si + 1]);  // first step of the next slice
            Quad<MODE, false> q[D + 1];
            int k = a;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                q[d].load(P, (uint32_t)(d * SW + lane) * 4u);
                __builtin_amdgcn_sched_barrier(0);  // issued in ring order, as the loop issues them
            }
            auto step = [&](Quad<MODE, false> &cur, Quad<MODE, false> &nxt) {
                nxt.load(P, (uint32_t)((k - a + D) * SW + lane) * 4u);
                cur.settle();
                if (k < b) {
                    // a lower layout's entries update y_i too
                    consume<M, MODE, R, true, false>(cur, reinterpret_cast<const char *>(xs),
                                                     reinterpret_cast<char *>(ys), xj, acc, lr);
                    if (k + 1 == snext && k + 1 < b) {
                        leave();
                        ++si;
                        enter();
                        snext = __builtin_amdgcn_readfirstlane(ss[si + 1]);
                    }
                }
                ++k;
            };
            while (k < b) ring_steps<0, D>(step, q);
#pragma unroll
            for (int d = 0; d <= D; ++d) q[d].settle();
            leave();
        }();
        __syncthreads();
        // the chunk's column sums: M consecutive lanes per column row
        for (int idx = t; idx < nsl * SW * M; idx += BS) {
            const int c = cl[idx / M];
            const int v = idx % M;
            const double y = cs[idx];
            if (c >= 0 && v < jb.m && y != 0.0) __builtin_amdgcn_global_atomic_fadd_f64(gy + (int64_t)c * M + v, y);
        }
        __syncthreads();  // the next chunk overwrites the tables
    }
    // the panel's rows, as contiguous runs of the row-major staging buffer
    for (int idx = t; idx < rows * M; idx += BS) {
        const int rr = idx / M, v = idx % M;
        const double y = ys[v * R + rr];
        if (v < jb.m && y != 0.0) __builtin_amdgcn_global_atomic_fadd_f64(gy + (it.row0 + rr) * M + v, y);
    }
}

// The SpMM work-item kernel (see the file header): the stream walk for full
// layouts, the chunked walk for lower-triangle ones.
template <int M, int MODE, int R, int BS, int D>
__global__ __launch_bounds__(BS) void k_sell(const SpmmItem *__restrict__ items, const SpmmJob *__restrict__ jobs,
                                             int NS, int NI) {
    extern __shared__ __align__(64) char smem[];
    const SpmmItem it = items[blockIdx.x];
    const SpmmJob jb = jobs[it.job];
    if (!jb.active || jb.mode != MODE) return;
    if (it.rows > R) return;  // never reached: items are cut for the shape
    if constexpr (MODE != GCI_OP_FP64) {  // split-precision matrices are always laid out full
        sell_stream<M, MODE, R, BS, D>(it, jb, smem, NI);
    } else {
        if (jb.full) sell_stream<M, MODE, R, BS, D>(it, jb, smem, NI);
        else sell_lower<M, R, BS, D>(it, jb, smem, NS);
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

int chunk_slices(const SpmmShape &sh, int M);
int item_slices(const SpmmShape &sh) { return std::min(sh.nslice, ITEM_SLICES_MAX); }

template <int M, int MODE, int R, int BS, int D>
void launch_depth(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, const SpmmShape &sh,
                  hipStream_t s) {
    using T = typename XT<MODE>::T;
    // FP64 launches may meet lower-triangle layouts too: room for both walks
    const int ns = MODE == GCI_OP_FP64 ? chunk_slices(sh, M) : 0;
    GCI_REQUIRE(MODE != GCI_OP_FP64 || ns >= 1, "spmm: panel height and launch width exceed the LDS");
    const int ni = item_slices(sh);
    size_t shm = stream_lds<M, T>(R, ni, BS / 64).total;
    if (MODE == GCI_OP_FP64) shm = std::max(shm, sell_lds<M, double>(R, ns).total);
    GCI_REQUIRE(shm <= LDS_MAX, "spmm: panel height and launch width exceed the LDS");
    static size_t attr = 0;  // one instantiation per template tuple
    if (shm > attr) {
        GCI_HIP(hipFuncSetAttribute((const void *)k_sell<M, MODE, R, BS, D>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        attr = shm;
    }
    hipLaunchKernelGGL((k_sell<M, MODE, R, BS, D>), dim3((unsigned)nitems), dim3(BS), shm, s, d_items, d_jobs, ns, ni);
}

template <int M, int MODE, int R, int BS>
void launch_shape(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, const SpmmShape &sh,
                  hipStream_t s) {
    if (sh.depth == 2) launch_depth<M, MODE, R, BS, 2>(d_items, nitems, d_jobs, sh, s);
    else if (sh.depth == 8) launch_depth<M, MODE, R, BS, 8>(d_items, nitems, d_jobs, sh, s);
    else launch_depth<M, MODE, R, BS, 4>(d_items, nitems, d_jobs, sh, s);
}

// 1024-thread workgroups only up to 4 columns (8 would not fit the 128
// registers per lane that 16 wavefronts leave); wider launches use 512
// More than four columns of a full layout can fill so much LDS that a
// 512-thread workgroup is alone on its CU (FP64 x rows of the split operator's
// refinement): such launches take 1024-thread workgroups at depth 2 (within
// the 128 registers per lane that 16 wavefronts leave), twice the wavefronts
// in flight.
template <int M, int MODE, int R>
bool wide_pays(const SpmmShape &sh) {
    using T = typename XT<MODE>::T;
    const int ni = item_slices(sh);
    return MODE != GCI_OP_FP64 && M > 4 && sh.block == 512 && 2 * stream_lds<M, T>(R, ni, 8).total > LDS_MAX &&
           stream_lds<M, T>(R, ni, 16).total <= LDS_MAX;
}

template <int M, int MODE, int R>
void launch_bs(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, const SpmmShape &sh, hipStream_t s) {
    if constexpr (MODE != GCI_OP_FP64 && M > 4) {
        if (wide_pays<M, MODE, R>(sh)) {
            launch_depth<M, MODE, R, 1024, 2>(d_items, nitems, d_jobs, sh, s);
            return;
        }
    }
    if (sh.block == 256) {
        launch_shape<M, MODE, R, 256>(d_items, nitems, d_jobs, sh, s);
    } else if constexpr (M <= 4) {
        if (sh.block == 1024) launch_shape<M, MODE, R, 1024>(d_items, nitems, d_jobs, sh, s);
        else launch_shape<M, MODE, R, 512>(d_items, nitems, d_jobs, sh, s);
    } else {
        launch_shape<M, MODE, R, 512>(d_items, nitems, d_jobs, sh, s);
    }
}

template <int M, int MODE>
void launch_r(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, const SpmmShape &sh, hipStream_t s) {
    if (sh.R == 2048) launch_bs<M, MODE, 2048>(d_items, nitems, d_jobs, sh, s);
    else if (sh.R == 4096) launch_bs<M, MODE, 4096>(d_items, nitems, d_jobs, sh, s);
    else launch_bs<M, MODE, 1024>(d_items, nitems, d_jobs, sh, s);
}

template <int MODE>
void launch_mode(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, int M, const SpmmShape &sh,
                 hipStream_t s) {
    switch (M) {
        case 1: launch_r<1, MODE>(d_items, nitems, d_jobs, sh, s); break;
        case 2: launch_r<2, MODE>(d_items, nitems, d_jobs, sh, s); break;
        case 4: launch_r<4, MODE>(d_items, nitems, d_jobs, sh, s); break;
        case 5: launch_r<5, MODE>(d_items, nitems, d_jobs, sh, s); break;
        default: launch_r<8, MODE>(d_items, nitems, d_jobs, sh, s); break;
    }
}

int env_int(const char *name, int dflt) {
    const char *v = std::getenv(name);
    return v && *v ? std::atoi(v) : dflt;
}

// LDS of the lower-triangle walk (FP64) with NS slices per chunk
size_t lower_bytes(int R, int NS, int M) {
    switch (M) {
        case 1: return sell_lds<1, double>(R, NS).total;
        case 2: return sell_lds<2, double>(R, NS).total;
        case 4: return sell_lds<4, double>(R, NS).total;
        case 5: return sell_lds<5, double>(R, NS).total;
        default: return sell_lds<8, double>(R, NS).total;
    }
}

// LDS of the stream walk (full layouts)
size_t stream_bytes(int R, int NI, int W, int M, int mode) {
    if (mode == GCI_OP_SPLIT_BULK) {
        switch (M) {
            case 1: return stream_lds<1, float>(R, NI, W).total;
            case 2: return stream_lds<2, float>(R, NI, W).total;
            case 4: return stream_lds<4, float>(R, NI, W).total;
            case 5: return stream_lds<5, float>(R, NI, W).total;
            default: return stream_lds<8, float>(R, NI, W).total;
        }
    }
    switch (M) {
        case 1: return stream_lds<1, double>(R, NI, W).total;
        case 2: return stream_lds<2, double>(R, NI, W).total;
        case 4: return stream_lds<4, double>(R, NI, W).total;
        case 5: return stream_lds<5, double>(R, NI, W).total;
        default: return stream_lds<8, double>(R, NI, W).total;
    }
}

template <class It, class O>
void scan_exclusive_it(It in, O *out, int64_t n, hipStream_t s) {
    size_t tb = 0;
    GCI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, s));
    SBuf<char> tmp(tb, s);
    GCI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, in, out, n, s));
}

template <class I, class O>
void scan_exclusive(const I *in, O *out, int64_t n, hipStream_t s) {
    size_t tb = 0;
    GCI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, s));
    SBuf<char> tmp(tb, s);
    GCI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, in, out, n, s));
}

// Sorts (k, v) in place between two buffer pairs (no third copy); returns
// true when the result is in the second pair.
template <class K, class V>
bool sort_pairs_db(K *k0, K *k1, V *v0, V *v1, int64_t n, hipStream_t s) {
    hipcub::DoubleBuffer<K> kb(k0, k1);
    hipcub::DoubleBuffer<V> vb(v0, v1);
    size_t tb = 0;
    GCI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kb, vb, n, 0, 8 * (int)sizeof(K), s));
    SBuf<char> tmp(tb, s);
    GCI_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kb, vb, n, 0, 8 * (int)sizeof(K), s));
    return kb.selector == 1;
}

template <class K, class V>
void sort_pairs(const K *k0, K *k1, const V *v0, V *v1, int64_t n, hipStream_t s) {
    size_t tb = 0;
    GCI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, n, 0, 8 * (int)sizeof(K), s));
    SBuf<char> tmp(tb, s);
    GCI_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k0, k1, v0, v1, n, 0, 8 * (int)sizeof(K), s));
}

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)); }

// Slices per chunk of an FP64 launch over a lower-triangle layout: the
// shape's count, or fewer if the LDS of this width holds fewer (0: not even
// one).
int chunk_slices(const SpmmShape &sh, int M) {
    int ns = sh.nslice;
    while (ns > 0 && lower_bytes(sh.R, ns, M) > LDS_MAX) ns = ns > 64 ? ns - ns / 8 : ns - 1;
    return ns;
}

// A launch width runs only if its chunks hold at least this many slices
// (narrower column blocks are cheaper than tiny chunks).
constexpr int NS_MIN = 8;

// whether a launch of M columns in this mode fits the LDS
bool width_fits(const SpmmShape &sh, int M, int mode) {
    return stream_bytes(sh.R, item_slices(sh), sh.block / 64, M, mode) <= LDS_MAX &&
           (mode != GCI_OP_FP64 || chunk_slices(sh, M) >= std::min(NS_MIN, sh.nslice));
}

}  // namespace

bool valid_shape(const SpmmShape &sh) {
    // the instantiated launch shapes: R 1024, 2048 or 4096, 256/512/1024 threads, depth 2, 4 or 8
    return (sh.R == 1024 || sh.R == 2048 || sh.R == 4096) && (sh.block == 256 || sh.block == 512 || sh.block == 1024) &&
           sh.target >= 4096 && sh.nslice >= 1 && sh.nslice <= 1024 &&
           (sh.depth == 2 || sh.depth == 4 || sh.depth == 8) &&
           width_fits(sh, 1, GCI_OP_FP64);
}

SpmmShape spmm_shape() {
    if (!g_shape_set) {
        SpmmShape sh;
        sh.R = env_int("GCI_SPMM_R", sh.R);
        sh.block = env_int("GCI_SPMM_BLOCK", sh.block);
        sh.target = env_int("GCI_SPMM_TARGET", (int)sh.target);
        sh.nslice = env_int("GCI_SPMM_NSLICE", sh.nslice);
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
    for (int w : {8, 5, 4, 2})
        if (width_fits(sh, w, mode)) return w;
    return 1;
}

// The instantiated launch widths: 1, 2, 4, 5 and 8 columns.  Five serves the
// Davidson start and restart blocks of NUME + 1 = 5 vectors, and is the
// widest launch whose LDS holds FP64 x rows at 2048-row panels.
int spmm_width(int m) { return m <= 1 ? 1 : m <= 2 ? 2 : m <= 4 ? 4 : m == 5 ? 5 : 8; }

void spmm_items(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, int M, int mode, hipStream_t s) {
    if (nitems <= 0) return;
    const SpmmShape sh = spmm_shape();
    switch (mode) {
        case GCI_OP_FP64: launch_mode<GCI_OP_FP64>(d_items, nitems, d_jobs, M, sh, s); break;
        case GCI_OP_SPLIT_BULK: launch_mode<GCI_OP_SPLIT_BULK>(d_items, nitems, d_jobs, M, sh, s); break;
        default: launch_mode<GCI_OP_SPLIT_FULL>(d_items, nitems, d_jobs, M, sh, s); break;
    }
}

void build_items(HMat &h, const SpmmShape &sh) {
    Pcsc &pc = h.pc;
    pc.items.clear();
    const int R = pc.R;
    const int64_t target = sh.target;
    const int ni = item_slices(sh);
    const int64_t ts = std::max<int64_t>(1, target / SE);  // steps per item
    const std::vector<int64_t> &T = pc.sstep_h;
    for (int64_t sa = 0; sa < pc.nslice;) {
        const int32_t p = pc.spanel_h[sa];
        int64_t sb = sa + 1;
        while (sb < pc.nslice && pc.spanel_h[sb] == p) ++sb;
        const int64_t row0 = (int64_t)p * R;
        const int32_t rows = (int32_t)std::min<int64_t>(R, h.n - row0);
        if (pc.full) {
            // full layouts: items of whole slices (a walk starts on a slice
            // boundary, where every lane's running row is known)
            for (int64_t c = sa; c < sb;) {
                int64_t c1 = c + 1;
                while (c1 < sb && T[c1] - T[c] < ts && c1 - c < ni) ++c1;
                pc.items.push_back(SpmmItem{T[c], T[c1], c, c1, row0, rows, 0});
                c = c1;
            }
            sa = sb;
            continue;
        }
        int64_t cur = sa;  // the slice holding step st
        for (int64_t st = T[sa]; st < T[sb];) {
            const int64_t en = std::min(st + ts, T[sb]);
            int64_t s1 = cur;
            while (s1 < sb && T[s1] < en) ++s1;
            pc.items.push_back(SpmmItem{st, en, cur, s1, row0, rows, 0});
            st = en;
            while (cur + 1 < sb && T[cur + 1] <= st) ++cur;
        }
        sa = sb;
    }
    pc.target = target;
    pc.item_slices = ni;
}

void recut_items(HMat &h) {
    const SpmmShape sh = spmm_shape();
    GCI_REQUIRE(h.pc.R == sh.R, "spmm: panel height changed while the matrix is in use");
    if (h.pc.target != sh.target || h.pc.item_slices != item_slices(sh)) {
        build_items(h, sh);
        h.pc.d_items.release();
    }
}

void layout_fp64_values(HMat &h) {
    if (h.precision != GCI_SPLIT32 || h.layout_fp64) return;
    GCI_REQUIRE(h.val.p != nullptr, "spmm: FP64 operator on a matrix without FP64 values");
    h.layout_fp64 = true;
    if (h.pc.R) {
        GCI_HIP(hipDeviceSynchronize());  // queued launches may still read the old layout
        h.pc = Pcsc();
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
    const size_t b = stream_bytes(sh.R, item_slices(sh), sh.block / 64, M, mode);
    return mode == GCI_OP_FP64 ? std::max(b, lower_bytes(sh.R, std::max(1, chunk_slices(sh, M)), M)) : b;
}

size_t spmm_lds_max() { return LDS_MAX; }

void retain_pool() {
    static bool done[64] = {};
    int dev = 0;
    GCI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64 || done[dev]) return;
    hipMemPool_t pool;
    GCI_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
    uint64_t keep = 4ull << 30;
    GCI_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    done[dev] = true;
}

// The full layout of a split-precision matrix: both triangles, each
// segment's entries in row order with two-bit row codes, no bank schedule
// (a full layout's walk makes no LDS atomics).
void build_full(HMat &h, int R) {
    Pcsc &pc = h.pc;
    const int64_t nnz = 2 * h.nnz, n = h.n;
    hipStream_t s = h.stream;
    pc.nnz = nnz;
    GCI_REQUIRE(h.nnz < ((int64_t)1 << 32), "spmm: more than 2^32 - 1 off-diagonal entries");
    const unsigned eb = grid_for(nnz);
    SBuf<uint64_t> k1(nnz, s);
    SBuf<uint32_t> i1(nnz, s);
    {
        SBuf<uint64_t> k0(nnz, s);
        SBuf<uint32_t> i0(nnz, s);
        const int64_t blocks = (n * 64 + 255) / 256;
        hipLaunchKernelGGL(k_keys_full, dim3((unsigned)blocks), dim3(256), 0, s, n, h.nnz, h.rp.p, h.col.p, R, k0.p,
                           i0.p);
        GCI_HIP(hipGetLastError());
        if (!sort_pairs_db(k0.p, k1.p, i0.p, i1.p, nnz, s)) {
            std::swap(k0.p, k1.p);
            std::swap(i0.p, i1.p);
        }
    }
    // segments and the escape steps before each entry (an exclusive prefix
    // of escapes(), evaluated on the fly)
    SBuf<int32_t> flag(nnz, s), escx(nnz, s);
    SBuf<int64_t> segid(nnz, s);
    hipLaunchKernelGGL(k_seg_flags, dim3(eb), dim3(256), 0, s, nnz, k1.p, flag.p);
    GCI_HIP(hipGetLastError());
    scan_exclusive(flag.p, segid.p, nnz, s);
    {
        hipcub::CountingInputIterator<int64_t> ids(0);
        hipcub::TransformInputIterator<int32_t, EscapesOf, hipcub::CountingInputIterator<int64_t>> escs(
            ids, EscapesOf{k1.p, flag.p});
        scan_exclusive_it(escs, escx.p, nnz, s);
    }
    int64_t last = 0;
    int32_t lastf = 0, lastescx = 0;
    uint64_t lastkeys[2] = {0, 0};
    GCI_HIP(hipMemcpyAsync(&last, segid.p + nnz - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    GCI_HIP(hipMemcpyAsync(&lastf, flag.p + nnz - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GCI_HIP(hipMemcpyAsync(&lastescx, escx.p + nnz - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GCI_HIP(hipMemcpyAsync(lastkeys, k1.p + nnz - 2, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    GCI_HIP(hipStreamSynchronize(s));
    const int32_t lastflag[2] = {0, lastf};
    const int64_t nseg = last + lastf, esc_total = (int64_t)lastescx + escapes(lastkeys, lastflag, 1);
    GCI_REQUIRE(esc_total < ((int64_t)1 << 31), "spmm: too many row escapes in a full layout");
    pc.nseg = nseg;
    SBuf<int32_t> seg_col(nseg, s), spanel(nseg, s), ulen(nseg, s);
    SBuf<int64_t> ustart(nseg, s);
    hipLaunchKernelGGL(k_seg_meta, dim3(eb), dim3(256), 0, s, nnz, k1.p, flag.p, segid.p, seg_col.p, ustart.p,
                       spanel.p);
    GCI_HIP(hipGetLastError());
    // segments of a panel, longest first (in stream positions)
    const int64_t npanel = (n + R - 1) / R;
    const unsigned sg = grid_for(nseg);
    SBuf<int64_t> order(nseg, s);
    {
        SBuf<uint64_t> okey(nseg, s), okey1(nseg, s);
        SBuf<int64_t> oval(nseg, s);
        hipLaunchKernelGGL(k_seg_order_full, dim3(sg), dim3(256), 0, s, nseg, nnz, ustart.p, spanel.p, escx.p,
                           esc_total, ulen.p, okey.p, oval.p);
        GCI_HIP(hipGetLastError());
        sort_pairs(okey.p, okey1.p, oval.p, order.p, nseg, s);
    }
    // slices
    SBuf<int64_t> pslices(npanel + 1, s), pslice0(npanel + 1, s), pfirst(npanel + 1, s);
    hipLaunchKernelGGL(k_panel_slices, dim3(grid_for(npanel + 1)), dim3(256), 0, s, npanel, nseg, spanel.p, pfirst.p,
                       pslices.p);
    GCI_HIP(hipGetLastError());
    scan_exclusive(pslices.p, pslice0.p, npanel + 1, s);
    int64_t nslice = 0;
    GCI_HIP(hipMemcpyAsync(&nslice, pslice0.p + npanel, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    GCI_HIP(hipStreamSynchronize(s));
    GCI_REQUIRE(nslice < 0x7FFFFFFF, "spmm: too many slices");
    pc.nslice = nslice;
    pc.scol.alloc((size_t)nslice * SW);
    pc.sfirst.alloc((size_t)nslice * SW);
    SBuf<int32_t> slice_panel(nslice, s), seg_lane(nseg, s);
    SBuf<int64_t> slane_seg((size_t)nslice * SW, s), slice_steps(nslice + 1, s), seg_slice(nseg, s),
        pcap(nslice + 1, s);
    GCI_HIP(hipMemsetAsync(pc.scol.p, 0xFF, (size_t)nslice * SW * sizeof(int32_t), s));  // -1: empty lane
    GCI_HIP(hipMemsetAsync(pc.sfirst.p, 0, (size_t)nslice * SW * sizeof(uint16_t), s));
    GCI_HIP(hipMemsetAsync(slice_steps.p, 0, (nslice + 1) * sizeof(int64_t), s));
    hipLaunchKernelGGL(k_seg_place, dim3(sg), dim3(256), 0, s, nseg, order.p, spanel.p, seg_col.p, ulen.p, pfirst.p,
                       pslice0.p, pc.scol.p, slane_seg.p, slice_panel.p, seg_slice.p, seg_lane.p, pcap.p);
    GCI_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_slice_full, dim3(sg), dim3(256), 0, s, nseg, ustart.p, ulen.p, seg_slice.p, seg_lane.p,
                       k1.p, slice_steps.p, pc.sfirst.p);
    GCI_HIP(hipGetLastError());
    pc.sstep.alloc(nslice + 1);
    scan_exclusive(slice_steps.p, pc.sstep.p, nslice + 1, s);
    pc.sstep_h.resize(nslice + 1);
    pc.spanel_h.resize(nslice);
    GCI_HIP(hipMemcpyAsync(pc.sstep_h.data(), pc.sstep.p, (nslice + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    GCI_HIP(hipMemcpyAsync(pc.spanel_h.data(), slice_panel.p, nslice * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GCI_HIP(hipStreamSynchronize(s));
    const int64_t nstep = pc.sstep_h[nslice];
    pc.nstep = nstep;
    const size_t nst = (size_t)nstep * SE;
    pc.drow.alloc(nst / 4);  // one code byte per lane and step
    GCI_HIP(hipMemsetAsync(pc.drow.p, 0, nst / 4, s));  // tail pads: escapes to row 0, value 0
    if (h.val.p && h.layout_fp64) {
        pc.val.alloc(nst);
        GCI_HIP(hipMemsetAsync(pc.val.p, 0, nst * sizeof(double), s));
    }
    pc.hi.alloc(nst);
    pc.lo.alloc(nst);
    GCI_HIP(hipMemsetAsync(pc.hi.p, 0, nst * sizeof(float), s));
    GCI_HIP(hipMemsetAsync(pc.lo.p, 0, nst * sizeof(float), s));
    hipLaunchKernelGGL(k_fill_full, dim3(eb), dim3(256), 0, s, nnz, k1.p, flag.p, segid.p, ustart.p, escx.p,
                       seg_slice.p, seg_lane.p, pc.sstep.p, i1.p, h.val.p, h.hi.p, h.lo.p,
                       reinterpret_cast<uint32_t *>(pc.drow.p), pc.val.p, pc.hi.p, pc.lo.p);
    GCI_HIP(hipGetLastError());
    GCI_HIP(hipStreamSynchronize(s));
}

void ensure_pcsc(HMat &h) {
    if (h.pc.R) return;
    const SpmmShape sh = spmm_shape();
    const int R = sh.R;
    Pcsc &pc = h.pc;
    // Split-precision matrices (the batched operators, several vectors per
    // product) lay out both triangles: every output is then a column sum
    // of gathered x, with no LDS atomics, at twice the entries (two-bit
    // row codes keep a stored entry at 4.25 bytes for fp32 values).  FP64
    // matrices (the plain single-vector form) keep the lower triangle.
    pc.full = h.precision == GCI_SPLIT32;
    if (pc.full) {
        pc.nseg = pc.nslice = pc.nstep = 0;
        pc.sstep_h.assign(1, 0);
        pc.spanel_h.clear();
        if (h.nnz > 0) build_full(h, R);
        pc.R = R;
        build_items(h, sh);
        return;
    }
    const int64_t nnz = pc.full ? 2 * h.nnz : h.nnz, n = h.n;
    hipStream_t s = h.stream;
    pc.nnz = nnz;
    pc.nseg = pc.nslice = pc.nstep = 0;
    pc.sstep_h.assign(1, 0);
    pc.spanel_h.clear();
    if (nnz > 0) {
        const unsigned eb = grid_for(nnz);
        // entries sorted by (panel, column, row class, row); 12 bytes per
        // entry and its double buffer (the largest temporaries of the build)
        GCI_REQUIRE(h.nnz < ((int64_t)1 << 32), "spmm: more than 2^32 - 1 off-diagonal entries");
        SBuf<uint64_t> k1(nnz, s);
        SBuf<uint32_t> i1(nnz, s);
        {
            SBuf<uint64_t> k0(nnz, s);
            SBuf<uint32_t> i0(nnz, s);
            const int64_t blocks = (n * 64 + 255) / 256;
            hipLaunchKernelGGL(k_keys, dim3((unsigned)blocks), dim3(256), 0, s, n, h.nnz, h.rp.p, h.col.p, R,
                               pc.full ? 1 : 0, k0.p, i0.p);
            GCI_HIP(hipGetLastError());
            if (!sort_pairs_db(k0.p, k1.p, i0.p, i1.p, nnz, s)) {
                std::swap(k0.p, k1.p);  // keep the sorted pair, free the other
                std::swap(i0.p, i1.p);
            }
        }
        // segments
        SBuf<int32_t> flag(nnz, s);
        SBuf<int64_t> segid(nnz, s);
        hipLaunchKernelGGL(k_seg_flags, dim3(eb), dim3(256), 0, s, nnz, k1.p, flag.p);
        GCI_HIP(hipGetLastError());
        scan_exclusive(flag.p, segid.p, nnz, s);
        int64_t last = 0;
        int32_t lastf = 0;
        GCI_HIP(hipMemcpyAsync(&last, segid.p + nnz - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipMemcpyAsync(&lastf, flag.p + nnz - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipStreamSynchronize(s));
        const int64_t nseg = last + lastf;
        pc.nseg = nseg;
        SBuf<int32_t> seg_col(nseg, s), spanel(nseg, s), ulen(nseg, s);
        SBuf<int64_t> ustart(nseg, s);
        hipLaunchKernelGGL(k_seg_meta, dim3(eb), dim3(256), 0, s, nnz, k1.p, flag.p, segid.p, seg_col.p, ustart.p,
                           spanel.p);
        GCI_HIP(hipGetLastError());
        // segments of a panel, longest first
        const int64_t npanel = (n + R - 1) / R;
        const unsigned sg = grid_for(nseg);
        SBuf<int64_t> order(nseg, s);
        {
            SBuf<uint64_t> okey(nseg, s), okey1(nseg, s);
            SBuf<int64_t> oval(nseg, s);
            hipLaunchKernelGGL(k_seg_order, dim3(sg), dim3(256), 0, s, nseg, nnz, ustart.p, spanel.p, ulen.p, okey.p,
                               oval.p);
            GCI_HIP(hipGetLastError());
            sort_pairs(okey.p, okey1.p, oval.p, order.p, nseg, s);
        }
        // slices
        SBuf<int64_t> pslices(npanel + 1, s), pslice0(npanel + 1, s), pfirst(npanel + 1, s);
        hipLaunchKernelGGL(k_panel_slices, dim3(grid_for(npanel + 1)), dim3(256), 0, s, npanel, nseg, spanel.p,
                           pfirst.p, pslices.p);
        GCI_HIP(hipGetLastError());
        scan_exclusive(pslices.p, pslice0.p, npanel + 1, s);
        int64_t nslice = 0;
        GCI_HIP(hipMemcpyAsync(&nslice, pslice0.p + npanel, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipStreamSynchronize(s));
        pc.nslice = nslice;
        pc.scol.alloc((size_t)nslice * SW);
        SBuf<int32_t> slice_panel(nslice, s), seg_lane(nseg, s);
        SBuf<int64_t> slane_seg((size_t)nslice * SW, s), slice_steps(nslice + 1, s), seg_slice(nseg, s),
            pcap(nslice + 1, s), padoff(nslice + 1, s);
        GCI_HIP(hipMemsetAsync(pc.scol.p, 0xFF, (size_t)nslice * SW * sizeof(int32_t), s));    // -1: empty lane
        GCI_HIP(hipMemsetAsync(slane_seg.p, 0xFF, (size_t)nslice * SW * sizeof(int64_t), s));  // -1: empty lane
        GCI_HIP(hipMemsetAsync(pcap.p, 0, (nslice + 1) * sizeof(int64_t), s));
        hipLaunchKernelGGL(k_seg_place, dim3(sg), dim3(256), 0, s, nseg, order.p, spanel.p, seg_col.p, ulen.p,
                           pfirst.p, pslice0.p, pc.scol.p, slane_seg.p, slice_panel.p, seg_slice.p, seg_lane.p,
                           pcap.p);
        GCI_HIP(hipGetLastError());
        // the bank schedule: each entry's position in its slice, the pads'
        // classes and the slices' step counts
        GCI_REQUIRE(nslice < 0x7FFFFFFF, "spmm: too many slices");
        scan_exclusive(pcap.p, padoff.p, nslice + 1, s);
        int64_t npadbuf = 0;
        GCI_HIP(hipMemcpyAsync(&npadbuf, padoff.p + nslice, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipStreamSynchronize(s));
        SBuf<uint8_t> padcls(npadbuf, s);
        SBuf<int32_t> qpos(nnz, s);
        SBuf<int> overflow(1, s);
        GCI_HIP(hipMemsetAsync(padcls.p, 0xFF, npadbuf, s));  // 0xFF: an entry's position
        GCI_HIP(hipMemsetAsync(overflow.p, 0, sizeof(int), s));
        hipLaunchKernelGGL(k_sched, dim3((unsigned)nslice), dim3(SW), 0, s, slane_seg.p, ustart.p, ulen.p, k1.p,
                           slice_panel.p, R, n, padoff.p, slice_steps.p, qpos.p, padcls.p, overflow.p);
        GCI_HIP(hipGetLastError());
        pc.sstep.alloc(nslice + 1);
        scan_exclusive(slice_steps.p, pc.sstep.p, nslice + 1, s);
        pc.sstep_h.resize(nslice + 1);
        pc.spanel_h.resize(nslice);
        GCI_HIP(hipMemcpyAsync(pc.sstep_h.data(), pc.sstep.p, (nslice + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                               s));
        GCI_HIP(hipMemcpyAsync(pc.spanel_h.data(), slice_panel.p, nslice * sizeof(int32_t), hipMemcpyDeviceToHost,
                               s));
        int ovf = 0;
        GCI_HIP(hipMemcpyAsync(&ovf, overflow.p, sizeof(int), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipStreamSynchronize(s));
        GCI_REQUIRE(ovf == 0, "spmm: layout schedule overflow");
        const int64_t nstep = pc.sstep_h[nslice];
        pc.nstep = nstep;
        const size_t nst = (size_t)nstep * SE;
        pc.lrow.alloc(nst);
        if (h.val.p && (h.precision != GCI_SPLIT32 || h.layout_fp64)) pc.val.alloc(nst);
        if (h.precision == GCI_SPLIT32) {
            pc.hi.alloc(nst);
            pc.lo.alloc(nst);
        }
        // pad entries keep value 0
        if (pc.val.p) GCI_HIP(hipMemsetAsync(pc.val.p, 0, nst * sizeof(double), s));
        if (pc.hi.p) {
            GCI_HIP(hipMemsetAsync(pc.hi.p, 0, nst * sizeof(float), s));
            GCI_HIP(hipMemsetAsync(pc.lo.p, 0, nst * sizeof(float), s));
        }
        hipLaunchKernelGGL(k_sell_fill, dim3(eb), dim3(256), 0, s, nnz, k1.p, flag.p, segid.p, seg_slice.p,
                           seg_lane.p, qpos.p, pc.sstep.p, i1.p, h.val.p, h.hi.p, h.lo.p, pc.lrow.p, pc.val.p,
                           pc.hi.p, pc.lo.p);
        GCI_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sell_pads, dim3(grid_for(nslice * SW)), dim3(256), 0, s, nslice, pc.sstep.p, padoff.p,
                           padcls.p, pc.lrow.p);
        GCI_HIP(hipGetLastError());
        GCI_HIP(hipStreamSynchronize(s));
    }
    pc.R = R;
    build_items(h, sh);
}

// The matrix's work items in device memory (uploaded if missing or stale).
void ensure_device_items(HMat &h, hipStream_t s) {
    Pcsc &pc = h.pc;
    if (!pc.items.empty() && pc.d_items.n < pc.items.size()) {
        GCI_HIP(hipStreamSynchronize(s));
        pc.d_items.alloc(pc.items.size());
        GCI_HIP(hipMemcpy(pc.d_items.p, pc.items.data(), pc.items.size() * sizeof(SpmmItem),
                          hipMemcpyHostToDevice));
    }
}

namespace {
__global__ void k_put_job(SpmmJob *dst, SpmmJob j) { *dst = j; }

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

size_t spmm_scratch_need(const HMat &h, int m, int mode) {
    const int W = spmm_max_width(mode);
    const int Mmax = spmm_width(std::min(W, m));
    const size_t xt = mode == GCI_OP_SPLIT_BULK ? sizeof(float) : sizeof(double);
    return round256((size_t)h.n * Mmax * xt) + round256((size_t)h.n * Mmax * sizeof(double)) +
           round256(sizeof(SpmmJob));
}

void spmm_noalloc(HMat &h, int m, const double *x, int ldx, double *y, int ldy, int mode, hipStream_t s,
                  char *scratch, size_t cap) {
    GCI_REQUIRE(m >= 1, "spmm: m < 1");
    GCI_REQUIRE(ldx >= h.n && ldy >= h.n, "spmm: leading dimension < n");
    GCI_REQUIRE(mode == GCI_OP_FP64 || mode == GCI_OP_SPLIT_BULK || mode == GCI_OP_SPLIT_FULL, "spmm: bad mode");
    if (mode != GCI_OP_FP64)
        GCI_REQUIRE(h.precision == GCI_SPLIT32, "spmm: split operator on an FP64-only matrix");
    GCI_REQUIRE(mode != GCI_OP_FP64 || h.val.p != nullptr, "spmm: FP64 operator on a matrix without FP64 values");
    GCI_REQUIRE(mode != GCI_OP_FP64 || h.precision != GCI_SPLIT32 || h.layout_fp64,
                "spmm_noalloc: the FP64 operator on a split matrix needs its FP64 layout; apply gci_spmm_device "
                "with GCI_OP_FP64 once first");
    if (h.n == 0) return;
    const SpmmShape sh = spmm_shape();
    Pcsc &pc = h.pc;
    GCI_REQUIRE(pc.R == sh.R && pc.target == sh.target && pc.item_slices == item_slices(sh) &&
                    (pc.items.empty() || pc.d_items.n >= pc.items.size()),
                "spmm_noalloc: the matrix's SpMM layout is not prepared for the shape in force; call "
                "gci_spmm_layout first");
    const size_t need = spmm_scratch_need(h, m, mode);
    GCI_REQUIRE(need <= cap, "spmm_noalloc: n x launch width needs " + std::to_string(need) +
                                 " bytes of scratch, above the " + std::to_string(cap) + " bytes available");
    const int W = spmm_max_width(mode);
    const int Mmax = spmm_width(std::min(W, m));
    const size_t xt = mode == GCI_OP_SPLIT_BULK ? sizeof(float) : sizeof(double);
    char *xs = scratch;
    double *ys = reinterpret_cast<double *>(scratch + round256((size_t)h.n * Mmax * xt));
    SpmmJob *job = reinterpret_cast<SpmmJob *>(reinterpret_cast<char *>(ys) +
                                               round256((size_t)h.n * Mmax * sizeof(double)));
    for (int c0 = 0; c0 < m; c0 += W) {
        const int mc = std::min(W, m - c0);
        const int M = spmm_width(mc);
        const double *xc = x + (int64_t)c0 * ldx;
        double *yc = y + (int64_t)c0 * ldy;
        // stream order keeps the previous column block's launch ahead of
        // this descriptor write
        hipLaunchKernelGGL(k_put_job, dim3(1), dim3(1), 0, s, job,
                           SpmmJob{pc.scol.p, pc.sstep.p, pc.lrow.p, pc.val.p, pc.hi.p, pc.lo.p, xs, ys, h.n, mc,
                                   mode, 1, pc.full ? 1 : 0, pc.drow.p, pc.sfirst.p});
        const unsigned pb = grid_for(h.n * M);
        if (mode == GCI_OP_SPLIT_BULK)
            hipLaunchKernelGGL(k_pack<float>, dim3(pb), dim3(256), 0, s, h.n, mc, M, xc, (int64_t)ldx,
                               (float *)xs, ys);
        else
            hipLaunchKernelGGL(k_pack<double>, dim3(pb), dim3(256), 0, s, h.n, mc, M, xc, (int64_t)ldx,
                               (double *)xs, ys);
        GCI_HIP(hipGetLastError());
        spmm_items(pc.d_items.p, (int64_t)pc.items.size(), job, M, mode, s);
        GCI_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_unpack, dim3(grid_for(h.n * mc)), dim3(256), 0, s, h.n, mc, M, (const double *)ys,
                           (const double *)h.diag.p, xc, (int64_t)ldx, yc, (int64_t)ldy);
        GCI_HIP(hipGetLastError());
    }
}

void spmm(HMat &h, int m, const double *x, int ldx, double *y, int ldy, int mode, hipStream_t s) {
    GCI_REQUIRE(m >= 1, "spmm: m < 1");
    GCI_REQUIRE(ldx >= h.n && ldy >= h.n, "spmm: leading dimension < n");
    if (mode != GCI_OP_FP64)
        GCI_REQUIRE(h.precision == GCI_SPLIT32, "spmm: split operator on an FP64-only matrix");
    else
        layout_fp64_values(h);
    GCI_REQUIRE(mode != GCI_OP_FP64 || h.val.p != nullptr, "spmm: FP64 operator on a matrix without FP64 values");
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
                jt[md * 8 + mc - 1] = SpmmJob{pc.scol.p, pc.sstep.p, pc.lrow.p, pc.val.p, pc.hi.p, pc.lo.p,
                                              pc.stage_x.p, pc.stage_y.p, h.n, mc, md, 1, pc.full ? 1 : 0,
                                              pc.drow.p, pc.sfirst.p};
        pc.d_jobs.alloc(jt.size());
        GCI_HIP(hipMemcpy(pc.d_jobs.p, jt.data(), jt.size() * sizeof(SpmmJob), hipMemcpyHostToDevice));
    }
    ensure_device_items(h, s);
    for (int c0 = 0; c0 < m; c0 += W) {
        const int mc = std::min(W, m - c0);
        const int M = spmm_width(mc);
        const double *xc = x + (int64_t)c0 * ldx;
        double *yc = y + (int64_t)c0 * ldy;
        const unsigned pb = grid_for(h.n * M);
        if (mode == GCI_OP_SPLIT_BULK)
            hipLaunchKernelGGL(k_pack<float>, dim3(pb), dim3(256), 0, s, h.n, mc, M, xc, (int64_t)ldx,
                               (float *)pc.stage_x.p, pc.stage_y.p);
        else
            hipLaunchKernelGGL(k_pack<double>, dim3(pb), dim3(256), 0, s, h.n, mc, M, xc, (int64_t)ldx,
                               (double *)pc.stage_x.p, pc.stage_y.p);
        GCI_HIP(hipGetLastError());
        // items carry job 0; the table offset selects the descriptor
        spmm_items(pc.d_items.p, (int64_t)pc.items.size(), pc.d_jobs.p + mode * 8 + mc - 1, M, mode, s);
        GCI_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_unpack, dim3(grid_for(h.n * mc)), dim3(256), 0, s, h.n, mc, M,
                           (const double *)pc.stage_y.p, (const double *)h.diag.p, xc, (int64_t)ldx, yc,
                           (int64_t)ldy);
        GCI_HIP(hipGetLastError());
    }
}

}  // namespace gci
