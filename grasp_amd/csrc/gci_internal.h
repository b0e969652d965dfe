// gci_internal.h -- shared internals of libgraspci (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/graspci.h"

namespace gci {

void set_error(const std::string &msg);

struct Error {
    int code;
    std::string msg;
};

#define GCI_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::gci::Error{GCI_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define GCI_REQUIRE(cond, msg)                                     \
    do {                                                           \
        if (!(cond)) throw ::gci::Error{GCI_ERR_ARG, (msg)};       \
    } while (0)

// Runs f and converts exceptions to status codes + thread-local message.
template <class F>
int guarded(F &&f) {
    try {
        f();
        return GCI_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        set_error(e.what());
        return GCI_ERR_ARG;
    }
}

// Owning device buffer.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{GCI_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Panel-CSC view of the strictly-lower off-diagonal part, the layout the
// symmetric SpMM streams: rows are cut into panels of R rows; inside a panel
// the entries are sorted by column, then row, so every column that the panel
// touches is one contiguous segment.  A segment's scatter contribution
// (H^T side) is reduced in registers and issued once; the panel's own rows
// (H side) accumulate in LDS.
struct SpmmItem {
    int64_t seg0, seg1;  // segment range of one work item (a slice of a panel)
    int64_t row0;        // first row of the panel
    int32_t rows;        // panel height
    int32_t job;         // owning job (filled per launch)
};

struct SpmmJob;

// Launch shape of the SpMM: panel height R (a power of two; the layout is
// built for it), lanes per segment group L (8 or 16), workgroup size,
// entries per work item, segments per LDS chunk, and prefetch depth.
struct SpmmShape {
    int R = 1024;
    int L = 16;       // lanes per segment group: L/M entry slots x M vectors
    int block = 512;
    int64_t target = 65536;
    int nseg = 512;   // segments per chunk (their tables are staged in LDS)
    int depth = 4;    // batches each lane keeps in flight
};

struct Pcsc {
    int R = 0;                 // rows per panel (0 = not built)
    int64_t nseg = 0, nnz = 0;
    int64_t npad = 0;          // stored entries including segment padding
    int64_t target = 0;        // entries per work item the items were cut for
    DBuf<int32_t> seg_col;     // nseg
    DBuf<int64_t> seg_start;   // nseg + 1, multiples of 4
    DBuf<uint16_t> lrow;       // npad, row - panel start
    DBuf<double> val;          // npad (FP64 values)
    DBuf<float> hi, lo;        // npad (split values)
    std::vector<int64_t> seg_start_h;  // host copies for cutting work items
    std::vector<int32_t> seg_panel_h;
    std::vector<SpmmItem> items;  // host copy of the work items
    // single-matrix operator state: staging (n x stage_M), job table, items
    DBuf<char> stage_x;
    DBuf<double> stage_y;
    int stage_M = 0;
    DBuf<SpmmJob> d_jobs;
    DBuf<SpmmItem> d_items;
    size_t bytes() const {
        return seg_col.bytes() + seg_start.bytes() + lrow.bytes() + val.bytes() + hi.bytes() + lo.bytes() +
               stage_x.bytes() + stage_y.bytes();
    }
};

// Device-resident block Hamiltonian: off-diagonal lower triangle in CSR
// (0-based columns, strictly below the diagonal, ascending), FP64 diagonal,
// and the panel-CSC copy the SpMM uses.
struct HMat {
    int64_t n = 0;
    int64_t nnz = 0;          // off-diagonal entries
    int precision = GCI_FP64;
    DBuf<int64_t> rp;         // n+1 row starts
    DBuf<int32_t> col;        // nnz
    DBuf<double> val;         // nnz (FP64 form)
    DBuf<float> hi, lo;       // nnz (split form)
    DBuf<double> diag;        // n (shifted)
    Pcsc pc;
    hipStream_t stream = nullptr;
    size_t device_bytes() const {
        return rp.bytes() + col.bytes() + val.bytes() + hi.bytes() + lo.bytes() + diag.bytes() + pc.bytes();
    }
};

// One SpMM job: y = H x for m (<= 8) columns of a matrix, through
// row-major staging buffers xs (fp32 for the bulk operator, else fp64) and
// ys (fp64), both n x M with M the launch width.
struct SpmmJob {
    const int32_t *seg_col;
    const int64_t *seg_start;
    const uint16_t *lrow;
    const double *val;
    const float *hi, *lo;
    const void *xs;
    double *ys;
    int m, mode, active, pad;
};

// The launch shape in force (defaults, overridable once through the
// GCI_SPMM_R / _L / _BLOCK / _TARGET / _NSEG / _DEPTH environment variables) and its
// setter.
SpmmShape spmm_shape();
void set_spmm_shape(const SpmmShape &sh);
bool valid_shape(const SpmmShape &sh);
// Builds h.pc for the current shape (idempotent).
void ensure_pcsc(HMat &h);
// Lays the matrix out again if it was built for another panel height,
// ensure_pcsc, then recut_items.
void ensure_items(HMat &h);
// Re-cuts the work items if the shape's entries per item changed since they
// were cut (drops the device copy of the items); the panel height must match.
void recut_items(HMat &h);
// LDS bytes one workgroup of the kernel needs for M columns in `mode`, and
// the per-workgroup LDS limit.
size_t spmm_lds_bytes(const SpmmShape &sh, int M, int mode);
size_t spmm_lds_max();
// Widest launch width (8, 4, 2, 1) whose LDS fits the shape in `mode`.
int spmm_max_width(int mode);
// Incremented by every set_spmm_shape (engines re-capture their graphs).
uint64_t spmm_shape_version();
// Re-cuts h.pc's work items for about `target` entries per item, inside one
// panel (host only).
void build_items(HMat &h, int64_t target);
// Launch width (1, 2, 4 or 8) for m columns.
int spmm_width(int m);
// The panel-CSC kernel over a list of work items whose jobs live in device
// memory (jobs[item.job]); stream-ordered, no synchronisation.
void spmm_items(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, int M, int mode, hipStream_t s);
// Single-matrix SpMM on device pointers (column-major x, y).
void spmm(HMat &h, int m, const double *x, int ldx, double *y, int ldy, int mode, hipStream_t s);

}  // namespace gci

struct gci_hmat : gci::HMat {};
