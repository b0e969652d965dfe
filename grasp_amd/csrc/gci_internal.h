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

// Keeps up to a few GiB of freed stream-ordered memory in the device's
// default pool (once per device), so temporaries are reused without returning
// to the driver.
void retain_pool();

// Stream-ordered temporary device buffer (hipMallocAsync / hipFreeAsync on
// one stream): allocation and release queue behind the stream's work and
// never synchronise the device, unlike hipFree.  For build temporaries that
// only that stream touches.
template <class T>
struct SBuf {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t s = nullptr;
    SBuf(size_t count, hipStream_t stream) : s(stream) {
        if (count == 0) count = 1;
        retain_pool();
        if (hipMallocAsync(reinterpret_cast<void **>(&p), count * sizeof(T), s) != hipSuccess) {
            p = nullptr;
            throw Error{GCI_ERR_NOMEM, "hipMallocAsync of " + std::to_string(count * sizeof(T)) + " bytes failed"};
        }
        n = count;
    }
    SBuf(const SBuf &) = delete;
    SBuf &operator=(const SBuf &) = delete;
    ~SBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
};

// Panel-slice view of the strictly-lower off-diagonal part, the layout the
// symmetric SpMM streams.  Rows are cut into panels of R rows; inside a panel
// the entries of one column form a segment.  A panel's segments, longest
// first, are dealt in groups of 64 (one per wavefront lane) into slices, and
// a slice is stored step-major: step k holds entries 4k..4k+3 of each of its
// 64 segments (256 entries), padded with zero-valued entries to the slice's
// longest segment.  The steps of all slices form one stream; a slice is a run
// of steps with the same 64 columns.
struct SpmmItem {
    int64_t step0, step1;    // step range of one work item (inside one panel)
    int64_t slice0, slice1;  // slices overlapping it
    int64_t row0;            // first row of the panel
    int32_t rows;            // panel height
    int32_t job;             // owning job (filled per launch)
};

struct SpmmJob;

// Launch shape of the SpMM: panel height R (1024, 2048 or 4096; the layout is
// built for it), workgroup size, entries per work item, slices per work item
// of a full layout (at most 128) and per chunk of a lower-triangle one (the
// most whose tables a launch stages in LDS at once; fewer when the launch
// width leaves less LDS), and prefetch depth (steps each wavefront keeps in
// flight).
struct SpmmShape {
    int R = 2048;
    int block = 512;
    int64_t target = 131072;
    int nslice = 64;
    int depth = 8;
};

struct Pcsc {
    int R = 0;                  // rows per panel (0 = not built)
    // Both triangles laid out (split-precision matrices).  A full layout
    // keeps each segment's entries in row order and codes rows in two bits
    // (drow: one byte per lane and step) from the segment's first row
    // (sfirst, per slice and lane): steps of 1..3 rows, or an escape entry
    // holding the next row in its value slot for a larger gap.  Work
    // items and wavefront walks start on slice boundaries, where every
    // lane's running row is its sfirst.  The lower-triangle layout (FP64
    // matrices) codes rows as 16-bit LDS offsets (lrow) in a bank schedule.
    bool full = false;
    int64_t nseg = 0, nnz = 0;  // segments, laid-out (unpadded) entries
    int64_t nslice = 0, nstep = 0;
    int64_t target = 0;         // entries per work item the items were cut for
    int item_slices = 0;        // and the most slices per item (full layouts)
    DBuf<int32_t> scol;         // nslice x 64 columns (-1: empty lane)
    DBuf<int64_t> sstep;        // nslice + 1 first steps
    DBuf<uint16_t> lrow;        // nstep x 256: (row - panel start) x 8 (lower layouts)
    DBuf<uint8_t> drow;         // nstep x 64: two-bit row codes, a lane's four per byte (full layouts)
    DBuf<uint16_t> sfirst;      // nslice x 64: first row - panel start of each lane's segment (full layouts)
    DBuf<double> val;           // nstep x 256 (FP64 values, pairs of lanes' halves)
    DBuf<float> hi, lo;         // nstep x 256 (split values)
    std::vector<int64_t> sstep_h;   // host copies for cutting work items
    std::vector<int32_t> spanel_h;  // panel of each slice
    std::vector<SpmmItem> items;    // host copy of the work items
    // single-matrix operator state: staging (n x stage_M), job table, items
    DBuf<char> stage_x;
    DBuf<double> stage_y;
    int stage_M = 0;
    DBuf<SpmmJob> d_jobs;
    DBuf<SpmmItem> d_items;
    size_t bytes() const {
        return scol.bytes() + sstep.bytes() + lrow.bytes() + drow.bytes() + sfirst.bytes() + val.bytes() +
               hi.bytes() + lo.bytes() +
               stage_x.bytes() + stage_y.bytes();
    }
};

// Device-resident block Hamiltonian: off-diagonal lower triangle in CSR
// (0-based columns, strictly below the diagonal, ascending), FP64 diagonal,
// and the panel-slice copy the SpMM uses.
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
    bool layout_fp64 = false; // a split matrix's layout also carries the FP64 values
    hipStream_t stream = nullptr;
    size_t device_bytes() const {
        return rp.bytes() + col.bytes() + val.bytes() + hi.bytes() + lo.bytes() + diag.bytes() + pc.bytes();
    }
};

// One SpMM job: y = H x for m (<= 8) columns of a matrix of n rows,
// through row-major staging buffers xs (fp32 for the bulk operator, else
// fp64) and ys (fp64), both n x M with M the launch width.
struct SpmmJob {
    const int32_t *scol;
    const int64_t *sstep;
    const uint16_t *lrow;
    const double *val;
    const float *hi, *lo;
    const void *xs;
    double *ys;
    int64_t n;
    int m, mode, active;
    int full;  // 1: both triangles laid out (column sums only), 0: the lower one
    const uint8_t *drow;     // full layouts: row deltas
    const uint16_t *sfirst;  // full layouts: first rows per slice and lane
};

// The launch shape in force (defaults, overridable once through the
// GCI_SPMM_R / _BLOCK / _TARGET / _NSLICE / _DEPTH environment variables) and
// its setter.
SpmmShape spmm_shape();
void set_spmm_shape(const SpmmShape &sh);
bool valid_shape(const SpmmShape &sh);
// Builds h.pc for the current shape (idempotent).
void ensure_pcsc(HMat &h);
// Lays the matrix out again if it was built for another panel height,
// ensure_pcsc, then recut_items.
void ensure_items(HMat &h);
// A split matrix's layout carries hi/lo only; before an FP64-operator use it
// is rebuilt once with the FP64 values too.
void layout_fp64_values(HMat &h);
// Re-cuts the work items if the shape's entries per item changed since they
// were cut (drops the device copy of the items); the panel height must match.
void recut_items(HMat &h);
// LDS bytes one workgroup of the kernel needs for M columns in `mode`, and
// the per-workgroup LDS limit.
size_t spmm_lds_bytes(const SpmmShape &sh, int M, int mode);
size_t spmm_lds_max();
// Widest launch width (8, 5, 4, 2, 1) whose LDS fits the shape in `mode`.
int spmm_max_width(int mode);
// Incremented by every set_spmm_shape (engines re-capture their graphs).
uint64_t spmm_shape_version();
// Re-cuts h.pc's work items for about `target` entries per item, inside one
// panel (host only).
void build_items(HMat &h, const SpmmShape &sh);
// Launch width (1, 2, 4, 5 or 8) for m columns.
int spmm_width(int m);
// The panel-slice kernel over a list of work items whose jobs live in device
// memory (jobs[item.job]); stream-ordered, no synchronisation.
void spmm_items(const SpmmItem *d_items, int64_t nitems, const SpmmJob *d_jobs, int M, int mode, hipStream_t s);
// Single-matrix SpMM on device pointers (column-major x, y).
void spmm(HMat &h, int m, const double *x, int ldx, double *y, int ldy, int mode, hipStream_t s);
// The matrix's work items in device memory (allocated and uploaded if missing).
void ensure_device_items(HMat &h, hipStream_t s);
// Bytes of caller-provided scratch spmm_noalloc needs for m columns in `mode`.
size_t spmm_scratch_need(const HMat &h, int m, int mode);
// spmm with its staging buffers and job descriptor in `scratch` (cap bytes,
// 256-byte aligned device memory): allocates nothing; the matrix's layout
// and device items must already exist for the shape in force.
void spmm_noalloc(HMat &h, int m, const double *x, int ldx, double *y, int ldy, int mode, hipStream_t s,
                  char *scratch, size_t cap);

// The value of every transverse Breit term of a block's pairs, in SETHAM's
// order (breit.hip): NaN where |COEFF| <= 1e-12 or the pair is skipped.
void breit_term_values(gci_breit *b, int64_t npairs, const int32_t *pic, const int32_t *pir, int lfordr, int iccut,
                       const gci_breit_lists *bl, std::vector<double> &tv);

}  // namespace gci

struct gci_hmat : gci::HMat {};
