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

// Device-resident block Hamiltonian: off-diagonal lower triangle in CSR
// (0-based columns, strictly below the diagonal, ascending), FP64 diagonal.
struct HMat {
    int64_t n = 0;
    int64_t nnz = 0;          // off-diagonal entries
    int precision = GCI_FP64;
    DBuf<int64_t> rp;         // n+1 row starts
    DBuf<int32_t> col;        // nnz
    DBuf<double> val;         // nnz (FP64 form)
    DBuf<float> hi, lo;       // nnz (split form)
    DBuf<double> diag;        // n (shifted)
    int group = 0;            // lanes per row chosen by the tuner (0 = not tuned)
    hipStream_t stream = nullptr;
    size_t device_bytes() const {
        return rp.bytes() + col.bytes() + val.bytes() + hi.bytes() + lo.bytes() + diag.bytes();
    }
};

// SpMM launcher (spmm.hip): y = H x for m columns, device pointers.
void spmm(const HMat &h, int m, const double *x, int ldx, double *y, int ldy, int mode,
          hipStream_t s);
// Pick the lanes-per-row group for h (autotuned once, cached on h).
int spmm_group(HMat &h, int m, int mode);

}  // namespace gci

struct gci_hmat : gci::HMat {};
