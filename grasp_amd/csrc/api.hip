// api.hip -- C ABI glue: handles, uploads, SPICMV2 entry points.
#include <cmath>
#include <cstring>
#include <mutex>

#include "gci_internal.h"

namespace gci {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// Scratch of the no-allocation SpMM entry point: one array per device,
// resident from module load until the process ends.  Host code never uses
// the array's name as an address: its device address is read once per
// device with hipGetSymbolAddress (gci_init, or the entry point's first
// call on a device) and kept in g_scratch_ptr.
constexpr size_t NOALLOC_SCRATCH_BYTES = GCI_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static char *g_scratch_ptr[MAX_DEVICES] = {};

char *scratch_of_current_device() {
    int dev = 0;
    GCI_HIP(hipGetDevice(&dev));
    GCI_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "scratch: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            throw Error{GCI_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        g_scratch_ptr[dev] = static_cast<char *>(p);
    }
    return g_scratch_ptr[dev];
}

namespace {

__global__ void k_split(int64_t nnz, const double *__restrict__ v, float *__restrict__ hi,
                        float *__restrict__ lo) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double a = v[i];
        const float h = (float)a;
        hi[i] = h;
        lo[i] = (float)(a - (double)h);
    }
}

}  // namespace

// Builds the device matrix from .res-style rows (diagonal last).
HMat *upload(int64_t n, const int64_t *iendc, const int32_t *irow, const double *emt, int base,
             double shift, int precision) {
    GCI_REQUIRE(n >= 0, "hmat_upload: n < 0");
    GCI_REQUIRE(precision == GCI_FP64 || precision == GCI_SPLIT32, "hmat_upload: bad precision");
    GCI_REQUIRE(base == 0 || base == 1, "hmat_upload: index_base must be 0 or 1");
    GCI_REQUIRE(n == 0 || (iendc && irow && emt), "hmat_upload: null array");
    GCI_REQUIRE(n == 0 || iendc[0] == 0, "hmat_upload: iendc[0] must be 0");
    const int64_t total = n ? iendc[n] : 0;
    const int64_t nnz = total - n;
    GCI_REQUIRE(nnz >= 0, "hmat_upload: fewer elements than rows");
    std::vector<int64_t> rp(n + 1);
    std::vector<int32_t> col((size_t)std::max<int64_t>(nnz, 1));
    std::vector<double> val((size_t)std::max<int64_t>(nnz, 1));
    std::vector<double> diag((size_t)std::max<int64_t>(n, 1));
    int64_t k = 0;
    rp[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t b = iendc[i], e = iendc[i + 1];
        GCI_REQUIRE(e > b, "hmat_upload: row " + std::to_string(i + 1) + " has no diagonal");
        GCI_REQUIRE((int64_t)irow[e - 1] - base == i,
                    "hmat_upload: last element of row " + std::to_string(i + 1) + " is not diagonal");
        int64_t prev = -1;
        for (int64_t t = b; t < e - 1; ++t) {
            const int64_t c = (int64_t)irow[t] - base;
            GCI_REQUIRE(c >= 0 && c < i, "hmat_upload: column outside the strict lower triangle in row " +
                                             std::to_string(i + 1));
            GCI_REQUIRE(c > prev, "hmat_upload: columns not ascending in row " + std::to_string(i + 1));
            prev = c;
            GCI_REQUIRE(std::isfinite(emt[t]), "hmat_upload: non-finite element in row " + std::to_string(i + 1));
            col[k] = (int32_t)c;
            val[k] = emt[t];
            ++k;
        }
        GCI_REQUIRE(std::isfinite(emt[e - 1]) && std::isfinite(shift),
                    "hmat_upload: non-finite diagonal in row " + std::to_string(i + 1));
        diag[i] = emt[e - 1] + shift;
        rp[i + 1] = k;
    }
    auto *h = new gci_hmat();
    try {
        h->n = n;
        h->nnz = nnz;
        h->precision = precision;
        // a non-blocking stream and stream-ordered copies: nothing here
        // touches the legacy stream, which a server's worker thread may be
        // capturing an iteration graph beside
        GCI_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        hipStream_t s = h->stream;
        h->rp.alloc(n + 1);
        h->col.alloc(nnz);
        h->diag.alloc(n);
        h->val.alloc(nnz);
        GCI_HIP(hipMemcpyAsync(h->rp.p, rp.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
        if (nnz) GCI_HIP(hipMemcpyAsync(h->col.p, col.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
        if (n) GCI_HIP(hipMemcpyAsync(h->diag.p, diag.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
        if (nnz) GCI_HIP(hipMemcpyAsync(h->val.p, val.data(), nnz * sizeof(double), hipMemcpyHostToDevice, s));
        if (precision == GCI_SPLIT32) {
            h->hi.alloc(nnz);
            h->lo.alloc(nnz);
            if (nnz) {
                int64_t blocks = std::min<int64_t>((nnz + 255) / 256, 8192);
                hipLaunchKernelGGL(k_split, dim3((unsigned)blocks), dim3(256), 0, h->stream, nnz, h->val.p,
                                   h->hi.p, h->lo.p);
                GCI_HIP(hipGetLastError());
            }
        }
        GCI_HIP(hipStreamSynchronize(h->stream));
    } catch (...) {
        if (h->stream) hipStreamDestroy(h->stream);
        delete h;
        throw;
    }
    return h;
}

void destroy(HMat *h) {
    if (!h) return;
    if (h->stream) hipStreamDestroy(h->stream);
    delete static_cast<gci_hmat *>(h);
}

}  // namespace gci

using namespace gci;

extern "C" {

const char *gci_last_error(void) { return gci::g_err.c_str(); }

int gci_init(int device) {
    return guarded([&] {
        GCI_HIP(hipSetDevice(device));
        GCI_HIP(hipFree(nullptr));
        (void)scratch_of_current_device();  // the no-allocation entry point's array on this device
    });
}

int gci_host_alloc(size_t bytes, void **ptr) {
    return guarded([&] {
        GCI_REQUIRE(ptr && bytes > 0, "gci_host_alloc: bad arguments");
        *ptr = nullptr;
        GCI_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    });
}

int gci_host_free(void *ptr) {
    return guarded([&] {
        if (ptr) GCI_HIP(hipHostFree(ptr));
    });
}

int gci_radgrd(int n, double rnt, double h, double *r, double *rp, double *rpor) {
    return guarded([&] {
        GCI_REQUIRE(n > 0 && r && rp && rpor, "gci_radgrd: bad arguments");
        const double eph = std::exp(h);
        double ett = 1.0;
        r[0] = 0.0;
        rpor[0] = 0.0;
        rp[0] = rnt;
        for (int i = 1; i < n + 10; ++i) {
            ett = eph * ett;
            const double ettm1 = ett - 1.0;
            r[i] = rnt * ettm1;
            rp[i] = rnt * ett;
            rpor[i] = ett / ettm1;
        }
    });
}

int gci_hmat_upload(int n, const int64_t *iendc, const int32_t *irow, const double *emt, int index_base,
                    double shift, int precision, gci_hmat **out) {
    return guarded([&] {
        GCI_REQUIRE(out, "gci_hmat_upload: out is null");
        *out = static_cast<gci_hmat *>(upload(n, iendc, irow, emt, index_base, shift, precision));
    });
}

int gci_hmat_free(gci_hmat *h) {
    return guarded([&] { destroy(h); });
}

int gci_hmat_info(const gci_hmat *h, int64_t *n, int64_t *nnz, int *precision) {
    return guarded([&] {
        GCI_REQUIRE(h, "gci_hmat_info: null handle");
        if (n) *n = h->n;
        if (nnz) *nnz = h->nnz;
        if (precision) *precision = h->precision;
    });
}

int gci_hmat_download(const gci_hmat *h, int64_t *iendc, int32_t *irow, double *emt, int base) {
    return guarded([&] {
        GCI_REQUIRE(h && iendc && irow && emt, "gci_hmat_download: null argument");
        std::vector<int64_t> rp(h->n + 1);
        std::vector<int32_t> col((size_t)std::max<int64_t>(h->nnz, 1));
        std::vector<double> val((size_t)std::max<int64_t>(h->nnz, 1)), diag((size_t)std::max<int64_t>(h->n, 1));
        GCI_HIP(hipMemcpy(rp.data(), h->rp.p, (h->n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (h->nnz) {
            GCI_HIP(hipMemcpy(col.data(), h->col.p, h->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
            GCI_HIP(hipMemcpy(val.data(), h->val.p, h->nnz * sizeof(double), hipMemcpyDeviceToHost));
        }
        if (h->n) GCI_HIP(hipMemcpy(diag.data(), h->diag.p, h->n * sizeof(double), hipMemcpyDeviceToHost));
        int64_t k = 0;
        iendc[0] = 0;
        for (int64_t i = 0; i < h->n; ++i) {
            for (int64_t t = rp[i]; t < rp[i + 1]; ++t) {
                irow[k] = col[t] + base;
                emt[k] = val[t];
                ++k;
            }
            irow[k] = (int32_t)(i + base);
            emt[k] = diag[i];
            ++k;
            iendc[i + 1] = k;
        }
    });
}

int gci_spmm_device(gci_hmat *h, int m, const double *dB, int ldb, double *dC, int ldc, int mode) {
    return guarded([&] {
        GCI_REQUIRE(h, "gci_spmm_device: null handle");
        spmm(*h, m, dB, ldb, dC, ldc, mode, h->stream);
    });
}

int gci_spmm_device_noalloc(gci_hmat *h, int m, const double *dB, int ldb, double *dC, int ldc, int mode) {
    return guarded([&] {
        GCI_REQUIRE(h && dB && dC, "gci_spmm_device_noalloc: null argument");
        spmm_noalloc(*h, m, dB, ldb, dC, ldc, mode, h->stream, scratch_of_current_device(),
                     NOALLOC_SCRATCH_BYTES);
    });
}

int gci_spmm(gci_hmat *h, int m, const double *B, double *C, int mode) {
    return guarded([&] {
        GCI_REQUIRE(h && B && C, "gci_spmm: null argument");
        GCI_REQUIRE(m >= 1, "gci_spmm: m < 1");
        const size_t bytes = (size_t)h->n * m * sizeof(double);
        if (h->n == 0) return;
        DBuf<double> dB((size_t)h->n * m), dC((size_t)h->n * m);
        GCI_HIP(hipMemcpyAsync(dB.p, B, bytes, hipMemcpyHostToDevice, h->stream));
        spmm(*h, m, dB.p, (int)h->n, dC.p, (int)h->n, mode, h->stream);
        GCI_HIP(hipMemcpyAsync(C, dC.p, bytes, hipMemcpyDeviceToHost, h->stream));
        GCI_HIP(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
