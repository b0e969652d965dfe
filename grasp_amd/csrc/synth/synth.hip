// synth.hip -- seeded synthetic CI blocks generated on the device.
//
// Benchmark/test input generation only (libgcisynth.so); not part of the
// rci90 path.  A block is a symmetric matrix whose rows are cut into CSF
// groups of uniform random size in [gmin, gmax]; a lower group pair (G, H)
// is coupled with probability fill/inner and each entry of a coupled pair is
// present with probability inner, so the expected lower fill is `fill`.
// Diagonal ~ N(diag_mean, diag_sd^2), off-diagonals ~ N(0, off_sd^2).  All
// draws come from a counter-based hash of (seed, indices), so a block is a
// pure function of its parameters and can be regenerated anywhere.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "../gci_internal.h"

namespace gci {

namespace {

__host__ __device__ inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

__host__ __device__ inline double u01(uint64_t seed, uint64_t a, uint64_t b, uint64_t tag) {
    uint64_t x = mix64(seed ^ mix64(a * 0x9E3779B97F4A7C15ull + mix64(b + (tag << 56))));
    return ((double)(x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline double gauss(uint64_t seed, uint64_t a, uint64_t b, uint64_t tag) {
    const double u1 = u01(seed, a, b, tag), u2 = u01(seed, a, b, tag + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct SynDev {
    int64_t n;
    uint64_t seed;
    double inner;
    const int64_t *gstart;      // ngroups+1 row starts
    const int32_t *group_of;    // row -> group
    const int64_t *cptr;        // coupled-group CSR per group
    const int32_t *clist;
};

// one wave per row: count (pass 0) or emit (pass 1) the row's entries
__global__ __launch_bounds__(256) void k_rows(SynDev d, int pass, int64_t *__restrict__ cnt,
                                              const int64_t *__restrict__ rp, int32_t *__restrict__ col,
                                              double *__restrict__ val, double off_sd) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    if (row >= d.n) return;
    const int g = d.group_of[row];
    int64_t pos = pass ? rp[row] : 0;
    int64_t total = 0;
    for (int64_t q = d.cptr[g]; q < d.cptr[g + 1]; ++q) {
        const int h = d.clist[q];
        const int64_t c0 = d.gstart[h], c1 = min(d.gstart[h + 1], row);  // strictly lower
        for (int64_t cb = c0; cb < c1; cb += 64) {
            const int64_t c = cb + lane;
            const bool on = c < c1 && u01(d.seed, (uint64_t)row, (uint64_t)c, 1) < d.inner;
            const unsigned long long m = __ballot(on);
            if (pass && on) {
                const int64_t k = pos + __popcll(m & ((1ull << lane) - 1ull));
                col[k] = (int32_t)c;
                val[k] = off_sd * gauss(d.seed, (uint64_t)row, (uint64_t)c, 2);
            }
            const int nb = __popcll(m);
            pos += nb;
            total += nb;
        }
    }
    if (!pass && lane == 0) cnt[row] = total;
}

__global__ void k_diag(int64_t n, uint64_t seed, double mean, double sd, double *__restrict__ diag) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        diag[i] = mean + sd * gauss(seed, (uint64_t)i, 0, 3);
}

__global__ void k_shift_split(int64_t n, double *__restrict__ diag, double shift, int64_t nnz,
                              const double *__restrict__ v, float *__restrict__ hi, float *__restrict__ lo) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n || i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < n) diag[i] = diag[i] + shift;
        if (hi && i < nnz) {
            const double a = v[i];
            const float h = (float)a;
            hi[i] = h;
            lo[i] = (float)(a - (double)h);
        }
    }
}

template <class T>
DBuf<T> to_dev(const std::vector<T> &h) {
    DBuf<T> d(h.size());
    if (!h.empty()) GCI_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

}  // namespace

}  // namespace gci

using namespace gci;

extern "C" {

// Generates a block on the device.  The diagonal is shifted by -EAV (mean
// of the diagonal) when shift_eav != 0, as MANEIG loads it; sorted_diag
// puts the diagonal in ascending order.
int gcs_synth_block(int64_t n, uint64_t seed, double fill, double diag_mean, double diag_sd, double off_sd, int gmin,
                    int gmax, double inner, int precision, int shift_eav, int sorted_diag, double *eav_out,
                    gci_hmat **out) {
    return guarded([&] {
        GCI_REQUIRE(n >= 1 && gmin >= 1 && gmax >= gmin && inner > 0.0 && inner <= 1.0 && out,
                    "gcs_synth_block: bad arguments");
        // groups and coupled pairs (host)
        std::vector<int64_t> gstart{0};
        uint64_t k = 0;
        while (gstart.back() < n) {
            const int64_t sz = gmin + (int64_t)(u01(seed, 7, k++, 4) * (gmax - gmin + 1));
            gstart.push_back(std::min<int64_t>(n, gstart.back() + std::min<int64_t>(sz, gmax)));
        }
        const int ng = (int)gstart.size() - 1;
        const double pc = std::min(1.0, fill / inner);
        std::vector<int64_t> cptr{0};
        std::vector<int32_t> clist;
        for (int g = 0; g < ng; ++g) {
            for (int h = 0; h <= g; ++h)
                if (u01(seed, (uint64_t)g, (uint64_t)h, 5) < pc) clist.push_back(h);
            cptr.push_back((int64_t)clist.size());
        }
        std::vector<int32_t> group_of(n);
        for (int g = 0; g < ng; ++g)
            for (int64_t r = gstart[g]; r < gstart[g + 1]; ++r) group_of[r] = g;
        auto dgs = to_dev(gstart);
        auto dgo = to_dev(group_of);
        auto dcp = to_dev(cptr);
        auto dcl = to_dev(clist);
        SynDev d{n, seed, inner, dgs.p, dgo.p, dcp.p, dcl.p};
        DBuf<int64_t> cnt(n + 1);
        auto *h = new gci_hmat();
        try {
            h->n = n;
            h->precision = precision;
            GCI_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
            hipStream_t s = h->stream;
            GCI_HIP(hipMemsetAsync(cnt.p, 0, (n + 1) * sizeof(int64_t), s));
            const int64_t blocks = (n * 64 + 255) / 256;
            hipLaunchKernelGGL(k_rows, dim3((unsigned)blocks), dim3(256), 0, s, d, 0, cnt.p, (const int64_t *)nullptr,
                               (int32_t *)nullptr, (double *)nullptr, off_sd);
            h->rp.alloc(n + 1);
            size_t tb = 0;
            GCI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.p, h->rp.p, n + 1, s));
            DBuf<char> tmp(tb);
            GCI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt.p, h->rp.p, n + 1, s));
            int64_t nnz = 0;
            GCI_HIP(hipMemcpyAsync(&nnz, h->rp.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
            GCI_HIP(hipStreamSynchronize(s));
            h->nnz = nnz;
            h->col.alloc(nnz);
            h->val.alloc(nnz);
            h->diag.alloc(n);
            hipLaunchKernelGGL(k_rows, dim3((unsigned)blocks), dim3(256), 0, s, d, 1, cnt.p, (const int64_t *)h->rp.p,
                               h->col.p, h->val.p, off_sd);
            hipLaunchKernelGGL(k_diag, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, s, n,
                               seed, diag_mean, diag_sd, h->diag.p);
            GCI_HIP(hipGetLastError());
            if (sorted_diag) {
                // ascending diagonal: the CSF list starts from the lowest
                // configurations, as a list built from a reference does
                DBuf<double> tmpd(n);
                size_t sb = 0;
                GCI_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, sb, h->diag.p, tmpd.p, n, 0, 64, s));
                DBuf<char> st(sb);
                GCI_HIP(hipcub::DeviceRadixSort::SortKeys(st.p, sb, h->diag.p, tmpd.p, n, 0, 64, s));
                GCI_HIP(hipMemcpyAsync(h->diag.p, tmpd.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
                GCI_HIP(hipStreamSynchronize(s));
            }
            std::vector<double> dg(n);
            GCI_HIP(hipMemcpyAsync(dg.data(), h->diag.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
            GCI_HIP(hipStreamSynchronize(s));
            double sum = 0.0;
            for (int64_t i = 0; i < n; ++i) sum += dg[i];
            const double eav = sum / (double)n;
            if (eav_out) *eav_out = eav;
            if (precision == GCI_SPLIT32) {
                h->hi.alloc(nnz);
                h->lo.alloc(nnz);
            }
            const int64_t nb = std::min<int64_t>((std::max<int64_t>(n, nnz) + 255) / 256, 65536);
            hipLaunchKernelGGL(k_shift_split, dim3((unsigned)nb), dim3(256), 0, s, n, h->diag.p,
                               shift_eav ? -eav : 0.0, nnz, h->val.p, precision == GCI_SPLIT32 ? h->hi.p : nullptr,
                               precision == GCI_SPLIT32 ? h->lo.p : nullptr);
            GCI_HIP(hipGetLastError());
            GCI_HIP(hipStreamSynchronize(s));
        } catch (...) {
            if (h->stream) (void)hipStreamDestroy(h->stream);
            delete h;
            throw;
        }
        *out = h;
    });
}

}  // extern "C"
