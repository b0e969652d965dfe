// hasm.hip -- Hamiltonian assembly from CPU-produced angular lists.
//
// Replaces the SETHAM element loop (rci90/setham_gg.f90:162-385) for the
// Dirac-Coulomb operator.  The angular lists (ONESCALAR's IA, IB, TSHELL and
// RKCO_GG's LABEL/COEFF per CSF pair) are consumed exactly as the producer
// emits them.  One lane reduces one pair's terms in list order (segmented
// gather-multiply-reduce: one-body I(a,b) and two-body R^k gathers from the
// device-resident tables, RKINTC canonicalisation and binary search), with
// floating-point contraction off, so each element equals the sequential
// Fortran sum bit for bit given the same integral tables.  Kept elements are
// compacted into the lower-triangle CSR by an exclusive scan over the pairs
// (pairs are sorted by row, then column, so the compacted order is the CSR
// order).
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

#include "gci_internal.h"

#pragma clang fp contract(off)

namespace gci {

HMat *upload(int64_t n, const int64_t *iendc, const int32_t *irow, const double *emt, int base, double shift,
             int precision);

namespace {

constexpr double CUTOFF = 1.0e-12;  // setham_gg.f90:105

struct AsmIn {
    int64_t npairs;
    int nw;
    int lfordr, iccut;
    const int32_t *pic, *pir, *oia, *oib;
    const int64_t *tsoff;
    const double *tsh;
    const int64_t *tbb, *tbe;
    const int32_t *lab;
    const double *cof;
    const double *iab;
    const int32_t *kst;
    const int32_t *keys;
    const double *vals;
    int j2max;
    // transverse Breit terms (optional): per pair [bbb, bbe) of the 6-label
    // entries, their COEFF and integral values (tv)
    const int64_t *bbb, *bbe;
    const int32_t *blab;
    const double *bcof, *btv;
    int bstride;  // element stride of bcof and btv (2 when interleaved)
};

// RKINTC (rci90/rkintc.f90:45-103); returns false where the Fortran STOPs.
__device__ bool rkintc(const AsmIn &in, int ia, int ib, int ic, int id, int k, double &out) {
    if (k < 0 || k > in.j2max) return false;
    const int key = in.nw + 1;
    int t;
    if (ia > ic) { t = ic; ic = ia; ia = t; }
    if (ib > id) { t = id; id = ib; ib = t; }
    if (ia > ib) {
        t = ib; ib = ia; ia = t;
        t = id; id = ic; ic = t;
    }
    const int index = ((ia * key + ib) * key + ic) * key + id;
    int jl = in.kst[k], ju = in.kst[k + 1] - 1;
    if (ju < jl) return false;
    if (index < in.keys[jl - 1] || index > in.keys[ju - 1]) return false;
    while (ju - jl > 1) {
        const int jm = (ju + jl) / 2;
        if (in.keys[jm - 1] > index) ju = jm;
        else jl = jm;
    }
    int loc;
    if (index == in.keys[ju - 1]) loc = ju;
    else if (index == in.keys[jl - 1]) loc = jl;
    else return false;
    out = in.vals[loc - 1];
    return true;
}

__global__ __launch_bounds__(256) void k_elements(AsmIn in, double *__restrict__ elem, int *__restrict__ keep_off,
                                                  int *__restrict__ err, double *__restrict__ core) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < in.npairs;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int ic = in.pic[p], ir = in.pir[p];
        if (in.lfordr && ir > in.iccut && ir != ic) {  // setham_gg.f90:172-174
            keep_off[p] = 0;
            elem[p] = 0.0;
            continue;
        }
        double e = 0.0;
        const int ia = in.oia[p], ib = in.oib[p];
        if (ia != 0) {  // setham_gg.f90:221-261
            const double *ts = in.tsh + in.tsoff[p];
            if (ia == ib) {
                for (int a = 0; a < in.nw; ++a) {
                    const double t = ts[a];
                    if (fabs(t) > CUTOFF) e = e + in.iab[(size_t)a * in.nw + a] * t;
                }
            } else {
                const double t = ts[0];
                if (fabs(t) > CUTOFF) {
                    const int lo = ia < ib ? ia : ib, hi = ia < ib ? ib : ia;  // IABINT swap
                    e = e + in.iab[(size_t)(lo - 1) * in.nw + (hi - 1)] * t;
                }
            }
        }
        for (int64_t q = in.tbb[p]; q < in.tbe[p]; ++q) {  // setham_gg.f90:273-292
            const double v = in.cof[q];
            if (fabs(v) > CUTOFF) {
                const int32_t *l = in.lab + 5 * q;
                double tegral = 0.0;
                if (!rkintc(in, l[0], l[1], l[2], l[3], l[4], tegral)) {
                    atomicOr(err, 1);
                    tegral = 0.0;
                }
                e = e + tegral * v;
            }
        }
        if (in.bbb) {  // transverse Breit, setham_gg.f90:299-357
            double els = 0.0;
            for (int64_t q = in.bbb[p]; q < in.bbe[p]; ++q) {
                const double v = in.bcof[in.bstride * q];
                if (fabs(v) > CUTOFF) {
                    const double contr = v * in.btv[in.bstride * q];
                    if (in.blab[6 * q + 5] > 0) e = e + contr;
                    else els = els + contr;
                }
            }
            core[p] = els;
        }
        elem[p] = e;
        keep_off[p] = (ir != ic) && (fabs(e) > CUTOFF);  // setham_gg.f90:364
    }
}

__global__ __launch_bounds__(256) void k_compact(int64_t npairs, const int32_t *__restrict__ pic,
                                                 const int32_t *__restrict__ pir, const double *__restrict__ elem,
                                                 const int *__restrict__ keep_off, const int64_t *__restrict__ pos,
                                                 int32_t *__restrict__ col, double *__restrict__ val,
                                                 double *__restrict__ diag) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npairs;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int ic = pic[p], ir = pir[p];
        if (ir == ic) {
            diag[ic - 1] = elem[p];
        } else if (keep_off[p]) {
            col[pos[p]] = ir - 1;
            val[pos[p]] = elem[p];
        }
    }
}

__global__ void k_row_ptr(int64_t n, const int64_t *__restrict__ row_first, const int64_t *__restrict__ pos,
                          int64_t total, int64_t *__restrict__ rp) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n; i += (int64_t)gridDim.x * blockDim.x)
        rp[i] = (i < n) ? pos[row_first[i]] : total;
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
DBuf<T> to_dev(const T *h, size_t n) {
    DBuf<T> d(n);
    if (n) GCI_HIP(hipMemcpy(d.p, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

}  // namespace

}  // namespace gci

using namespace gci;

namespace {

// gci_h_assemble, with the transverse Breit terms when bl is given
void assemble(int ncf, int nw, const gci_pair_lists *L, const gci_breit_lists *bl, gci_breit *br,
              const double *iab, const int32_t *kstart, const int32_t *keys, const double *vals, int64_t nkeys,
              int lfordr, int iccut, int precision, gci_hmat **out, double *eav, double *elsto, int64_t *nelmnt) {
    {
        GCI_REQUIRE(L && iab && kstart && out && eav, "gci_h_assemble: null argument");
        GCI_REQUIRE(ncf >= 1 && nw >= 1 && nw <= 127, "gci_h_assemble: bad ncf/nw");
        GCI_REQUIRE(precision == GCI_FP64 || precision == GCI_SPLIT32, "gci_h_assemble: bad precision");
        const int64_t np = L->npairs;
        GCI_REQUIRE(np >= ncf, "gci_h_assemble: fewer pairs than rows");
        // list validation (host, O(npairs)): rows ascending, columns ascending
        // within a row, ir <= ic, the diagonal last in each row
        std::vector<int64_t> row_first(ncf, -1);
        int prev_ic = 0, prev_ir = 0;
        for (int64_t p = 0; p < np; ++p) {
            const int ic = L->pair_ic[p], ir = L->pair_ir[p];
            GCI_REQUIRE(ic >= 1 && ic <= ncf && ir >= 1 && ir <= ic, "gci_h_assemble: pair out of the lower triangle");
            if (ic != prev_ic) {
                GCI_REQUIRE(ic == prev_ic + 1 && (prev_ic == 0 || prev_ir == prev_ic),
                            "gci_h_assemble: rows must be consecutive and end with their diagonal");
                row_first[ic - 1] = p;
            } else {
                GCI_REQUIRE(ir > prev_ir, "gci_h_assemble: columns not ascending");
            }
            prev_ic = ic;
            prev_ir = ir;
            const int ia = L->one_ia[p], ib = L->one_ib[p];
            GCI_REQUIRE(ia >= 0 && ia <= nw && ib >= 0 && ib <= nw, "gci_h_assemble: one-body label out of range");
            const int64_t cnt = ia == 0 ? 0 : (ia == ib ? nw : 1);
            GCI_REQUIRE(cnt == 0 || (L->tsh_off[p] >= 0 && L->tsh_off[p] + cnt <= L->ntshell),
                        "gci_h_assemble: TSHELL offset out of range");
            GCI_REQUIRE(L->tb_beg[p] >= 0 && L->tb_beg[p] <= L->tb_end[p] && L->tb_end[p] <= L->nterms,
                        "gci_h_assemble: two-body range out of bounds");
        }
        GCI_REQUIRE(prev_ic == ncf && prev_ir == ncf, "gci_h_assemble: last row incomplete");
        int j2max = 0;
        while (j2max < 21 && kstart[j2max + 1] <= nkeys + 1 && kstart[j2max + 1] >= kstart[j2max]) {
            if (kstart[j2max + 1] == nkeys + 1) break;
            ++j2max;
        }
        for (int64_t q = 0; q < L->nterms; ++q) {
            const int32_t *l = L->label + 5 * q;
            GCI_REQUIRE(l[0] >= 1 && l[0] <= nw && l[1] >= 1 && l[1] <= nw && l[2] >= 1 && l[2] <= nw && l[3] >= 1 &&
                            l[3] <= nw && l[4] >= 0 && l[4] <= 20,
                        "gci_h_assemble: two-body label out of range");
        }
        std::vector<double> tv;
        if (bl) {
            GCI_REQUIRE(br && bl->bt_beg && bl->bt_end && (bl->nterms == 0 || (bl->label && bl->coeff)),
                        "gci_h_assemble_tb: null Breit argument");
            for (int64_t p = 0; p < np; ++p)
                GCI_REQUIRE(bl->bt_beg[p] >= 0 && bl->bt_beg[p] <= bl->bt_end[p] && bl->bt_end[p] <= bl->nterms,
                            "gci_h_assemble_tb: Breit range out of bounds");
            for (int64_t q = 0; q < bl->nterms; ++q) {
                const int32_t *l = bl->label + 6 * q;
                const int t = l[5] < 0 ? -l[5] : l[5];
                GCI_REQUIRE(l[0] >= 1 && l[0] <= nw && l[1] >= 1 && l[1] <= nw && l[2] >= 1 && l[2] <= nw &&
                                l[3] >= 1 && l[3] <= nw && l[4] >= 0 && l[4] <= 20 && t >= 1 && t <= 6,
                            "gci_h_assemble_tb: Breit label out of range");
            }
            // the integrals every kept term needs, in SETHAM's order
            breit_term_values(br, np, L->pair_ic, L->pair_ir, lfordr, iccut, bl, tv);
        }
        hipStream_t s = nullptr;
        auto dpic = to_dev(L->pair_ic, np), dpir = to_dev(L->pair_ir, np);
        auto doia = to_dev(L->one_ia, np), doib = to_dev(L->one_ib, np);
        auto dtso = to_dev(L->tsh_off, np);
        auto dtsh = to_dev(L->tshell, (size_t)L->ntshell);
        auto dtbb = to_dev(L->tb_beg, np), dtbe = to_dev(L->tb_end, np);
        auto dlab = to_dev(L->label, (size_t)L->nterms * 5);
        auto dcof = to_dev(L->coeff, (size_t)L->nterms);
        auto diab = to_dev(iab, (size_t)nw * nw);
        auto dkst = to_dev(kstart, (size_t)j2max + 2);
        auto dkeys = to_dev(keys, (size_t)nkeys);
        auto dvals = to_dev(vals, (size_t)nkeys);
        AsmIn in{np,     nw,     lfordr,  iccut,  dpic.p,  dpir.p,  doia.p, doib.p, dtso.p, dtsh.p,
                 dtbb.p, dtbe.p, dlab.p,  dcof.p, diab.p,  dkst.p,  dkeys.p, dvals.p, j2max,
                 nullptr, nullptr, nullptr, nullptr, nullptr, 1};
        DBuf<int64_t> dbbb, dbbe;
        DBuf<int32_t> dblab;
        DBuf<double> dbcof, dbtv, dcore;
        if (bl) {
            dbbb = to_dev(bl->bt_beg, np);
            dbbe = to_dev(bl->bt_end, np);
            dblab = to_dev(bl->label, (size_t)bl->nterms * 6);
            dbcof = to_dev(bl->coeff, (size_t)bl->nterms);
            dbtv = to_dev(tv.data(), tv.size());
            dcore.alloc(np);
            in.bbb = dbbb.p;
            in.bbe = dbbe.p;
            in.blab = dblab.p;
            in.bcof = dbcof.p;
            in.btv = dbtv.p;
        }
        DBuf<double> elem(np);
        DBuf<int> keep(np), err(1);
        DBuf<int64_t> pos(np + 1);
        GCI_HIP(hipMemset(err.p, 0, sizeof(int)));
        const int64_t blocks = std::min<int64_t>((np + 255) / 256, 8192);
        hipLaunchKernelGGL(k_elements, dim3((unsigned)blocks), dim3(256), 0, s, in, elem.p, keep.p, err.p, dcore.p);
        GCI_HIP(hipGetLastError());
        // exclusive scan of the off-diagonal keep flags -> CSR positions
        size_t tmp_bytes = 0;
        GCI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, keep.p, pos.p, np + 1, s));
        DBuf<char> tmp(tmp_bytes);
        // keep has np entries; scan np+1 so pos[np] is the total: use a
        // transform iterator over keep with a zero past the end
        {
            DBuf<int> keep1(np + 1);
            GCI_HIP(hipMemcpyAsync(keep1.p, keep.p, np * sizeof(int), hipMemcpyDeviceToDevice, s));
            GCI_HIP(hipMemsetAsync(keep1.p + np, 0, sizeof(int), s));
            GCI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, keep1.p, pos.p, np + 1, s));
        }
        int herr = 0;
        int64_t total = 0;
        GCI_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipMemcpyAsync(&total, pos.p + np, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipStreamSynchronize(s));
        if (herr) throw Error{GCI_ERR_NOTFOUND, "Rkintc Integral not found"};
        auto *h = new gci_hmat();
        try {
            h->n = ncf;
            h->nnz = total;
            h->precision = precision;
            GCI_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
            h->rp.alloc(ncf + 1);
            h->col.alloc(total);
            h->val.alloc(total);
            h->diag.alloc(ncf);
            auto drf = to_dev(row_first.data(), row_first.size());
            hipLaunchKernelGGL(k_compact, dim3((unsigned)blocks), dim3(256), 0, s, np, dpic.p, dpir.p, elem.p, keep.p,
                               pos.p, h->col.p, h->val.p, h->diag.p);
            hipLaunchKernelGGL(k_row_ptr, dim3((unsigned)std::min<int64_t>((ncf + 256) / 256, 4096)), dim3(256), 0, s,
                               (int64_t)ncf, drf.p, pos.p, total, h->rp.p);
            GCI_HIP(hipGetLastError());
            // EAV = sum of diagonals / NCF (setham_gg.f90:382, genmat2.f90:73)
            std::vector<double> dg(ncf);
            GCI_HIP(hipMemcpyAsync(dg.data(), h->diag.p, ncf * sizeof(double), hipMemcpyDeviceToHost, s));
            GCI_HIP(hipStreamSynchronize(s));
            double sum = 0.0;
            for (int i = 0; i < ncf; ++i) sum = sum + dg[i];
            // ELSTO: the core Breit terms summed in SETHAM's pair order
            double els = 0.0;
            if (bl) {
                std::vector<double> cs(np);
                GCI_HIP(hipMemcpy(cs.data(), dcore.p, np * sizeof(double), hipMemcpyDeviceToHost));
                for (int64_t p = 0; p < np; ++p)
                    if (!(lfordr && L->pair_ir[p] > iccut && L->pair_ir[p] != L->pair_ic[p])) els = els + cs[p];
            }
            if (elsto) *elsto = els;
            // EAV = sum / NCF + ELSTO (genmat2.f90:73)
            const double e = sum / (double)ncf + els;
            *eav = e;
            if (precision == GCI_SPLIT32) {
                h->hi.alloc(total);
                h->lo.alloc(total);
            }
            const int64_t nb = std::min<int64_t>((std::max<int64_t>(ncf, total) + 255) / 256, 8192);
            hipLaunchKernelGGL(k_shift_split, dim3((unsigned)nb), dim3(256), 0, s, (int64_t)ncf, h->diag.p, -e, total,
                               h->val.p, precision == GCI_SPLIT32 ? h->hi.p : nullptr,
                               precision == GCI_SPLIT32 ? h->lo.p : nullptr);
            GCI_HIP(hipGetLastError());
            GCI_HIP(hipStreamSynchronize(s));
            if (nelmnt) *nelmnt = total + ncf;
        } catch (...) {
            if (h->stream) (void)hipStreamDestroy(h->stream);
            delete h;
            throw;
        }
        *out = h;
    }
}

}  // namespace

extern "C" int gci_h_assemble(int ncf, int nw, const gci_pair_lists *L, const double *iab, const int32_t *kstart,
                              const int32_t *keys, const double *vals, int64_t nkeys, int lfordr, int iccut,
                              int precision, gci_hmat **out, double *eav, int64_t *nelmnt) {
    return guarded([&] {
        assemble(ncf, nw, L, nullptr, nullptr, iab, kstart, keys, vals, nkeys, lfordr, iccut, precision, out, eav,
                 nullptr, nelmnt);
    });
}

extern "C" int gci_h_assemble_tb(int ncf, int nw, const gci_pair_lists *L, const gci_breit_lists *bl,
                                 gci_breit *br, const double *iab, const int32_t *kstart, const int32_t *keys,
                                 const double *vals, int64_t nkeys, int lfordr, int iccut, int precision,
                                 gci_hmat **out, double *eav, double *elsto, int64_t *nelmnt) {
    return guarded([&] {
        GCI_REQUIRE(bl && br, "gci_h_assemble_tb: null Breit argument");
        assemble(ncf, nw, L, bl, br, iab, kstart, keys, vals, nkeys, lfordr, iccut, precision, out, eav, elsto,
                 nelmnt);
    });
}

// ------------------------------------------------ streamed assembly
//
// The angular producer runs on the CPU row by row (setham_gg.f90:162-395).
// gci_hstream_* let it hand its lists over in chunks of pairs while it
// produces the next ones: each push copies the chunk into one of two
// page-locked staging slots, queues its upload and the element kernel on the
// stream, and returns; the slot is reused only after the kernel that read it
// has finished (an event).  Element values and keep flags accumulate on the
// device; the end compacts them into the block exactly as gci_h_assemble
// does, so the block is identical for any chunking.

namespace {

struct Stage {
    size_t cap = 0;
    char *host = nullptr;
    DBuf<char> dev;
    hipEvent_t done = nullptr;
    ~Stage() {
        if (done) {
            (void)hipEventSynchronize(done);
            (void)hipEventDestroy(done);
        }
        if (host) (void)hipHostFree(host);
    }
    void ensure(size_t bytes) {
        if (!done) GCI_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        GCI_HIP(hipEventSynchronize(done));  // the last kernel that read this slot
        if (bytes <= cap) return;
        if (host) GCI_HIP(hipHostFree(host));
        host = nullptr;
        const size_t c = std::max<size_t>(bytes, 2 * cap);
        GCI_HIP(hipHostMalloc(reinterpret_cast<void **>(&host), c, hipHostMallocDefault));
        dev.alloc(c);
        cap = c;
    }
};

// byte layout of one chunk in a staging slot (each array 256-byte aligned)
struct ChunkLayout {
    size_t off[14];
    size_t total;
};

template <class T>
std::unique_ptr<SBuf<T>> grow(std::unique_ptr<SBuf<T>> old, size_t used, size_t want, hipStream_t s) {
    auto nb = std::make_unique<SBuf<T>>(want, s);
    if (old && used) GCI_HIP(hipMemcpyAsync(nb->p, old->p, used * sizeof(T), hipMemcpyDeviceToDevice, s));
    return nb;  // the old buffer is released in stream order
}

}  // namespace

struct gci_hstream {
    int ncf = 0, nw = 0, lfordr = 0, iccut = 0, precision = 0, j2max = 0;
    gci_breit *br = nullptr;
    DBuf<double> iab, vals;
    DBuf<int32_t> kst, keys;
    hipStream_t s = nullptr;
    Stage st[2];
    int cur = 0;
    std::unique_ptr<SBuf<double>> elem, core;
    std::unique_ptr<SBuf<int>> keep;
    std::unique_ptr<SBuf<int32_t>> pic, pir;
    int64_t np = 0, cap = 0;
    std::vector<int64_t> row_first;
    std::vector<int32_t> pir_h, pic_h;  // for LFORDR's ELSTO rule at the end
    int prev_ic = 0, prev_ir = 0;
    DBuf<int> err;
    bool with_breit = false;
    ~gci_hstream() {
        if (s) {
            (void)hipStreamSynchronize(s);
            elem.reset();
            core.reset();
            keep.reset();
            pic.reset();
            pir.reset();
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

extern "C" {

int gci_hstream_begin(int ncf, int nw, const double *iab, const int32_t *kstart, const int32_t *keys,
                      const double *vals, int64_t nkeys, int lfordr, int iccut, int precision, gci_breit *br,
                      gci_hstream **out) {
    return guarded([&] {
        GCI_REQUIRE(iab && kstart && out, "gci_hstream_begin: null argument");
        GCI_REQUIRE(ncf >= 1 && nw >= 1 && nw <= 127, "gci_hstream_begin: bad ncf/nw");
        GCI_REQUIRE(precision == GCI_FP64 || precision == GCI_SPLIT32, "gci_hstream_begin: bad precision");
        auto h = std::make_unique<gci_hstream>();
        h->ncf = ncf;
        h->nw = nw;
        h->lfordr = lfordr;
        h->iccut = iccut;
        h->precision = precision;
        h->br = br;
        h->with_breit = br != nullptr;
        int j2max = 0;
        while (j2max < 21 && kstart[j2max + 1] <= nkeys + 1 && kstart[j2max + 1] >= kstart[j2max]) {
            if (kstart[j2max + 1] == nkeys + 1) break;
            ++j2max;
        }
        h->j2max = j2max;
        h->iab = to_dev(iab, (size_t)nw * nw);
        h->kst = to_dev(kstart, (size_t)j2max + 2);
        h->keys = to_dev(keys, (size_t)nkeys);
        h->vals = to_dev(vals, (size_t)nkeys);
        h->row_first.assign(ncf, -1);
        h->err.alloc(1);
        GCI_HIP(hipMemset(h->err.p, 0, sizeof(int)));
        GCI_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
        *out = h.release();
    });
}

int gci_hstream_push(gci_hstream *h, const gci_pair_lists *L, const gci_breit_lists *bl) {
    return guarded([&] {
        GCI_REQUIRE(h && L, "gci_hstream_push: null argument");
        GCI_REQUIRE((bl != nullptr) == h->with_breit, "gci_hstream_push: Breit lists given iff the stream has tables");
        const int64_t cn = L->npairs;
        if (cn == 0) return;
        const int nw = h->nw, ncf = h->ncf;
        // the same validation as gci_h_assemble, continued across chunks
        for (int64_t p = 0; p < cn; ++p) {
            const int ic = L->pair_ic[p], ir = L->pair_ir[p];
            GCI_REQUIRE(ic >= 1 && ic <= ncf && ir >= 1 && ir <= ic, "gci_hstream_push: pair out of the lower triangle");
            if (ic != h->prev_ic) {
                GCI_REQUIRE(ic == h->prev_ic + 1 && (h->prev_ic == 0 || h->prev_ir == h->prev_ic),
                            "gci_hstream_push: rows must be consecutive and end with their diagonal");
                h->row_first[ic - 1] = h->np + p;
            } else {
                GCI_REQUIRE(ir > h->prev_ir, "gci_hstream_push: columns not ascending");
            }
            h->prev_ic = ic;
            h->prev_ir = ir;
            const int ia = L->one_ia[p], ib = L->one_ib[p];
            GCI_REQUIRE(ia >= 0 && ia <= nw && ib >= 0 && ib <= nw, "gci_hstream_push: one-body label out of range");
            const int64_t cnt = ia == 0 ? 0 : (ia == ib ? nw : 1);
            GCI_REQUIRE(cnt == 0 || (L->tsh_off[p] >= 0 && L->tsh_off[p] + cnt <= L->ntshell),
                        "gci_hstream_push: TSHELL offset out of range");
            GCI_REQUIRE(L->tb_beg[p] >= 0 && L->tb_beg[p] <= L->tb_end[p] && L->tb_end[p] <= L->nterms,
                        "gci_hstream_push: two-body range out of bounds");
            if (bl)
                GCI_REQUIRE(bl->bt_beg[p] >= 0 && bl->bt_beg[p] <= bl->bt_end[p] && bl->bt_end[p] <= bl->nterms,
                            "gci_hstream_push: Breit range out of bounds");
        }
        for (int64_t q = 0; q < L->nterms; ++q) {
            const int32_t *l = L->label + 5 * q;
            GCI_REQUIRE(l[0] >= 1 && l[0] <= nw && l[1] >= 1 && l[1] <= nw && l[2] >= 1 && l[2] <= nw && l[3] >= 1 &&
                            l[3] <= nw && l[4] >= 0 && l[4] <= 20,
                        "gci_hstream_push: two-body label out of range");
        }
        std::vector<double> tv;
        if (bl) {
            for (int64_t q = 0; q < bl->nterms; ++q) {
                const int32_t *l = bl->label + 6 * q;
                const int t = l[5] < 0 ? -l[5] : l[5];
                GCI_REQUIRE(l[0] >= 1 && l[0] <= nw && l[1] >= 1 && l[1] <= nw && l[2] >= 1 && l[2] <= nw &&
                                l[3] >= 1 && l[3] <= nw && l[4] >= 0 && l[4] <= 20 && t >= 1 && t <= 6,
                            "gci_hstream_push: Breit label out of range");
            }
            breit_term_values(h->br, cn, L->pair_ic, L->pair_ir, h->lfordr, h->iccut, bl, tv);
        }
        h->pic_h.insert(h->pic_h.end(), L->pair_ic, L->pair_ic + cn);
        h->pir_h.insert(h->pir_h.end(), L->pair_ir, L->pair_ir + cn);
        // result arrays grow in stream order
        if (h->np + cn > h->cap) {
            const int64_t c = std::max<int64_t>(2 * h->cap, h->np + cn);
            h->elem = grow(std::move(h->elem), h->np, c, h->s);
            h->keep = grow(std::move(h->keep), h->np, c, h->s);
            h->pic = grow(std::move(h->pic), h->np, c, h->s);
            h->pir = grow(std::move(h->pir), h->np, c, h->s);
            if (bl) h->core = grow(std::move(h->core), h->np, c, h->s);
            h->cap = c;
        }
        // the chunk into a staging slot: pinned copy, one upload
        const int64_t nb = bl ? bl->nterms : 0;
        const size_t sizes[14] = {cn * 4ul, cn * 4ul, cn * 4ul, cn * 4ul, cn * 8ul, (size_t)L->ntshell * 8,
                                  cn * 8ul, cn * 8ul, (size_t)L->nterms * 20, (size_t)L->nterms * 8,
                                  bl ? cn * 8ul : 0, bl ? cn * 8ul : 0, (size_t)nb * 24, (size_t)nb * 16};
        const void *srcs[14] = {L->pair_ic, L->pair_ir, L->one_ia, L->one_ib, L->tsh_off, L->tshell,
                                L->tb_beg, L->tb_end, L->label, L->coeff, bl ? bl->bt_beg : nullptr,
                                bl ? bl->bt_end : nullptr, bl ? bl->label : nullptr, nullptr};
        ChunkLayout lay{};
        size_t total = 0;
        for (int i = 0; i < 14; ++i) {
            lay.off[i] = total;
            total += (sizes[i] + 255) & ~(size_t)255;
        }
        lay.total = total;
        Stage &sg = h->st[h->cur];
        sg.ensure(total);
        // the host copy into page-locked memory is the push's main cost: large
        // chunks are copied by several threads, in 1 MiB pieces
        {
            struct Piece {
                char *dst;
                const char *src;
                size_t n;
            };
            std::vector<Piece> pieces;
            for (int i = 0; i < 13; ++i)
                for (size_t o = 0; o < sizes[i]; o += (1u << 20))
                    pieces.push_back(Piece{sg.host + lay.off[i] + o, static_cast<const char *>(srcs[i]) + o,
                                           std::min<size_t>(1u << 20, sizes[i] - o)});
            const int nth = (int)std::min<size_t>(8, std::max<size_t>(1, total >> 23));
            auto work = [&](int t) {
                for (size_t j = t; j < pieces.size(); j += nth) std::memcpy(pieces[j].dst, pieces[j].src, pieces[j].n);
            };
            std::vector<std::thread> th;
            for (int t = 1; t < nth; ++t) th.emplace_back(work, t);
            work(0);
            for (auto &x : th) x.join();
        }
        if (bl && nb) {
            // coefficients and their integral values side by side
            double *cv = reinterpret_cast<double *>(sg.host + lay.off[13]);
            for (int64_t q = 0; q < nb; ++q) {
                cv[2 * q] = bl->coeff[q];
                cv[2 * q + 1] = tv[q];
            }
        }
        GCI_HIP(hipMemcpyAsync(sg.dev.p, sg.host, total, hipMemcpyHostToDevice, h->s));
        auto D = [&](int i) { return sg.dev.p + lay.off[i]; };
        GCI_HIP(hipMemcpyAsync(h->pic->p + h->np, D(0), cn * 4, hipMemcpyDeviceToDevice, h->s));
        GCI_HIP(hipMemcpyAsync(h->pir->p + h->np, D(1), cn * 4, hipMemcpyDeviceToDevice, h->s));
        AsmIn in{cn,
                 nw,
                 h->lfordr,
                 h->iccut,
                 reinterpret_cast<const int32_t *>(D(0)),
                 reinterpret_cast<const int32_t *>(D(1)),
                 reinterpret_cast<const int32_t *>(D(2)),
                 reinterpret_cast<const int32_t *>(D(3)),
                 reinterpret_cast<const int64_t *>(D(4)),
                 reinterpret_cast<const double *>(D(5)),
                 reinterpret_cast<const int64_t *>(D(6)),
                 reinterpret_cast<const int64_t *>(D(7)),
                 reinterpret_cast<const int32_t *>(D(8)),
                 reinterpret_cast<const double *>(D(9)),
                 h->iab.p,
                 h->kst.p,
                 h->keys.p,
                 h->vals.p,
                 h->j2max,
                 bl ? reinterpret_cast<const int64_t *>(D(10)) : nullptr,
                 bl ? reinterpret_cast<const int64_t *>(D(11)) : nullptr,
                 bl ? reinterpret_cast<const int32_t *>(D(12)) : nullptr,
                 bl ? reinterpret_cast<const double *>(D(13)) : nullptr,
                 bl ? reinterpret_cast<const double *>(D(13)) + 1 : nullptr,
                 2};
        const int64_t blocks = std::min<int64_t>((cn + 255) / 256, 8192);
        hipLaunchKernelGGL(k_elements, dim3((unsigned)blocks), dim3(256), 0, h->s, in, h->elem->p + h->np,
                           h->keep->p + h->np, h->err.p, bl ? h->core->p + h->np : nullptr);
        GCI_HIP(hipGetLastError());
        GCI_HIP(hipEventRecord(sg.done, h->s));
        h->np += cn;
        h->cur ^= 1;
    });
}

int gci_hstream_end(gci_hstream *h, gci_hmat **out, double *eav, double *elsto, int64_t *nelmnt) {
    return guarded([&] {
        GCI_REQUIRE(h && out && eav, "gci_hstream_end: null argument");
        GCI_REQUIRE(h->prev_ic == h->ncf && h->prev_ir == h->ncf, "gci_hstream_end: last row incomplete");
        hipStream_t s = h->s;
        const int64_t np = h->np, ncf = h->ncf;
        SBuf<int64_t> pos(np + 1, s);
        SBuf<int> keep1(np + 1, s);
        GCI_HIP(hipMemcpyAsync(keep1.p, h->keep->p, np * sizeof(int), hipMemcpyDeviceToDevice, s));
        GCI_HIP(hipMemsetAsync(keep1.p + np, 0, sizeof(int), s));
        size_t tmp_bytes = 0;
        GCI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, keep1.p, pos.p, np + 1, s));
        SBuf<char> tmp(tmp_bytes, s);
        GCI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, keep1.p, pos.p, np + 1, s));
        int herr = 0;
        int64_t total = 0;
        GCI_HIP(hipMemcpyAsync(&herr, h->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipMemcpyAsync(&total, pos.p + np, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        GCI_HIP(hipStreamSynchronize(s));
        if (herr) throw Error{GCI_ERR_NOTFOUND, "Rkintc Integral not found"};
        auto *m = new gci_hmat();
        try {
            m->n = ncf;
            m->nnz = total;
            m->precision = h->precision;
            GCI_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
            m->rp.alloc(ncf + 1);
            m->col.alloc(total);
            m->val.alloc(total);
            m->diag.alloc(ncf);
            auto drf = to_dev(h->row_first.data(), h->row_first.size());
            const int64_t blocks = std::min<int64_t>((np + 255) / 256, 8192);
            hipLaunchKernelGGL(k_compact, dim3((unsigned)blocks), dim3(256), 0, s, np, h->pic->p, h->pir->p,
                               h->elem->p, h->keep->p, pos.p, m->col.p, m->val.p, m->diag.p);
            hipLaunchKernelGGL(k_row_ptr, dim3((unsigned)std::min<int64_t>((ncf + 256) / 256, 4096)), dim3(256), 0, s,
                               ncf, drf.p, pos.p, total, m->rp.p);
            GCI_HIP(hipGetLastError());
            std::vector<double> dg(ncf);
            GCI_HIP(hipMemcpyAsync(dg.data(), m->diag.p, ncf * sizeof(double), hipMemcpyDeviceToHost, s));
            std::vector<double> cs;
            if (h->with_breit) {
                cs.resize(np);
                GCI_HIP(hipMemcpyAsync(cs.data(), h->core->p, np * sizeof(double), hipMemcpyDeviceToHost, s));
            }
            GCI_HIP(hipStreamSynchronize(s));
            double sum = 0.0;
            for (int64_t i = 0; i < ncf; ++i) sum = sum + dg[i];
            double els = 0.0;
            for (int64_t p = 0; p < (int64_t)cs.size(); ++p)
                if (!(h->lfordr && h->pir_h[p] > h->iccut && h->pir_h[p] != h->pic_h[p])) els = els + cs[p];
            if (elsto) *elsto = els;
            const double e = sum / (double)ncf + els;
            *eav = e;
            if (h->precision == GCI_SPLIT32) {
                m->hi.alloc(total);
                m->lo.alloc(total);
            }
            const int64_t nb = std::min<int64_t>((std::max<int64_t>(ncf, total) + 255) / 256, 8192);
            hipLaunchKernelGGL(k_shift_split, dim3((unsigned)nb), dim3(256), 0, s, ncf, m->diag.p, -e, total, m->val.p,
                               h->precision == GCI_SPLIT32 ? m->hi.p : nullptr,
                               h->precision == GCI_SPLIT32 ? m->lo.p : nullptr);
            GCI_HIP(hipGetLastError());
            GCI_HIP(hipStreamSynchronize(s));
            if (nelmnt) *nelmnt = total + ncf;
        } catch (...) {
            if (m->stream) (void)hipStreamDestroy(m->stream);
            delete m;
            throw;
        }
        *out = m;
    });
}

int gci_hstream_free(gci_hstream *h) {
    return guarded([&] { delete h; });
}

}  // extern "C"
