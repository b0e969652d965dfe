// tune.hip -- launch shape of the symmetric SpMM and its autotuner.
//
// No counterpart in the reference: SPICMV2 (lib9290/spicmv2.f90) is a plain
// loop.  Here the SpMM's shape -- panel height R, workgroup size, slices
// and entries per work item, prefetch depth -- decides LDS occupancy, the
// number of wavefronts that hide HBM latency and the share of global atomics,
// and the best choice depends on the matrices' sizes and fill.  The tuner times the operator on the caller's matrices and walks
// each parameter in turn (coordinate search), keeping the fastest shape.
#include <algorithm>
#include <vector>

#include "gci_internal.h"

namespace gci {

namespace {

template <class T>
__global__ void k_seeded(int64_t count, uint64_t seed, T *__restrict__ x) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(q + 1);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        x[q] = (T)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
    }
}

// The operator as the Davidson engine runs it: every matrix's work items in
// one launch per column block, over row-major staging buffers (seeded x).
struct TuneSet {
    std::vector<HMat *> mats;
    std::vector<DBuf<char>> xs;
    std::vector<DBuf<double>> ys;
    DBuf<SpmmItem> d_items;
    DBuf<SpmmJob> d_jobs;
    int64_t nitems = 0;
    int m = 1, mode = 0, M = 1, W = 1;
    hipStream_t s = nullptr;

    TuneSet(gci_hmat *const *hs, int nh, int m_, int mode_) : m(m_), mode(mode_) {
        GCI_REQUIRE(hs && nh >= 1 && m >= 1 && m <= 64, "spmm tuner: bad arguments");
        GCI_HIP(hipStreamCreate(&s));
        for (int i = 0; i < nh; ++i) {
            GCI_REQUIRE(hs[i], "spmm tuner: null matrix");
            HMat *h = hs[i];
            if (mode != GCI_OP_FP64)
                GCI_REQUIRE(h->precision == GCI_SPLIT32, "spmm tuner: split operator on an FP64-only matrix");
            else
                GCI_REQUIRE(h->val.p != nullptr, "spmm tuner: FP64 operator on a matrix without FP64 values");
            mats.push_back(h);
            const int64_t cnt = std::max<int64_t>(h->n, 1) * 8;
            xs.emplace_back((size_t)cnt * sizeof(double));
            ys.emplace_back((size_t)cnt);
            const unsigned gb = (unsigned)std::min<int64_t>((cnt + 255) / 256, 4096);
            if (mode == GCI_OP_SPLIT_BULK)
                hipLaunchKernelGGL(k_seeded<float>, dim3(gb), dim3(256), 0, s, cnt, (uint64_t)(i + 1),
                                   (float *)xs.back().p);
            else
                hipLaunchKernelGGL(k_seeded<double>, dim3(gb), dim3(256), 0, s, cnt, (uint64_t)(i + 1),
                                   (double *)xs.back().p);
            GCI_HIP(hipGetLastError());
            GCI_HIP(hipMemsetAsync(ys.back().p, 0, (size_t)cnt * sizeof(double), s));
        }
        GCI_HIP(hipStreamSynchronize(s));
    }
    ~TuneSet() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
    // lays the matrices out for the shape in force and uploads the combined
    // work items and the jobs
    void prepare() {
        W = spmm_max_width(mode);
        M = spmm_width(std::min(W, m));
        std::vector<SpmmItem> items;
        std::vector<SpmmJob> jobs;
        for (size_t i = 0; i < mats.size(); ++i) {
            HMat &h = *mats[i];
            ensure_items(h);
            const Pcsc &pc = h.pc;
            jobs.push_back(SpmmJob{pc.scol.p, pc.sstep.p, pc.lrow.p, pc.val.p, pc.hi.p, pc.lo.p, xs[i].p,
                                   ys[i].p, h.n, std::min(W, m), mode, 1, pc.full ? 1 : 0, pc.drow.p,
                                   pc.sfirst.p});
            for (SpmmItem it : pc.items) {
                it.job = (int32_t)i;
                items.push_back(it);
            }
        }
        nitems = (int64_t)items.size();
        d_items.alloc(std::max<int64_t>(nitems, 1));
        d_jobs.alloc(jobs.size());
        if (nitems)
            GCI_HIP(hipMemcpy(d_items.p, items.data(), nitems * sizeof(SpmmItem), hipMemcpyHostToDevice));
        GCI_HIP(hipMemcpy(d_jobs.p, jobs.data(), jobs.size() * sizeof(SpmmJob), hipMemcpyHostToDevice));
    }
    void run_once() {
        for (int c0 = 0; c0 < m; c0 += W) spmm_items(d_items.p, nitems, d_jobs.p, M, mode, s);
    }
    // mean ms of one m-column application (kernel launches only), after one
    // untimed application
    double time(int reps) {
        prepare();
        run_once();
        GCI_HIP(hipStreamSynchronize(s));
        hipEvent_t e0, e1;
        GCI_HIP(hipEventCreate(&e0));
        GCI_HIP(hipEventCreate(&e1));
        GCI_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) run_once();
        GCI_HIP(hipEventRecord(e1, s));
        GCI_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        GCI_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return (double)ms / std::max(reps, 1);
    }
};

// a valid shape always runs: the operator narrows its column blocks until
// the launch fits the LDS (at worst one column)
bool fits(const SpmmShape &sh, int, int) { return valid_shape(sh); }

gci_spmm_shape to_c(const SpmmShape &sh) {
    return gci_spmm_shape{sh.R, sh.block, sh.nslice, sh.depth, sh.target};
}

SpmmShape from_c(const gci_spmm_shape &c) {
    SpmmShape sh;
    sh.R = c.R;
    sh.block = c.block;
    sh.nslice = c.nslice;
    sh.depth = c.depth;
    sh.target = c.target;
    return sh;
}

}  // namespace

}  // namespace gci

using namespace gci;

extern "C" {

int gci_spmm_get_shape(gci_spmm_shape *out) {
    return guarded([&] {
        GCI_REQUIRE(out, "gci_spmm_get_shape: null argument");
        *out = to_c(spmm_shape());
    });
}

int gci_spmm_set_shape(const gci_spmm_shape *shape) {
    return guarded([&] {
        GCI_REQUIRE(shape, "gci_spmm_set_shape: null argument");
        set_spmm_shape(from_c(*shape));
    });
}

int gci_spmm_layout(gci_hmat *h, int64_t *nseg, int64_t *nslice, int64_t *stored) {
    return guarded([&] {
        GCI_REQUIRE(h && nseg && nslice && stored, "gci_spmm_layout: null argument");
        ensure_items(*h);
        ensure_device_items(*h, h->stream);
        *nseg = h->pc.nseg;
        *nslice = h->pc.nslice;
        *stored = h->pc.nstep * 256;
    });
}

int gci_spmm_time(gci_hmat *const *mats, int nmats, int m, int mode, int reps, double *ms) {
    return guarded([&] {
        GCI_REQUIRE(ms && reps >= 1, "gci_spmm_time: bad arguments");
        TuneSet ts(mats, nmats, m, mode);
        *ms = ts.time(reps);
    });
}

int gci_spmm_autotune(gci_hmat *const *mats, int nmats, int m, int mode, int reps, gci_spmm_shape *best,
                      double *best_ms) {
    return guarded([&] {
        GCI_REQUIRE(reps >= 1, "gci_spmm_autotune: reps < 1");
        TuneSet ts(mats, nmats, m, mode);
        SpmmShape cur = spmm_shape();
        if (!fits(cur, m, mode)) cur = SpmmShape{};
        GCI_REQUIRE(fits(cur, m, mode), "gci_spmm_autotune: no shape fits this width");
        set_spmm_shape(cur);
        double tbest = ts.time(reps);
        const std::vector<int> blocks{256, 512, 1024}, nslices{8, 16, 32, 64}, Rs{1024, 2048, 4096}, depths{2, 4, 8};
        const std::vector<int64_t> targets{32768, 65536, 131072, 262144};
        auto try_shape = [&](const SpmmShape &cand) {
            if (!fits(cand, m, mode)) return;
            set_spmm_shape(cand);
            const double t = ts.time(reps);
            if (t < tbest) {
                tbest = t;
                cur = cand;
            }
        };
        for (int pass = 0; pass < 2; ++pass) {
            for (int v : depths) { SpmmShape c = cur; c.depth = v; try_shape(c); }
            for (int v : blocks) { SpmmShape c = cur; c.block = v; try_shape(c); }
            for (int v : nslices) { SpmmShape c = cur; c.nslice = v; try_shape(c); }
            for (int64_t v : targets) { SpmmShape c = cur; c.target = v; try_shape(c); }
            for (int v : Rs) { SpmmShape c = cur; c.R = v; try_shape(c); }
        }
        set_spmm_shape(cur);
        ts.prepare();  // leave the matrices laid out for the chosen shape
        if (best) *best = to_c(cur);
        if (best_ms) *best_ms = tbest;
    });
}

}  // extern "C"
