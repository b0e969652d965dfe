// maneig.hip -- MANEIG dispatch, GDVD entry point and the batched solve.
#include <algorithm>
#include <limits>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <mutex>

#include "dvd.h"

namespace gci {

namespace {

// Dense ns x ns leading block of the symmetric H (column-major, both
// triangles), as INIEST2 and the NCF <= 2000 branch of MANEIG build it
// (lib9290/iniest2.f90:62-72, rci90/maneig.f90:200-213).
__global__ void k_dense_block(int64_t ns, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                              const double *__restrict__ val, const double *__restrict__ diag,
                              double *__restrict__ a) {
    const int64_t i = blockIdx.x;
    if (i >= ns) return;
    for (int64_t e = rp[i] + threadIdx.x; e < rp[i + 1]; e += blockDim.x) {
        const int64_t c = col[e];
        if (c < ns) {
            a[i + c * ns] = val[e];
            a[c + i * ns] = val[e];
        }
    }
    if (threadIdx.x == 0) a[i + i * ns] = diag[i];
}

void lowest_of_leading_block(const HMat &h, int64_t ns, int nev, std::vector<double> &eval,
                             std::vector<double> &evec) {
    DBuf<double> a((size_t)ns * ns), w(nev), z((size_t)ns * nev);
    GCI_HIP(hipMemsetAsync(a.p, 0, a.bytes(), h.stream));
    hipLaunchKernelGGL(k_dense_block, dim3((unsigned)ns), dim3(64), 0, h.stream, ns, h.rp.p, h.col.p, h.val.p,
                       h.diag.p, a.p);
    GCI_HIP(hipGetLastError());
    dense_lowest((int)ns, a.p, nev, w.p, z.p, h.stream);
    eval.resize(nev);
    evec.resize((size_t)ns * nev);
    GCI_HIP(hipMemcpy(eval.data(), w.p, nev * sizeof(double), hipMemcpyDeviceToHost));
    GCI_HIP(hipMemcpy(evec.data(), z.p, evec.size() * sizeof(double), hipMemcpyDeviceToHost));
}

void sign_fix(int64_t n, int nvec, double *evec) {
    // rci90/maneig.f90:532-555: dominant component made positive
    for (int j = 0; j < nvec; ++j) {
        double *v = evec + (size_t)j * n, amax = 0.0;
        int64_t ia = 0;
        for (int64_t i = 0; i < n; ++i) {
            const double wa = std::fabs(v[i]);
            if (wa <= amax) continue;
            amax = wa;
            ia = i;
        }
        if (v[ia] >= 0.0) continue;
        for (int64_t i = 0; i < n; ++i) v[i] = -v[i];
    }
}

DvdSpec resolve(const HMat &h, const gci_dvd_params *p) {
    DvdSpec s;
    s.h = &h;
    s.nume = p->nvec;
    GCI_REQUIRE(s.nume >= 1 && s.nume <= MAXNUME, "davidson: nvec must be in 1.." + std::to_string(MAXNUME));
    s.lim = p->lim > 0 ? p->lim : (int)std::min<int64_t>(h.n, 2 * s.nume + 60);
    s.mblock = p->mblock > 0 ? p->mblock : 1;
    s.maxiter = p->maxiter > 0 ? p->maxiter : (int)std::max<int64_t>(200 * s.nume, h.n / 10);
    s.crite = p->crite > 0 ? p->crite : 1.0e-17;
    s.critc = p->critc > 0 ? p->critc : 1.0e-9;
    s.critr = p->critr > 0 ? p->critr : 1.0e-9;
    s.mode = p->mode;
    s.refine = p->mode != GCI_OP_FP64 && p->refine;
    if (s.mode != GCI_OP_FP64) {
        // the split-precision bulk phase stops at a residual the fp32
        // operator can certify (1e-7); the FP64 refinement applies CRITR
        s.critr_final = s.critr;
        s.critr = std::max(s.critr, 1.0e-7);
    }
    return s;
}

// Runs an engine until every slot is done, restarting split-precision slots
// once with the FP64 operator.  The host stays one iteration ahead: it
// queues the next iteration before reading the state after the previous
// one, so the device does not wait for the host's decisions.
// Slots whose bulk phase has ended wait for their FP64 refinement restart
// until no slot of the engine is still iterating (GCI_REFINE_BATCH=n
// restarts as soon as n slots wait), so that one restart launch sequence
// and one wide FP64 product serve every block, and no iteration carries a
// second operator mode: on the 64-block batch, 8.08 ms per solve against
// 8.20-8.24 ms when half of the slots start the restart and 8.75 ms with a
// restart at every poll.
int refine_batch() {
    const char *v = std::getenv("GCI_REFINE_BATCH");
    return v && *v ? std::max(1, std::atoi(v)) : std::numeric_limits<int>::max();
}

void collect_refinements(DavidsonEngine &eng, const std::vector<int> &done, std::vector<int> &pending, int batch) {
    for (int sl : done)
        if (eng.needs_refine(sl)) pending.push_back(sl);
    if (!pending.empty() && ((int)pending.size() >= batch || eng.iterating() == 0)) {
        eng.refine_restart(pending);
        pending.clear();
    }
}

void run_to_completion(DavidsonEngine &eng, int guard_iters) {
    const int batch = refine_batch();
    std::vector<int> pending;
    int it = 0;
    eng.step(1);
    ++it;
    for (;;) {
        eng.step(1);
        ++it;
        collect_refinements(eng, eng.poll_lagged(), pending, batch);
        if (pending.empty() && eng.all_done()) break;
        GCI_REQUIRE(it <= guard_iters, "davidson: iteration guard exceeded");
    }
    eng.sync();
}
// Several engines, each on its own stream, driven in turn: while one
// engine's problems stream their SpMM the other's latency-bound Davidson
// kernels (Rayleigh-Ritz, Gram-Schmidt, NEWVEC) run beside them.  Each
// problem's arithmetic is the same in any grouping.
void run_to_completion_multi(const std::vector<DavidsonEngine *> &engs, int guard_iters) {
    const size_t E = engs.size();
    const int batch = refine_batch();
    std::vector<std::vector<int>> pending(E);
    std::vector<char> finished(E, 0);
    for (DavidsonEngine *e : engs) e->step(1);
    int it = 1;
    for (;;) {
        bool all = true;
        for (size_t i = 0; i < E; ++i) {
            if (finished[i]) continue;
            DavidsonEngine &eng = *engs[i];
            eng.step(1);
            collect_refinements(eng, eng.poll_lagged(), pending[i], batch);
            if (pending[i].empty() && eng.all_done()) finished[i] = 1;
            else all = false;
        }
        ++it;
        if (all) break;
        GCI_REQUIRE(it <= guard_iters, "davidson: iteration guard exceeded");
    }
    for (DavidsonEngine *e : engs) e->sync();
}

// One GDVD solve on one block; init: n x niv initial vectors (host) or null
// for INIEST2.
void solve_one(HMat &h, const DvdSpec &sp, const double *init, int niv, double *eval, double *evec,
               double *resid, gci_dvd_stats *st) {
    std::vector<double> iv;
    if (!init || niv <= 0) {
        // INIEST2 (lib9290/iniest2.f90): leading min(2000, n) block
        const int64_t ns = std::min<int64_t>(2000, h.n);
        std::vector<double> w, z;
        lowest_of_leading_block(h, ns, sp.nume, w, z);
        iv.assign((size_t)h.n * sp.nume, 0.0);
        for (int j = 0; j < sp.nume; ++j)
            std::copy(z.begin() + (size_t)j * ns, z.begin() + (size_t)(j + 1) * ns, iv.begin() + (size_t)j * h.n);
        init = iv.data();
        niv = sp.nume;
    }
    // the engine gets its own non-blocking stream (it captures graphs on it);
    // the matrix's data were completed on the matrix's stream
    hipStream_t es = nullptr;
    GCI_HIP(hipStreamCreateWithFlags(&es, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } guard{es};
    DavidsonEngine eng(1, es);
    std::vector<int> slots = eng.admit({sp}, {init}, {niv});
    run_to_completion(eng, 2 * sp.maxiter + 128);
    eng.extract(slots[0], eval, evec, resid, st);
}

}  // namespace

}  // namespace gci

using namespace gci;

extern "C" {

int gci_davidson(gci_hmat *h, const gci_dvd_params *p, double *eval, double *evec, double *resid,
                 gci_dvd_stats *st) {
    return guarded([&] {
        GCI_REQUIRE(h && p && eval, "gci_davidson: null argument");
        const auto t0 = std::chrono::steady_clock::now();
        DvdSpec sp = resolve(*h, p);
        if (sp.mode != GCI_OP_FP64) GCI_REQUIRE(h->precision == GCI_SPLIT32, "gci_davidson: split mode needs GCI_SPLIT32");
        gci_dvd_stats local{};
        solve_one(*h, sp, p->niv > 0 ? evec : nullptr, p->niv, eval, evec, resid, &local);
        local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (st) *st = local;
        if (local.ierr & 2048) throw Error{GCI_ERR_NOCONV, "gci_davidson: NLOOPS > MAXITER (IERR 2048)"};
    });
}

int gci_maneig(gci_hmat *h, int nvec, double *eval, double *evec, gci_dvd_stats *st) {
    return guarded([&] {
        GCI_REQUIRE(h && eval && evec, "gci_maneig: null argument");
        GCI_REQUIRE(nvec >= 1 && nvec <= h->n, "gci_maneig: nvec out of range");
        const auto t0 = std::chrono::steady_clock::now();
        gci_dvd_stats local{};
        if (h->n == 1) {
            // trivial case (maneig.f90:136-156)
            GCI_HIP(hipStreamSynchronize(h->stream));
            GCI_HIP(hipMemcpy(eval, h->diag.p, sizeof(double), hipMemcpyDeviceToHost));
            evec[0] = 1.0;
        } else if (h->n <= 2000) {
            // LAPACK branch (maneig.f90:157-242)
            std::vector<double> w, z;
            lowest_of_leading_block(*h, h->n, nvec, w, z);
            std::copy(w.begin(), w.end(), eval);
            std::copy(z.begin(), z.end(), evec);
        } else {
            // DVDSON branch (maneig.f90:262-451).  The reference keeps H in
            // core while NSTORE = NELMNT + NELMNT/2 + (NCF+1)/2 is at most
            // NINCOR = 2^28 and runs GDVD with MBLOCK = 1; above it H stays
            // on disk and MBLOCK = NVEX (maneig.f90:93, 262-266, 416-420).
            // The device holds H either way; the block size follows the rule.
            const int64_t nelmnt = h->nnz + h->n;
            const int64_t nstore = nelmnt + nelmnt / 2 + (h->n + 1) / 2;
            gci_dvd_params p{};
            p.nvec = nvec;
            p.mblock = nstore > NINCOR ? nvec : 1;
            p.mode = GCI_OP_FP64;
            DvdSpec sp = resolve(*h, &p);
            solve_one(*h, sp, nullptr, 0, eval, evec, nullptr, &local);
            if (local.ierr & 2048) throw Error{GCI_ERR_NOCONV, "maneig: DVDSON wrong (IERR 2048)"};
        }
        sign_fix(h->n, nvec, evec);
        local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (st) *st = local;
    });
}

int gci_maneig_mcdhf(gci_hmat *h, int nvect, int nlev, const int32_t *iccmin, int warm, double *eval,
                     double *evec, gci_dvd_stats *st) {
    return guarded([&] {
        GCI_REQUIRE(h && iccmin && eval && evec, "gci_maneig_mcdhf: null argument");
        GCI_REQUIRE(nlev >= 1 && nvect >= 1, "gci_maneig_mcdhf: nlev and nvect must be positive");
        const int64_t n = h->n;
        const int nvex = (int)std::min<int64_t>(nvect, n);
        for (int j = 0; j < nlev; ++j)
            GCI_REQUIRE(iccmin[j] >= 1 && iccmin[j] <= nvex, "gci_maneig_mcdhf: level outside 1..min(nvect, NCF)");
        const auto t0 = std::chrono::steady_clock::now();
        gci_dvd_stats local{};
        if (n == 1) {
            // rmcdhf90/maneig.f90:78-82: EVAL = 0, EVEC = 1
            for (int j = 0; j < nlev; ++j) {
                eval[j] = 0.0;
                evec[j] = 1.0;
            }
        } else {
            // NVEX = NIV = min(NVECT, NCF), LIM = min(NCF, 2 NVECT + 40),
            // MAXITR = min(200 NVECT, NCF), N1000 = 4000
            // (rmcdhf90/maneig.f90:103-113)
            GCI_REQUIRE(nvex <= MAXNUME, "gci_maneig_mcdhf: nvect must be at most " + std::to_string(MAXNUME));
            std::vector<double> all_val(nvex), all_vec;
            std::vector<double> iv;
            const int64_t n1000 = 4000;
            if (!warm || n <= n1000) {
                // INIEST2(N1000) (rmcdhf90/maneig.f90:115-116): for NCF <= 4000
                // the whole block, which is then the answer
                const int64_t ns = std::min<int64_t>(n1000, n);
                std::vector<double> w, z;
                lowest_of_leading_block(*h, ns, nvex, w, z);
                iv.assign((size_t)n * nvex, 0.0);
                for (int j = 0; j < nvex; ++j) {
                    std::copy(z.begin() + (size_t)j * ns, z.begin() + (size_t)(j + 1) * ns, iv.begin() + (size_t)j * n);
                    all_val[j] = w[j];
                }
            } else {
                // warm start: the previous iteration's vectors (orthonormal)
                // at columns ICCMIN (rmcdhf90/maneig.f90:118-124).  The
                // reference leaves the other columns as allocated; here column
                // c holds the unit vector at the c-th smallest diagonal
                // element.
                iv.assign((size_t)n * nvex, 0.0);
                std::vector<char> given(nvex, 0);
                for (int j = 0; j < nlev; ++j) {
                    const int c = iccmin[j] - 1;
                    std::copy(evec + (size_t)j * n, evec + (size_t)(j + 1) * n, iv.begin() + (size_t)c * n);
                    given[c] = 1;
                }
                std::vector<double> d(n);
                GCI_HIP(hipStreamSynchronize(h->stream));
                GCI_HIP(hipMemcpy(d.data(), h->diag.p, n * sizeof(double), hipMemcpyDeviceToHost));
                std::vector<int64_t> idx(n);
                for (int64_t i = 0; i < n; ++i) idx[i] = i;
                std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return d[a] < d[b]; });
                for (int c = 0; c < nvex; ++c)
                    if (!given[c]) iv[(size_t)c * n + idx[c]] = 1.0;
                // INITDVD takes its start block as orthonormal
                // (libdvd90/dvdson.f90:886-905): a filled column is made
                // orthogonal to the given columns and the filled ones before
                // it, two classical Gram-Schmidt passes
                for (int c = 0; c < nvex; ++c) {
                    if (given[c]) continue;
                    double *v = iv.data() + (size_t)c * n;
                    for (int pass = 0; pass < 2; ++pass)
                        for (int q = 0; q < nvex; ++q) {
                            if (q == c || (!given[q] && q > c)) continue;
                            const double *u = iv.data() + (size_t)q * n;
                            double dot = 0.0;
                            for (int64_t i = 0; i < n; ++i) dot += u[i] * v[i];
                            for (int64_t i = 0; i < n; ++i) v[i] -= dot * u[i];
                        }
                    double nrm = 0.0;
                    for (int64_t i = 0; i < n; ++i) nrm += v[i] * v[i];
                    nrm = std::sqrt(nrm);
                    for (int64_t i = 0; i < n; ++i) v[i] /= nrm;
                }
            }
            if (n > n1000) {
                // GDVD with MBLOCK = 1, CRITE 1e-17, CRITC = CRITR = 1e-9
                // (rmcdhf90/maneig.f90:139-162)
                gci_dvd_params p{};
                p.nvec = nvex;
                p.lim = (int)std::min<int64_t>(n, 2 * (int64_t)nvect + 40);
                p.maxiter = (int)std::min<int64_t>(200 * (int64_t)nvect, n);
                p.mblock = 1;
                p.mode = GCI_OP_FP64;
                DvdSpec sp = resolve(*h, &p);
                all_vec.assign((size_t)n * nvex, 0.0);
                solve_one(*h, sp, iv.data(), nvex, all_val.data(), all_vec.data(), nullptr, &local);
                if (local.ierr & 2048) throw Error{GCI_ERR_NOCONV, "maneig (rmcdhf90): DVDSON wrong (IERR 2048)"};
            } else {
                all_vec.swap(iv);
            }
            // the levels ICCMIN (rmcdhf90/maneig.f90:171-176)
            for (int j = 0; j < nlev; ++j) {
                const int c = iccmin[j] - 1;
                eval[j] = all_val[c];
                std::copy(all_vec.begin() + (size_t)c * n, all_vec.begin() + (size_t)(c + 1) * n, evec + (size_t)j * n);
            }
        }
        sign_fix(n, nlev, evec);  // rmcdhf90/maneig.f90:179-196
        local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (st) *st = local;
    });
}

// Batched form: every block's Davidson runs interleaved in one engine
// (split-precision bulk operator when the block holds hi/lo, block
// Davidson with MBLOCK = NEV, convergence on the relative residual
// ||r|| < tol |lambda + EAV|, then an FP64 refinement pass).
int gci_batch_solve(int nblocks, gci_hmat **blocks, const int *nev, const double *tol, const double *eav,
                    double *eval, int64_t *eval_off, double *evec, int64_t *evec_off, gci_dvd_stats *stats) {
    return guarded([&] {
        GCI_REQUIRE(nblocks >= 1 && blocks && nev && tol && eval && eval_off, "gci_batch_solve: null argument");
        const auto t0 = std::chrono::steady_clock::now();
        // the blocks in E contiguous groups of about equal matrix size, one
        // engine (and stream) per group; the engines and their workspaces
        // are kept for the process and reused by later batches.  E = 1 by
        // default; GCI_BATCH_ENGINES (up to 4) overlaps the groups' phases
        // (two engines: 8.6 against 9.25 ms per 64-block solve with graph
        // replay, but every kernel then shares the device, and with
        // per-launch event timing, which replays no graphs, the doubled
        // host launches make it slower)
        constexpr int MAXE = 4;
        static std::mutex mu;
        static std::unique_ptr<DavidsonEngine> eng_cache[MAXE];
        static hipStream_t streams[MAXE] = {};
        std::lock_guard<std::mutex> lk(mu);
        int E = 1;
        if (const char *v = std::getenv("GCI_BATCH_ENGINES")) E = std::atoi(v);
        E = std::max(1, std::min({E, MAXE, nblocks}));
        double total = 0.0;
        for (int b = 0; b < nblocks; ++b) {
            GCI_REQUIRE(blocks[b] && blocks[b]->n >= 2, "gci_batch_solve: block order < 2");
            total += (double)blocks[b]->nnz + (double)blocks[b]->n;
        }
        std::vector<int> gbeg(E + 1, nblocks);
        gbeg[0] = 0;
        {
            double acc = 0.0;
            int g = 1;
            for (int b = 0; b < nblocks && g < E; ++b) {
                acc += (double)blocks[b]->nnz + (double)blocks[b]->n;
                if (acc >= total * g / E && b + 1 < nblocks) gbeg[g++] = b + 1;
            }
            for (; g < E; ++g) gbeg[g] = nblocks;
        }
        std::vector<DavidsonEngine *> engs;
        for (int g = 0; g < E; ++g) {
            const int nb = gbeg[g + 1] - gbeg[g];
            if (nb <= 0) continue;
            // non-blocking: the engine captures graphs on it, and a capture
            // must not be joined by other threads' legacy-stream calls
            if (!streams[g]) GCI_HIP(hipStreamCreateWithFlags(&streams[g], hipStreamNonBlocking));
            if (!eng_cache[g] || eng_cache[g]->capacity() < nb)
                eng_cache[g] = std::make_unique<DavidsonEngine>(std::max(nb, 64 / E), streams[g]);
        }
        std::vector<DvdSpec> specs(nblocks);
        std::vector<int> niv(nblocks);
        int guard = 0;
        for (int b = 0; b < nblocks; ++b) {
            DvdSpec &sp = specs[b];
            sp.h = blocks[b];
            sp.nume = nev[b];
            sp.mblock = nev[b];
            sp.lim = (int)std::min<int64_t>(blocks[b]->n, 2 * nev[b] + 60);
            sp.maxiter = (int)std::max<int64_t>(200 * nev[b], blocks[b]->n / 10);
            sp.crite = 0.0;  // convergence on the residual alone
            sp.critc = 0.0;
            const bool split = blocks[b]->precision == GCI_SPLIT32;
            sp.mode = split ? GCI_OP_SPLIT_BULK : GCI_OP_FP64;
            sp.refine = split;
            sp.tol_rel = split ? bulk_tolerance(tol[b]) : tol[b];
            sp.tol_final = tol[b];
            sp.critr = 0.0;
            sp.eshift = eav ? eav[b] : 0.0;
            niv[b] = start_vectors(nev[b], sp.lim);
            guard = std::max(guard, 2 * sp.maxiter + 128);
        }
        std::vector<int> slots(nblocks);
        try {
            for (int g = 0; g < E; ++g) {
                const int b0 = gbeg[g], b1 = gbeg[g + 1];
                if (b1 <= b0) continue;
                std::vector<DvdSpec> sg(specs.begin() + b0, specs.begin() + b1);
                std::vector<const double *> ig(b1 - b0, nullptr);
                std::vector<int> ng(niv.begin() + b0, niv.begin() + b1);
                std::vector<int> sl = eng_cache[g]->admit(sg, ig, ng);
                std::copy(sl.begin(), sl.end(), slots.begin() + b0);
                engs.push_back(eng_cache[g].get());
            }
            run_to_completion_multi(engs, guard);
            int64_t eo = 0, vo = 0;
            for (int g = 0; g < E; ++g) {
                const int b0 = gbeg[g], b1 = gbeg[g + 1];
                if (b1 <= b0) continue;
                DavidsonEngine &eng = *eng_cache[g];
                std::vector<double> res;
                eng.extract_all(res);
                const int64_t vo0 = vo;
                for (int b = b0; b < b1; ++b) {
                    const double *r = res.data() + (size_t)slots[b] * RESULT_STRIDE;
                    eval_off[b] = eo;
                    if (evec_off) evec_off[b] = vo;
                    for (int q = 0; q < nev[b]; ++q) eval[eo + q] = r[5 + q];
                    if (stats) {
                        gci_dvd_stats st{};
                        const bool refined = r[3] != 0.0;
                        st.nloops = (int)(refined ? r[4] + r[0] : r[0]);
                        st.nmv = (int)r[1];
                        st.ierr = (int)r[2];
                        st.refine_loops = refined ? (int)r[0] : 0;
                        stats[b] = st;
                    }
                    eo += nev[b];
                    vo += blocks[b]->n * nev[b];
                }
                if (evec) eng.copy_vectors(std::vector<int>(slots.begin() + b0, slots.begin() + b1), evec + vo0);
                eng.sync();
            }
            const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (stats)
                for (int b = 0; b < nblocks; ++b) stats[b].seconds = secs;
        } catch (...) {
            for (auto &e : eng_cache) e.reset();  // state unknown after a failure
            throw;
        }
        for (DavidsonEngine *e : engs) e->release_all();
    });
}

int gci_profile_enable(int on) {
    return guarded([&] {
        GCI_HIP(hipDeviceSynchronize());  // engines run on non-blocking streams
        OpProfile &pf = op_profile();
        pf = OpProfile{};
        pf.on = on != 0;
        GCI_HIP(hipMemset(op_bytes_device(), 0, 3 * sizeof(double)));
        GCI_HIP(hipDeviceSynchronize());
    });
}

int gci_profile_read(double *stats) {
    return guarded([&] {
        GCI_REQUIRE(stats, "gci_profile_read: null argument");
        GCI_HIP(hipDeviceSynchronize());
        OpProfile &pf = op_profile();
        double b[3];
        GCI_HIP(hipMemcpy(b, op_bytes_device(), sizeof(b), hipMemcpyDeviceToHost));
        for (int m = 0; m < 3; ++m) {
            stats[3 * m + 0] = pf.launches[m];
            stats[3 * m + 1] = pf.ms[m];
            stats[3 * m + 2] = b[m];
        }
    });
}

}  // extern "C"
