// serve.hip -- continuous-batching CI solver (request loop).
//
// A worker thread owns one DavidsonEngine with a fixed number of slots.
// Requests (a device-resident block Hamiltonian, NEV, tolerance) queue
// FIFO; between engine steps the worker admits queued requests while a slot
// is free and the HBM budget (matrix + Davidson workspace) allows, so a
// request that arrives while others are iterating joins the running batch
// at the next step and its iterations interleave with theirs on the device.
// Finished slots are refined (FP64 operator), extracted, and released,
// which frees their budget for the next arrivals.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "dvd.h"

namespace gci {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    HMat *h = nullptr;
    bool own = true;
    int nev = 1;
    double tol = 1e-9, eav = 0.0;
    size_t bytes = 0;
    Clock::time_point t_submit, t_done;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int code = GCI_OK;
    std::string err;
    int slot = -1;
    std::vector<double> eval, evec;
};

}  // namespace gci

struct gci_server {
    size_t budget = 0, used = 0;
    int max_active = 1;
    hipStream_t stream = nullptr;
    std::unique_ptr<gci::DavidsonEngine> eng;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<gci::Request *> queue;
    std::map<int64_t, std::unique_ptr<gci::Request>> reqs;
    std::map<int, gci::Request *> running;  // slot -> request
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    std::string fatal;
    // finished requests' own matrices, freed by the reaper thread: hipFree
    // synchronises the device, which must not stall the worker's steps
    std::vector<gci::HMat *> graveyard;
    std::condition_variable cv_reap;
    bool stop_reaper = false;
    std::thread reaper;
};

namespace gci {

namespace {

DvdSpec spec_for(const Request &r) {
    DvdSpec sp;
    sp.h = r.h;
    sp.nume = r.nev;
    sp.mblock = r.nev;
    sp.lim = (int)std::min<int64_t>(r.h->n, 2 * r.nev + 60);
    sp.maxiter = (int)std::max<int64_t>(200 * r.nev, r.h->n / 10);
    sp.crite = 0.0;
    sp.critc = 0.0;
    sp.critr = 0.0;
    const bool split = r.h->precision == GCI_SPLIT32;
    sp.mode = split ? GCI_OP_SPLIT_BULK : GCI_OP_FP64;
    sp.refine = split;
    sp.tol_rel = split ? bulk_tolerance(r.tol) : r.tol;
    sp.tol_final = r.tol;
    sp.eshift = r.eav;
    return sp;
}

void finish(gci_server *s, Request *r, int code, const std::string &msg) {
    r->t_done = Clock::now();
    r->code = code;
    r->err = msg;
    r->state = code == GCI_OK ? 2 : 3;
    if (r->own && r->h) {  // called with s->mu held
        s->graveyard.push_back(r->h);
        r->h = nullptr;
        s->cv_reap.notify_all();
    }
}

void free_matrix(HMat *h) {
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete static_cast<gci_hmat *>(h);
}

void reaper_loop(gci_server *s) {
    for (;;) {
        std::vector<HMat *> dead;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_reap.wait(lk, [&] { return s->stop_reaper || !s->graveyard.empty(); });
            dead.swap(s->graveyard);
            if (dead.empty() && s->stop_reaper) break;
        }
        for (HMat *h : dead) free_matrix(h);
    }
}

void worker_loop(gci_server *s) {
    try {
        for (;;) {
            std::vector<Request *> admit;
            {
                std::unique_lock<std::mutex> lk(s->mu);
                s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->running.empty(); });
                if (s->stop && s->running.empty()) break;
                int free_slots = s->max_active - (int)s->running.size();
                while (!s->queue.empty() && free_slots > 0 && s->used + s->queue.front()->bytes <= s->budget) {
                    Request *r = s->queue.front();
                    s->queue.pop_front();
                    s->used += r->bytes;
                    r->state = 1;
                    admit.push_back(r);
                    --free_slots;
                }
            }
            if (!admit.empty()) {
                std::vector<DvdSpec> specs;
                std::vector<const double *> init;
                std::vector<int> niv;
                for (Request *r : admit) {
                    specs.push_back(spec_for(*r));
                    init.push_back(nullptr);
                    niv.push_back(start_vectors(r->nev, specs.back().lim));
                }
                std::vector<int> slots = s->eng->admit(specs, init, niv);
                std::lock_guard<std::mutex> lk(s->mu);
                for (size_t i = 0; i < admit.size(); ++i) {
                    admit[i]->slot = slots[i];
                    s->running[slots[i]] = admit[i];
                }
            }
            bool any_running;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                any_running = !s->running.empty();
            }
            if (!any_running) continue;
            // one iteration queued ahead of the state the host reads, so the
            // device does not idle while finished requests are handled
            s->eng->step(1);
            std::vector<int> done = s->eng->poll_lagged();
            std::vector<int> ref;
            for (int sl : done)
                if (s->eng->needs_refine(sl)) ref.push_back(sl);
            if (!ref.empty()) s->eng->refine_restart(ref);
            for (int sl : done) {
                if (s->eng->needs_refine(sl) || s->eng->phase(sl) != PH_DONE) continue;
                Request *r;
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    r = s->running[sl];
                }
                r->eval.resize(r->nev);
                r->evec.resize((size_t)r->h->n * r->nev);
                gci_dvd_stats st{};
                s->eng->extract(sl, r->eval.data(), r->evec.data(), nullptr, &st);
                s->eng->release(sl);
                std::lock_guard<std::mutex> lk(s->mu);
                s->running.erase(sl);
                s->used -= r->bytes;
                if (st.ierr & 2048) finish(s, r, GCI_ERR_NOCONV, "NLOOPS > MAXITER");
                else finish(s, r, GCI_OK, "");
                s->cv_done.notify_all();
            }
        }
    } catch (const Error &e) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->fatal = e.msg;
        for (auto &kv : s->reqs)
            if (kv.second->state < 2) finish(s, kv.second.get(), e.code, e.msg);
        s->queue.clear();
        s->running.clear();
        s->cv_done.notify_all();
    }
}

}  // namespace

}  // namespace gci

using namespace gci;

extern "C" {

int gci_server_create(size_t hbm_budget_bytes, int max_active, gci_server **out) {
    return guarded([&] {
        GCI_REQUIRE(out && max_active >= 1 && hbm_budget_bytes > 0, "gci_server_create: bad arguments");
        auto s = std::make_unique<gci_server>();
        s->budget = hbm_budget_bytes;
        s->max_active = max_active;
        // non-blocking: the worker captures graphs on it while the caller's
        // thread uploads matrices through the legacy stream
        GCI_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->eng = std::make_unique<DavidsonEngine>(max_active, s->stream);
        gci_server *raw = s.release();
        raw->worker = std::thread(worker_loop, raw);
        raw->reaper = std::thread(reaper_loop, raw);
        *out = raw;
    });
}

int gci_server_destroy(gci_server *s) {
    return guarded([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            // queued requests that never started are dropped
            for (Request *r : s->queue) finish(s, r, GCI_ERR_ARG, "server destroyed");
            s->queue.clear();
        }
        s->cv_work.notify_all();
        if (s->worker.joinable()) s->worker.join();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop_reaper = true;
        }
        s->cv_reap.notify_all();
        if (s->reaper.joinable()) s->reaper.join();  // frees what is left
        s->eng.reset();
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
    });
}

namespace {

int submit(gci_server *s, gci_hmat *h, int nev, double tol, double eav, int64_t *req, bool own) {
    return guarded([&] {
        GCI_REQUIRE(s && h && req, "gci_server_submit: null argument");
        GCI_REQUIRE(tol > 0.0, "gci_server_submit: tolerance must be positive");
        {
            // the engine's admission rules, checked here so that a request
            // the worker could not admit is refused to its caller alone
            Request probe;
            probe.h = h;
            probe.nev = nev;
            probe.tol = tol;
            const DvdSpec sp = spec_for(probe);
            try {
                DavidsonEngine::check_spec(sp, start_vectors(nev, sp.lim));
            } catch (const Error &e) {
                throw Error{GCI_ERR_ARG, "gci_server_submit: " + e.msg};
            }
        }
        const Clock::time_point t0 = Clock::now();
        // the SpMM layout is built here (once per matrix), on the caller's
        // thread and the matrix's stream, so the worker's engine steps never
        // wait for it
        ensure_items(*h);
        auto r = std::make_unique<Request>();
        r->h = h;
        r->own = own;
        r->nev = nev;
        r->tol = tol;
        r->eav = eav;
        // a shared matrix stays the caller's: only the workspace is charged
        r->bytes = (own ? h->device_bytes() : 0) + DavidsonEngine::workspace_bytes(spec_for(*r));
        r->t_submit = t0;
        std::lock_guard<std::mutex> lk(s->mu);
        GCI_REQUIRE(s->fatal.empty(), "gci_server_submit: server failed: " + s->fatal);
        if (r->bytes > s->budget) {
            r->own = false;  // the caller keeps the matrix on rejection
            throw Error{GCI_ERR_NOMEM, "gci_server_submit: request exceeds the HBM budget"};
        }
        r->id = s->next_id++;
        *req = r->id;
        s->queue.push_back(r.get());
        s->reqs[r->id] = std::move(r);
        s->cv_work.notify_all();
    });
}

}  // namespace

int gci_server_submit(gci_server *s, gci_hmat *h, int nev, double tol, double eav, int64_t *req) {
    return submit(s, h, nev, tol, eav, req, true);
}

int gci_server_submit_shared(gci_server *s, gci_hmat *h, int nev, double tol, double eav, int64_t *req) {
    return submit(s, h, nev, tol, eav, req, false);
}

int gci_server_poll(gci_server *s, int64_t req, double *eval, double *evec, double *latency_s) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end()) {
        set_error("gci_server_poll: unknown request");
        return GCI_ERR_ARG;
    }
    Request *r = it->second.get();
    if (r->state < 2) return GCI_ERR_BUSY;
    const int code = r->code;
    if (code == GCI_OK) {
        if (eval) std::copy(r->eval.begin(), r->eval.end(), eval);
        if (evec) std::copy(r->evec.begin(), r->evec.end(), evec);
    } else {
        set_error(r->err);
    }
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    s->reqs.erase(it);
    return code;
}

int gci_server_drain(gci_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        if (!s->fatal.empty()) return true;
        for (auto &kv : s->reqs)
            if (kv.second->state < 2) return false;
        return true;
    });
    if (!s->fatal.empty()) {
        set_error(s->fatal);
        return GCI_ERR_DEVICE;
    }
    return GCI_OK;
}

}  // extern "C"
