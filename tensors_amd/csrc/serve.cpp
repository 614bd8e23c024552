// Serving loop over the batched CG engine and the GMRES solver.
//
// Requests arrive with host arrays.  A scheduler thread admits them in arrival
// order while their device bytes fit in the budget, uploads each admitted CG
// request and adds it to one batched CG engine (tg_batch_*), and advances that
// engine between admissions; GMRES requests of up to serve_gmres_batch_rows rows
// join one batched GMRES engine advanced by a thread of its own, larger ones go
// to worker threads that run tg_gmres_solve on their own streams.  So requests
// that arrive while others run are admitted as memory frees up and their work
// overlaps on the device.
// A finished request's X is downloaded and its device memory freed at once.
//
// Nothing a request computes depends on the other requests: the batched engine
// keeps every system's arithmetic to itself, and a GMRES solve runs alone on
// its stream.  The loop is host code only; it launches no kernels itself.

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"
#include "krylov.hpp"

namespace tg {

// GMRES requests of servers created from now on: 1 = requests of up to
// g_serve_gmres_batch_rows rows join one batched GMRES engine (tg_gbatch_*)
// advanced by a dedicated thread, larger ones run on the gmres_workers threads
// with tg_gmres_solve; 0 = every request on a worker thread.  The batched engine
// has every workgroup of a system add all of that system's tile partials (so a
// system's scalars are bitwise the same in any mix), which costs tiles^2 reads per
// reduction: right for many small systems, wrong for a system of 1e6 rows (977
// tiles), which the single-system solver reduces in two levels.
static int g_serve_gmres_batch = 1;
static int g_serve_gmres_batch_rows = 131072;  // 128 tiles of the batched engine
// GMRES requests submitted from now on that run on a worker thread: 1 (default) = the
// mixed solver (tg_gmres_solve_mixed: FP32 cycles, the FP64 true residual as stop
// test), 0 = tg_gmres_solve.  The worker-thread requests are the large ones, whose
// steps stream the basis, and several of them share the HBM bandwidth at once: FP32
// halves those bytes (160-request mix: 58.8 -> 78.5 solves/s saturated).
static int g_serve_gmres_mixed = 1;

int serve_tune_set(const char* key, int32_t value, bool* known) {
    const std::string k(key);
    *known = k == "serve_gmres_batch" || k == "serve_gmres_batch_rows" || k == "serve_gmres_mixed";
    if (!*known) return TG_OK;
    if (k == "serve_gmres_mixed") {
        if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: serve_gmres_mixed must be 0 or 1");
        g_serve_gmres_mixed = value;
        return TG_OK;
    }
    if (k == "serve_gmres_batch_rows") {
        if (value < 1) return fail(TG_ERR_INVALID, "tg_tune_set: serve_gmres_batch_rows must be >= 1");
        g_serve_gmres_batch_rows = value;
        return TG_OK;
    }
    if (value != 0 && value != 1) return fail(TG_ERR_INVALID, "tg_tune_set: serve_gmres_batch must be 0 or 1");
    g_serve_gmres_batch = value;
    return TG_OK;
}

int32_t serve_tune_get(const char* key) {
    const std::string k(key);
    if (k == "serve_gmres_batch") return g_serve_gmres_batch;
    if (k == "serve_gmres_batch_rows") return g_serve_gmres_batch_rows;
    if (k == "serve_gmres_mixed") return g_serve_gmres_mixed;
    return -1;
}

// Whether a GMRES request of n rows goes to the batched engine under the current knobs.
static bool gmres_goes_batched(int32_t n) { return g_serve_gmres_batch != 0 && n <= g_serve_gmres_batch_rows; }

namespace {

enum : int { kCG = 0, kGMRES = 1 };
enum : int { kQueued = 0, kRunning = 1, kDone = 2 };

struct Request {
    int64_t id = 0;
    int kind = kCG;
    int32_t n = 0, k = 1, max_iter = 0, mixed = 0, restart = 30, max_restarts = 0;
    int64_t nnz = 0;
    const int32_t* rp = nullptr;
    const int32_t* ci = nullptr;
    const double* val = nullptr;
    const double* diag = nullptr;
    const double* B = nullptr;
    double* X = nullptr;
    double tol = 0.0;
    size_t bytes = 0;  // charged against the budget
    // device side
    char* dmem = nullptr;
    int64_t batch_id = -1;
    bool batched = false;  // GMRES: in the batched engine (else on a worker thread)
    // outcome
    int state = kQueued;
    int status = TG_OK;
    int32_t succeeded = 0, iterations = 0, restarts = 0;
    std::vector<double> rel;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

size_t input_bytes(int32_t n, int64_t nnz, int32_t k, bool has_diag) {
    return align256((size_t)(n + 1) * 4) + align256((size_t)nnz * 4) + align256((size_t)nnz * 8) +
           (has_diag ? align256((size_t)n * 8) : 0) + 2 * align256((size_t)n * k * 8);
}

}  // namespace

struct Server {
    int device = 0;
    int64_t budget = 0;
    int32_t cg_iters = 16;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    std::mutex m;
    std::condition_variable cv_sched, cv_work, cv_done;
    std::deque<Request*> pending;  // arrival order
    std::deque<Request*> gmres_ready;   // for the worker threads
    std::deque<Request*> gbatch_ready;  // for the batched GMRES thread
    std::map<int64_t, std::unique_ptr<Request>> all;
    std::vector<int64_t> finished;
    int64_t next_id = 1;
    int64_t in_use = 0, peak = 0;
    int32_t running = 0;
    bool stop = false;
    bool sched_done = false;  // the scheduler admitted its last request and exited

    // scheduler-thread state
    hipStream_t cg_stream = nullptr;
    tg_batch_t batch = nullptr;
    std::map<int64_t, Request*> by_batch_id;
    // set when a batched step failed: the engine is not used again, CG requests are refused
    std::atomic<bool> cg_failed{false};

    // batched GMRES (gmres_batched): engine, stream and the requests in it, owned by the
    // GMRES thread
    bool gmres_batched = false;
    int32_t gbatch_max_rows = 0;  // requests of up to this many rows join the engine
    bool gmres_enabled = false;
    hipStream_t gm_stream = nullptr;
    tg_gbatch_t gbatch = nullptr;
    std::map<int64_t, Request*> by_gbatch_id;

    std::thread sched;
    std::vector<std::thread> workers;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    // m held: moves admissible requests out of `pending`.
    void admit_locked(std::vector<Request*>& cg_out) {
        while (!pending.empty()) {
            Request* r = pending.front();
            if ((int64_t)r->bytes > budget) {
                pending.pop_front();
                r->status = fail(TG_ERR_WORKSPACE, "tg_server: request %lld needs %zu bytes, more than the budget",
                                 (long long)r->id, r->bytes);
                r->state = kDone;
                r->t_admit = r->t_done = now();
                finished.push_back(r->id);
                cv_done.notify_all();
                continue;
            }
            if (in_use + (int64_t)r->bytes > budget) break;
            pending.pop_front();
            in_use += (int64_t)r->bytes;
            peak = std::max(peak, in_use);
            ++running;
            r->state = kRunning;
            r->t_admit = now();
            if (r->kind == kCG) {
                cg_out.push_back(r);
            } else {
                (r->batched ? gbatch_ready : gmres_ready).push_back(r);
                cv_work.notify_all();
            }
        }
    }

    void complete(Request* r) {
        std::lock_guard<std::mutex> lk(m);
        in_use -= (int64_t)r->bytes;
        --running;
        r->state = kDone;
        r->t_done = now();
        finished.push_back(r->id);
        cv_done.notify_all();
        cv_sched.notify_all();  // memory freed: more may fit
    }

    int upload(Request* r, hipStream_t s, bool with_diag) {
        const size_t N = (size_t)r->n * r->k;
        TG_HIP(hipMallocAsync((void**)&r->dmem, input_bytes(r->n, r->nnz, r->k, with_diag), s));
        char* p = r->dmem;
        auto put = [&](size_t nb) -> char* {
            char* d = p;
            p += align256(nb);
            return d;
        };
        char* rp = put((size_t)(r->n + 1) * 4);
        char* ci = put((size_t)r->nnz * 4);
        char* val = put((size_t)r->nnz * 8);
        char* dg = with_diag ? put((size_t)r->n * 8) : nullptr;
        char* B = put(N * 8);
        TG_HIP(hipMemcpyAsync(rp, r->rp, (size_t)(r->n + 1) * 4, hipMemcpyHostToDevice, s));
        TG_HIP(hipMemcpyAsync(ci, r->ci, (size_t)r->nnz * 4, hipMemcpyHostToDevice, s));
        TG_HIP(hipMemcpyAsync(val, r->val, (size_t)r->nnz * 8, hipMemcpyHostToDevice, s));
        if (dg) TG_HIP(hipMemcpyAsync(dg, r->diag, (size_t)r->n * 8, hipMemcpyHostToDevice, s));
        TG_HIP(hipMemcpyAsync(B, r->B, N * 8, hipMemcpyHostToDevice, s));
        return TG_OK;
    }

    // Device pointers inside a request's allocation (same layout as upload()).
    struct DevPtrs {
        int32_t* rp;
        int32_t* ci;
        double* val;
        double* diag;
        double* B;
        double* X;
    };
    static DevPtrs ptrs(const Request* r, bool with_diag) {
        const size_t N = (size_t)r->n * r->k;
        char* p = r->dmem;
        DevPtrs d{};
        d.rp = (int32_t*)p;
        p += align256((size_t)(r->n + 1) * 4);
        d.ci = (int32_t*)p;
        p += align256((size_t)r->nnz * 4);
        d.val = (double*)p;
        p += align256((size_t)r->nnz * 8);
        if (with_diag) {
            d.diag = (double*)p;
            p += align256((size_t)r->n * 8);
        }
        d.B = (double*)p;
        p += align256(N * 8);
        d.X = (double*)p;
        return d;
    }

    void start_cg(Request* r) {
        if (cg_failed) {
            r->status = fail(TG_ERR_HIP, "tg_server: the batched CG engine failed earlier; CG requests are refused");
            complete(r);
            return;
        }
        const bool wd = r->diag != nullptr;
        int rc = upload(r, cg_stream, wd);
        if (rc == TG_OK) {
            DevPtrs d = ptrs(r, wd);
            rc = tg_batch_add_cg(batch, r->n, r->nnz, d.rp, d.ci, d.val, d.diag, r->k, d.B, d.X, r->tol, r->max_iter,
                                 r->mixed, &r->batch_id);
        }
        if (rc != TG_OK) {
            r->status = rc;
            if (r->dmem) (void)hipFreeAsync(r->dmem, cg_stream);
            r->dmem = nullptr;
            complete(r);
            return;
        }
        by_batch_id[r->batch_id] = r;
    }

    void finish_cg(int64_t bid) {
        auto it = by_batch_id.find(bid);
        if (it == by_batch_id.end()) return;
        Request* r = it->second;
        by_batch_id.erase(it);
        int32_t fin = 0, outer = 0;
        r->rel.assign(r->k, 0.0);
        int rc = tg_batch_result(batch, bid, &fin, &r->succeeded, &r->iterations, &outer, r->rel.data());
        r->restarts = outer;
        if (rc == TG_OK) {
            DevPtrs d = ptrs(r, r->diag != nullptr);
            rc = hipMemcpyAsync(r->X, d.X, (size_t)r->n * r->k * 8, hipMemcpyDeviceToHost, cg_stream) == hipSuccess
                     ? TG_OK
                     : fail(TG_ERR_HIP, "tg_server: download failed");
        }
        if (rc == TG_OK && hipStreamSynchronize(cg_stream) != hipSuccess) rc = fail(TG_ERR_HIP, "tg_server: sync failed");
        (void)tg_batch_release(batch, bid);
        (void)hipFreeAsync(r->dmem, cg_stream);
        r->dmem = nullptr;
        r->status = rc;
        complete(r);
    }

    void sched_loop() {
        (void)hipSetDevice(device);
        std::vector<Request*> add;
        std::vector<int64_t> done(256);
        for (;;) {
            add.clear();
            {
                std::unique_lock<std::mutex> lk(m);
                admit_locked(add);
                if (add.empty() && by_batch_id.empty()) {
                    if (stop && pending.empty()) {
                        sched_done = true;
                        cv_work.notify_all();
                        break;
                    }
                    cv_sched.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
            }
            for (Request* r : add) start_cg(r);
            if (by_batch_id.empty()) continue;
            if (tg_batch_step(batch, cg_iters) != TG_OK) {
                // the engine failed: release every CG system and its device memory, finish
                // the requests with that error, and refuse CG work from now on
                const int rc = TG_ERR_HIP;
                cg_failed = true;
                for (auto& kv : by_batch_id) {
                    Request* r = kv.second;
                    (void)tg_batch_release(batch, kv.first);
                    if (r->dmem) (void)hipFreeAsync(r->dmem, cg_stream);
                    r->dmem = nullptr;
                    r->status = rc;
                    complete(r);
                }
                by_batch_id.clear();
                (void)hipStreamSynchronize(cg_stream);
                continue;
            }
            for (;;) {
                int32_t c = 0;
                if (tg_batch_finished(batch, done.data(), (int32_t)done.size(), &c) != TG_OK || c == 0) break;
                for (int32_t i = 0; i < c; ++i) finish_cg(done[i]);
            }
        }
    }

    void gmres_loop() {
        (void)hipSetDevice(device);
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
        for (;;) {
            Request* r = nullptr;
            {
                std::unique_lock<std::mutex> lk(m);
                cv_work.wait(lk, [&] { return sched_done || !gmres_ready.empty(); });
                if (gmres_ready.empty()) break;  // sched_done: nothing more can be admitted
                r = gmres_ready.front();
                gmres_ready.pop_front();
            }
            solve_gmres(r, s);
        }
        if (s) (void)hipStreamDestroy(s);
    }

    // The batched GMRES thread: admitted GMRES requests join the engine at the next
    // restart cycle; one cycle of all of them per step.
    void gmres_batch_loop() {
        (void)hipSetDevice(device);
        std::vector<Request*> add;
        std::vector<int64_t> done(256);
        for (;;) {
            add.clear();
            {
                std::unique_lock<std::mutex> lk(m);
                if (by_gbatch_id.empty())
                    cv_work.wait(lk, [&] { return sched_done || !gbatch_ready.empty(); });
                if (gbatch_ready.empty() && by_gbatch_id.empty()) break;  // sched_done, nothing left
                while (!gbatch_ready.empty()) {
                    add.push_back(gbatch_ready.front());
                    gbatch_ready.pop_front();
                }
            }
            for (Request* r : add) start_gmres_batched(r);
            if (by_gbatch_id.empty()) continue;
            if (tg_gbatch_step(gbatch, 1) != TG_OK) {
                // the engine failed: finish every GMRES request in it with that error
                for (auto& kv : by_gbatch_id) {
                    Request* r = kv.second;
                    if (r->dmem) (void)hipFreeAsync(r->dmem, gm_stream);
                    r->dmem = nullptr;
                    r->status = TG_ERR_HIP;
                    complete(r);
                }
                by_gbatch_id.clear();
                (void)hipStreamSynchronize(gm_stream);
                continue;
            }
            for (;;) {
                int32_t c = 0;
                if (tg_gbatch_finished(gbatch, done.data(), (int32_t)done.size(), &c) != TG_OK || c == 0) break;
                for (int32_t i = 0; i < c; ++i) finish_gmres_batched(done[i]);
            }
        }
    }

    void start_gmres_batched(Request* r) {
        int rc = upload(r, gm_stream, false);
        if (rc == TG_OK) {
            DevPtrs d = ptrs(r, false);
            rc = tg_gbatch_add(gbatch, r->n, r->nnz, d.rp, d.ci, d.val, nullptr, r->k, d.B, d.X, r->tol, r->restart,
                               r->max_restarts, &r->batch_id);
        }
        if (rc != TG_OK) {
            r->status = rc;
            if (r->dmem) (void)hipFreeAsync(r->dmem, gm_stream);
            r->dmem = nullptr;
            complete(r);
            return;
        }
        by_gbatch_id[r->batch_id] = r;
    }

    void finish_gmres_batched(int64_t bid) {
        auto it = by_gbatch_id.find(bid);
        if (it == by_gbatch_id.end()) return;
        Request* r = it->second;
        by_gbatch_id.erase(it);
        int32_t fin = 0;
        r->rel.assign(r->k, 0.0);
        int rc = tg_gbatch_result(gbatch, bid, &fin, &r->succeeded, &r->iterations, &r->restarts, r->rel.data());
        if (rc == TG_OK) {
            DevPtrs d = ptrs(r, false);
            rc = hipMemcpyAsync(r->X, d.X, (size_t)r->n * r->k * 8, hipMemcpyDeviceToHost, gm_stream) == hipSuccess
                     ? TG_OK
                     : fail(TG_ERR_HIP, "tg_server: download failed");
        }
        if (rc == TG_OK && hipStreamSynchronize(gm_stream) != hipSuccess) rc = fail(TG_ERR_HIP, "tg_server: sync failed");
        (void)tg_gbatch_release(gbatch, bid);
        (void)hipFreeAsync(r->dmem, gm_stream);
        r->dmem = nullptr;
        r->status = rc;
        complete(r);
    }

    void solve_gmres(Request* r, hipStream_t s) {
        int rc = s ? TG_OK : fail(TG_ERR_HIP, "tg_server: no stream for a GMRES worker");
        void* ws = nullptr;
        const size_t wsb = r->mixed ? tg_gmres_mixed_workspace_bytes(r->n, r->nnz, r->restart, r->k, 0, 0)
                                    : tg_gmres_workspace_bytes(r->n, r->restart, r->k, TG_F64);
        if (rc == TG_OK) rc = upload(r, s, false);
        if (rc == TG_OK && hipMallocAsync(&ws, wsb, s) != hipSuccess) rc = fail(TG_ERR_HIP, "tg_server: workspace");
        r->rel.assign(r->k, 0.0);
        if (rc == TG_OK) {
            DevPtrs d = ptrs(r, false);
            if (r->mixed)  // iterations: the FP32 Arnoldi steps over all cycles
                rc = tg_gmres_solve_mixed(r->n, r->nnz, d.rp, TG_I32, d.ci, d.val, nullptr, 0, nullptr, 0, r->k,
                                          r->restart, r->max_restarts, 1.0, d.B, r->k, 0.0, d.X, r->k, r->tol, 0, ws,
                                          wsb, &r->succeeded, &r->restarts, &r->iterations, r->rel.data(), s);
            else
                rc = tg_gmres_solve(r->n, r->nnz, d.rp, TG_I32, d.ci, d.val, TG_F64, 0, nullptr, 0, r->k, r->restart,
                                    1.0, d.B, r->k, 0.0, d.X, r->k, r->tol, r->max_restarts, 0, ws, wsb,
                                    &r->succeeded, &r->iterations, &r->restarts, r->rel.data(), s);
            if (rc == TG_OK && hipMemcpyAsync(r->X, d.X, (size_t)r->n * r->k * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
                rc = fail(TG_ERR_HIP, "tg_server: download failed");
        }
        if (ws) (void)hipFreeAsync(ws, s);
        if (r->dmem) (void)hipFreeAsync(r->dmem, s);
        r->dmem = nullptr;
        if (s && hipStreamSynchronize(s) != hipSuccess && rc == TG_OK) rc = fail(TG_ERR_HIP, "tg_server: sync failed");
        r->status = rc;
        complete(r);
    }
};

}  // namespace tg

using namespace tg;

struct tg_server_s {
    Server* s;
};

extern "C" size_t tg_server_request_bytes(int32_t kind, int32_t n, int64_t nnz, int32_t nrhs, int32_t has_diag,
                                          int32_t mixed, int32_t restart) {
    if (n < 1 || nrhs < 1 || nnz < 0) return 0;
    const size_t in = input_bytes(n, nnz, nrhs, kind == kCG && has_diag);
    if (kind == kCG) return in + tg_batch_request_bytes(n, nnz, nrhs, has_diag, mixed);
    if (gmres_goes_batched(n)) return in + tg_gbatch_request_bytes(n, nrhs, restart);
    if (g_serve_gmres_mixed) return in + tg_gmres_mixed_workspace_bytes(n, nnz, restart, nrhs, 0, 0);
    return in + tg_gmres_workspace_bytes(n, restart, nrhs, TG_F64);
}

extern "C" int tg_server_create(int64_t budget_bytes, int32_t gmres_workers, int32_t cg_iterations_per_poll,
                                tg_server_t* out) {
    if (!out || budget_bytes <= 0 || gmres_workers < 0 || gmres_workers > 16 || cg_iterations_per_poll < 1)
        return fail(TG_ERR_INVALID, "tg_server_create: bad argument (budget > 0, 0..16 GMRES workers, iterations >= 1)");
    auto S = std::make_unique<Server>();
    S->budget = budget_bytes;
    S->cg_iters = cg_iterations_per_poll;
    TG_HIP(hipGetDevice(&S->device));
    TG_HIP(hipStreamCreateWithFlags(&S->cg_stream, hipStreamNonBlocking));
    int rc = tg_batch_create(&S->batch, S->cg_stream);
    if (rc != TG_OK) {
        (void)hipStreamDestroy(S->cg_stream);
        return rc;
    }
    S->gmres_enabled = gmres_workers > 0;
    S->gmres_batched = g_serve_gmres_batch != 0 && gmres_workers > 0;
    S->gbatch_max_rows = g_serve_gmres_batch_rows;
    if (S->gmres_batched) {
        rc = hipStreamCreateWithFlags(&S->gm_stream, hipStreamNonBlocking) == hipSuccess
                 ? tg_gbatch_create(&S->gbatch, S->gm_stream)
                 : fail(TG_ERR_HIP, "tg_server_create: stream creation failed");
        if (rc != TG_OK) {
            (void)tg_batch_destroy(S->batch);
            if (S->gm_stream) (void)hipStreamDestroy(S->gm_stream);
            (void)hipStreamDestroy(S->cg_stream);
            return rc;
        }
    }
    Server* raw = S.get();
    S->sched = std::thread([raw] { raw->sched_loop(); });
    if (S->gmres_batched) S->workers.emplace_back([raw] { raw->gmres_batch_loop(); });
    for (int i = 0; i < gmres_workers; ++i) S->workers.emplace_back([raw] { raw->gmres_loop(); });
    *out = new tg_server_s{S.release()};
    return TG_OK;
}

extern "C" int tg_server_destroy(tg_server_t h) {
    if (!h) return TG_OK;
    Server* S = h->s;
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->stop = true;
    }
    S->cv_sched.notify_all();
    S->cv_work.notify_all();
    if (S->sched.joinable()) S->sched.join();  // returns once every request has been admitted
    for (auto& t : S->workers)
        if (t.joinable()) t.join();
    (void)tg_batch_destroy(S->batch);
    (void)hipStreamSynchronize(S->cg_stream);
    (void)hipStreamDestroy(S->cg_stream);
    if (S->gbatch) (void)tg_gbatch_destroy(S->gbatch);
    if (S->gm_stream) {
        (void)hipStreamSynchronize(S->gm_stream);
        (void)hipStreamDestroy(S->gm_stream);
    }
    delete S;
    delete h;
    return TG_OK;
}

static int submit(tg_server_t h, std::unique_ptr<Request> r, int64_t* id_out) {
    Server* S = h->s;
    {
        std::lock_guard<std::mutex> lk(S->m);
        if (S->stop) return fail(TG_ERR_INVALID, "tg_server: server is shutting down");
        r->id = S->next_id++;
        r->t_submit = S->now();
        *id_out = r->id;
        S->pending.push_back(r.get());
        S->all[r->id] = std::move(r);
    }
    S->cv_sched.notify_all();
    return TG_OK;
}

extern "C" int tg_server_submit_cg(tg_server_t h, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                                   const double* val, const double* inv_diag, int32_t nrhs, const double* B, double* X,
                                   double tol, int32_t max_iter, int32_t mixed, int64_t* id_out) {
    if (!h || !id_out || n < 1 || nnz < 0 || nrhs < 1 || nrhs > 8 || !rp || !ci || !val || !B || !X)
        return fail(TG_ERR_INVALID, "tg_server_submit_cg: bad argument (host arrays, 1 <= nrhs <= 8)");
    auto r = std::make_unique<Request>();
    r->kind = kCG;
    r->n = n;
    r->nnz = nnz;
    r->rp = rp;
    r->ci = ci;
    r->val = val;
    r->diag = inv_diag;
    r->k = nrhs;
    r->B = B;
    r->X = X;
    r->tol = tol;
    r->max_iter = max_iter;
    r->mixed = mixed == 2 ? 2 : mixed ? 1 : 0;  // the batched engine's modes (tg_batch_add_cg)
    r->bytes = tg_server_request_bytes(kCG, n, nnz, nrhs, inv_diag != nullptr, r->mixed, 0);
    if (h->s->cg_failed)
        return fail(TG_ERR_HIP, "tg_server_submit_cg: the batched CG engine failed earlier; CG requests are refused");
    return submit(h, std::move(r), id_out);
}

extern "C" int tg_server_submit_gmres(tg_server_t h, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                                      const double* val, int32_t nrhs, const double* B, double* X, double tol,
                                      int32_t restart, int32_t max_restarts, int64_t* id_out) {
    if (!h || !id_out || n < 1 || nnz < 0 || nrhs < 1 || restart < 1 || !rp || !ci || !val || !B || !X)
        return fail(TG_ERR_INVALID, "tg_server_submit_gmres: bad argument");
    if (!h->s->gmres_enabled) return fail(TG_ERR_INVALID, "tg_server_submit_gmres: server has no GMRES workers");
    auto r = std::make_unique<Request>();
    r->kind = kGMRES;
    r->n = n;
    r->nnz = nnz;
    r->rp = rp;
    r->ci = ci;
    r->val = val;
    r->k = nrhs;
    r->B = B;
    r->X = X;
    r->tol = tol;
    r->restart = restart;
    r->max_restarts = max_restarts;
    r->batched = h->s->gmres_batched && n <= h->s->gbatch_max_rows;
    r->mixed = (!r->batched && g_serve_gmres_mixed) ? 1 : 0;
    r->bytes = r->batched ? input_bytes(n, nnz, nrhs, false) + tg_gbatch_request_bytes(n, nrhs, restart)
               : r->mixed ? input_bytes(n, nnz, nrhs, false) + tg_gmres_mixed_workspace_bytes(n, nnz, restart, nrhs, 0, 0)
                          : input_bytes(n, nnz, nrhs, false) + tg_gmres_workspace_bytes(n, restart, nrhs, TG_F64);
    return submit(h, std::move(r), id_out);
}

extern "C" int tg_server_wait(tg_server_t h, int64_t* ids, int32_t cap, int32_t* count, double timeout_s) {
    if (!h || !count || (cap > 0 && !ids)) return fail(TG_ERR_INVALID, "tg_server_wait: bad argument");
    Server* S = h->s;
    std::unique_lock<std::mutex> lk(S->m);
    if (S->finished.empty() && timeout_s > 0)
        S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return !S->finished.empty(); });
    const int32_t c = (int32_t)std::min<size_t>(S->finished.size(), (size_t)std::max(cap, 0));
    for (int32_t i = 0; i < c; ++i) ids[i] = S->finished[i];
    S->finished.erase(S->finished.begin(), S->finished.begin() + c);
    *count = c;
    return TG_OK;
}

extern "C" int tg_server_result(tg_server_t h, int64_t id, int32_t* status, int32_t* succeeded, int32_t* iterations,
                                int32_t* restarts, double* rel_residuals, double* times) {
    if (!h) return fail(TG_ERR_INVALID, "tg_server_result: null handle");
    Server* S = h->s;
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(TG_ERR_INVALID, "tg_server_result: unknown id");
    const Request& r = *it->second;
    if (r.state != kDone) return fail(TG_ERR_INVALID, "tg_server_result: request %lld not finished", (long long)id);
    if (status) *status = r.status;
    if (succeeded) *succeeded = r.succeeded;
    if (iterations) *iterations = r.iterations;
    if (restarts) *restarts = r.restarts;
    if (rel_residuals)
        for (size_t c = 0; c < r.rel.size(); ++c) rel_residuals[c] = r.rel[c];
    if (times) {
        times[0] = r.t_submit;
        times[1] = r.t_admit;
        times[2] = r.t_done;
    }
    return TG_OK;
}

extern "C" int tg_server_release(tg_server_t h, int64_t id) {
    if (!h) return fail(TG_ERR_INVALID, "tg_server_release: null handle");
    Server* S = h->s;
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->all.find(id);
    if (it == S->all.end()) return fail(TG_ERR_INVALID, "tg_server_release: unknown id");
    if (it->second->state != kDone) return fail(TG_ERR_INVALID, "tg_server_release: request still running");
    S->all.erase(it);
    return TG_OK;
}

extern "C" int tg_server_stats(tg_server_t h, int64_t* bytes_in_use, int64_t* peak_bytes, int32_t* queued,
                               int32_t* running) {
    if (!h) return fail(TG_ERR_INVALID, "tg_server_stats: null handle");
    Server* S = h->s;
    std::lock_guard<std::mutex> lk(S->m);
    if (bytes_in_use) *bytes_in_use = S->in_use;
    if (peak_bytes) *peak_bytes = S->peak;
    if (queued) *queued = (int32_t)S->pending.size();
    if (running) *running = S->running;
    return TG_OK;
}
