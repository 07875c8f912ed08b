// Solve server (vol_server_*): a request loop that keeps meshes, their GAMG hierarchies
// and matrices resident on the GPU and runs requests concurrently.
//
// Each of `slots` worker threads owns a context (its own HIP stream and solve scratch),
// so admitted requests run side by side on the device. A request names a registered
// mesh and carries one matrix and 1–8 right-hand sides (PISO correctors) that reuse the
// matrix and the hierarchy. Requests are admitted in arrival order as soon as a worker
// is free and their device memory fits the HBM budget; a mesh copy and a matrix are
// cached per (worker, mesh) and evicted, least recently used first, when memory is
// needed. Every request runs on one worker's own buffers and stream, so its result is
// the same whether it runs alone or beside others.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "host_common.hpp"

using namespace vol;
using namespace vol::host;

namespace {

struct ProtoMesh {
    vol_mesh mesh = nullptr;  // on the server's setup context; workers use copies
    int n = 0, nf = 0;
    uint64_t bytes = 0;       // device estimate of one resident copy + matrix
};

struct CacheEntry {
    vol_mesh mesh = nullptr;  // this worker's copy
    vol_matrix A = nullptr;
    uint64_t bytes = 0;
    uint64_t lastUse = 0;
    bool busy = false;
};

struct Req {
    int64_t id = 0;
    vol_request r{};
    std::vector<const double*> b;
    std::vector<double*> x;
    int status = VOL_OK;
    std::string error;
    bool done = false;
    double tQueued = 0, tStart = 0, tDone = 0;  // seconds since the server started
};

struct Worker {
    vol_context ctx = nullptr;
    std::thread th;
    std::map<int, CacheEntry> cache;
};

uint64_t estimateBytes(int n, int nf) {
    // mesh copy (SELL rows, levels, transfer lists) + matrix (coefficients on every
    // level in FP64 and FP32, LU, work vectors): a generous per-cell/per-face bound
    return uint64_t(n) * 400 + uint64_t(nf) * 96 + (uint64_t(1) << 20);
}

}  // namespace

struct vol_server_s {
    int device = 0;
    vol_context setup = nullptr;  // registered meshes (hierarchies built once)
    uint64_t budget = 0, inUse = 0, registered = 0, clock = 0;
    std::vector<std::unique_ptr<Worker>> workers;
    std::vector<ProtoMesh> meshes;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<Req*> queue;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    int64_t nextId = 1;
    bool stop = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

// Under the lock: can request q run on worker w now? Reserves memory if so (evicting
// idle cache entries of w, least recently used first).
bool admit(vol_server S, Worker& w, Req* q) {
    const int mid = q->r.mesh;
    auto it = w.cache.find(mid);
    if (it != w.cache.end()) return !it->second.busy;
    const uint64_t need = S->meshes[mid].bytes;
    while (S->inUse + need > S->budget) {
        auto victim = w.cache.end();
        for (auto e = w.cache.begin(); e != w.cache.end(); ++e)
            if (!e->second.busy && (victim == w.cache.end() || e->second.lastUse < victim->second.lastUse)) victim = e;
        if (victim == w.cache.end()) return false;
        if (victim->second.A) vol_matrix_destroy(victim->second.A);
        if (victim->second.mesh) vol_mesh_destroy(victim->second.mesh);
        S->inUse -= victim->second.bytes;
        w.cache.erase(victim);
    }
    CacheEntry e;
    e.bytes = need;
    w.cache[mid] = e;
    S->inUse += need;
    return true;
}

void runRequest(vol_server S, Worker& w, Req* q) {
    CacheEntry* e;
    vol_mesh proto;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        e = &w.cache[q->r.mesh];
        proto = S->meshes[q->r.mesh].mesh;
    }
    int rc = VOL_OK;
    std::string err;
    auto chk = [&](int c) {
        if (c != VOL_OK && rc == VOL_OK) {
            rc = c;
            err = vol_last_error();
        }
        return c == VOL_OK;
    };
    if (!e->mesh) {
        try {
            e->mesh = meshCloneTo(proto, w.ctx);
        } catch (const std::exception& ex) {
            rc = VOL_ERR_HIP;
            err = ex.what();
        }
    }
    if (rc == VOL_OK && !e->A) chk(vol_matrix_create(e->mesh, &e->A));
    if (rc == VOL_OK && chk(vol_matrix_set(e->A, q->r.diag, q->r.upper, q->r.lower, VOL_HOST))) {
        for (int k = 0; k < q->r.nRhs && rc == VOL_OK; ++k)
            chk(vol_solve(e->A, q->b[k], q->x[k], &q->r.controls, q->r.perf ? &q->r.perf[k] : nullptr, VOL_HOST));
    }
    std::lock_guard<std::mutex> lk(S->mu);
    e->busy = false;
    e->lastUse = ++S->clock;
    q->status = rc;
    q->error = err;
    q->tDone = S->now();
    q->done = true;
    S->cvDone.notify_all();
    S->cvWork.notify_all();
}

void workerLoop(vol_server S, int wi) {
    Worker& w = *S->workers[wi];
    for (;;) {
        Req* q = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop) return;
                for (auto it = S->queue.begin(); it != S->queue.end(); ++it)
                    if (admit(S, w, *it)) {
                        q = *it;
                        S->queue.erase(it);
                        break;
                    }
                if (q) break;
                S->cvWork.wait(lk);
            }
            w.cache[q->r.mesh].busy = true;
            q->tStart = S->now();
        }
        runRequest(S, w, q);
    }
}

}  // namespace

extern "C" {

int vol_server_create(int device, int slots, uint64_t hbmBudgetBytes, vol_server* out) {
    API_BEGIN
    if (!out || slots < 1 || slots > 16 || hbmBudgetBytes == 0)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "slots must be 1..16 and the budget positive");
    auto S = std::make_unique<vol_server_s>();
    S->device = device;
    S->budget = hbmBudgetBytes;
    int rc = vol_context_create(device, &S->setup);
    if (rc != VOL_OK) throw VolError(rc, vol_last_error());
    for (int i = 0; i < slots; ++i) {
        auto w = std::make_unique<Worker>();
        rc = vol_context_create(device, &w->ctx);
        if (rc != VOL_OK) {
            for (auto& p : S->workers) vol_context_destroy(p->ctx);
            vol_context_destroy(S->setup);
            throw VolError(rc, vol_last_error());
        }
        S->workers.push_back(std::move(w));
    }
    vol_server raw = S.get();
    for (int i = 0; i < slots; ++i) raw->workers[i]->th = std::thread(workerLoop, raw, i);
    *out = S.release();
    API_END
}

int vol_server_add_mesh(vol_server S, int nCells, int nInternalFaces, const int32_t* owner, const int32_t* neighbour,
                        const double* gamgWeights, const int32_t* boundaryOwner, int nBoundaryFaces, int* meshId) {
    API_BEGIN
    if (!S || !meshId) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    ProtoMesh p;
    int rc = vol_mesh_create(S->setup, nCells, nInternalFaces, owner, neighbour, gamgWeights, 0, &p.mesh);
    if (rc != VOL_OK) throw VolError(rc, vol_last_error());
    if (boundaryOwner && nBoundaryFaces > 0) {
        rc = vol_mesh_set_boundary(p.mesh, nBoundaryFaces, boundaryOwner);
        if (rc != VOL_OK) {
            vol_mesh_destroy(p.mesh);
            throw VolError(rc, vol_last_error());
        }
    }
    p.n = nCells;
    p.nf = nInternalFaces;
    p.bytes = estimateBytes(nCells, nInternalFaces);
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->registered + p.bytes / 2 > S->budget) {
        vol_mesh_destroy(p.mesh);
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "registered meshes exceed the HBM budget");
    }
    S->registered += p.bytes / 2;  // the registered copy (addressing and hierarchy)
    S->inUse += p.bytes / 2;
    S->meshes.push_back(p);
    *meshId = int(S->meshes.size()) - 1;
    API_END
}

int vol_server_submit(vol_server S, const vol_request* r, int64_t* ticket) {
    API_BEGIN
    if (!S || !r || !ticket) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (r->nRhs < 1 || r->nRhs > 8 || !r->b || !r->x || !r->diag || !r->upper)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "request needs diag, upper and 1..8 right-hand sides");
    auto q = std::make_unique<Req>();
    q->r = *r;
    for (int k = 0; k < r->nRhs; ++k) {
        if (!r->b[k] || !r->x[k]) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null right-hand side or solution");
        q->b.push_back(r->b[k]);
        q->x.push_back(r->x[k]);
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (r->mesh < 0 || r->mesh >= int(S->meshes.size())) throw VolError(VOL_ERR_INVALID_ARGUMENT, "unknown mesh id");
    if (S->meshes[r->mesh].bytes + S->registered > S->budget)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "request can never fit the HBM budget");
    q->id = S->nextId++;
    q->tQueued = S->now();
    *ticket = q->id;
    S->queue.push_back(q.get());
    S->reqs[q->id] = std::move(q);
    S->cvWork.notify_all();
    API_END
}

int vol_server_wait(vol_server S, int64_t ticket, double timeoutSeconds, int* requestStatus, double* times) {
    API_BEGIN
    if (!S) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) throw VolError(VOL_ERR_INVALID_ARGUMENT, "unknown ticket");
    Req* q = it->second.get();
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeoutSeconds);
    while (!q->done)
        if (S->cvDone.wait_until(lk, until) == std::cv_status::timeout && !q->done)
            throw VolError(VOL_ERR_TIMEOUT, "request not finished within the timeout");
    if (requestStatus) *requestStatus = q->status;
    if (times) {
        times[0] = q->tQueued;
        times[1] = q->tStart;
        times[2] = q->tDone;
    }
    const std::string err = q->error;
    const int st = q->status;
    S->reqs.erase(it);
    if (st != VOL_OK) throw VolError(st, "request failed: " + err);
    API_END
}

int vol_server_stats(vol_server S, uint64_t* bytesInUse, int* queued, int* cachedEntries) {
    API_BEGIN
    if (!S) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (bytesInUse) *bytesInUse = S->inUse;
    if (queued) *queued = int(S->queue.size());
    if (cachedEntries) {
        int c = 0;
        for (auto& w : S->workers) c += int(w->cache.size());
        *cachedEntries = c;
    }
    API_END
}

int vol_server_destroy(vol_server S) {
    API_BEGIN
    if (!S) return VOL_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cvWork.notify_all();
    }
    for (auto& w : S->workers)
        if (w->th.joinable()) w->th.join();
    for (auto& w : S->workers)
        for (auto& [mid, e] : w->cache) {
            if (e.A) vol_matrix_destroy(e.A);
            if (e.mesh) vol_mesh_destroy(e.mesh);
        }
    for (auto& p : S->meshes) vol_mesh_destroy(p.mesh);
    for (auto& w : S->workers) vol_context_destroy(w->ctx);
    vol_context_destroy(S->setup);
    delete S;
    API_END
}

}  // extern "C"
