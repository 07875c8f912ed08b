// Solve server (vol_server_*): a request loop that keeps meshes, their GAMG hierarchies
// and matrices resident on the GPU and runs requests concurrently.
//
// Each of `slots` worker threads owns a context (its own HIP stream and solve scratch),
// so admitted requests run side by side on the device. A request names a registered
// mesh and carries one matrix and 1–8 right-hand sides (PISO correctors) that reuse the
// matrix and the hierarchy. Requests are admitted in arrival order as soon as a worker
// is free and their device memory fits the HBM budget; a mesh copy and a matrix are
// cached per (worker, mesh) and evicted, least recently used first, when memory is
// needed.
//
// Continuous batching: pressure requests (PCG + GAMG with the damped-Jacobi V-cycle) on
// meshes of at most BATCH_MAX_CELLS cells are solved as systems of a batch of
// BATCH_SYSTEMS copies of their mesh (vol_batch, the mesh's hierarchy reused): a worker
// that takes one such request also takes the requests queued behind it for the same mesh
// with the same controls, as many as the free systems hold, and their right-hand sides
// run in the same launches; systems left over get b = 0 and finish at once. A system's
// arithmetic in a batch does not depend on its batch mates or its position, so a
// request's result is the same bits alone and in the mix. Other requests (PBiCGStab
// momentum solves, pressure on larger meshes, which fill the GPU on their own and run
// faster through vol_solve's fused iteration) solve their right-hand sides one after the
// other with vol_solve.
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
    bool batched = false;     // has a hierarchy and ≤ BATCH_MAX_CELLS cells: pressure requests batch
    uint64_t bytes = 0;       // device estimate of one resident copy + matrix (+ batch)
};

constexpr int BATCH_SYSTEMS = 4;           // right-hand sides a worker solves together per mesh
constexpr int BATCH_MAX_CELLS = 1 << 18;  // larger meshes: one right-hand side at a time

struct CacheEntry {
    vol_mesh mesh = nullptr;  // this worker's copy
    vol_matrix A = nullptr;
    vol_batch batch = nullptr;  // BATCH_SYSTEMS copies of the mesh (pressure requests)
    DevBuf<double> bb, bx;      // the batch's right-hand sides and solutions
    // device bytes counted against the budget: the estimate until the first request on
    // the entry ends, then what the entry's allocations measured (grown with later ones)
    uint64_t bytes = 0;
    bool measured = false;
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
    int prefetchWorker = -1;  // ≥ 0: no solve, make the mesh resident on that worker
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

// A request that the batched path serves: PCG + GAMG, damped-Jacobi V-cycle, fast sums.
bool batchable(const Req* q, bool meshHasGamg) {
    const vol_solver_controls& c = q->r.controls;
    const std::string solver = c.solver ? c.solver : "PCG", pre = c.preconditioner ? c.preconditioner : "diagonal";
    return meshHasGamg && solver == "PCG" && pre == "GAMG" && c.gamgSmoother == VOL_GAMG_JACOBI &&
           (c.sums == VOL_SUMS_FAST || c.sums == VOL_SUMS_CONTEXT);
}

// Two requests whose right-hand sides can share one batched solve.
bool sameSolve(const Req* a, const Req* b) {
    const vol_solver_controls &x = a->r.controls, &y = b->r.controls;
    auto str = [](const char* s) { return std::string(s ? s : ""); };
    return a->r.mesh == b->r.mesh && str(x.solver) == str(y.solver) && str(x.preconditioner) == str(y.preconditioner) &&
           x.tolerance == y.tolerance && x.relTol == y.relTol && x.maxIter == y.maxIter && x.minIter == y.minIter &&
           x.nSweeps == y.nSweeps && x.sums == y.sums && x.precision == y.precision &&
           x.gamgPreSweeps == y.gamgPreSweeps && x.gamgPostSweeps == y.gamgPostSweeps && x.gamgOmega == y.gamgOmega &&
           x.innerRelTol == y.innerRelTol && x.gamgSmoother == y.gamgSmoother;
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

// Frees a cache entry's device objects (under the lock).
void evict(vol_server S, Worker& w, std::map<int, CacheEntry>::iterator victim) {
    if (victim->second.batch) vol_batch_destroy(victim->second.batch);
    if (victim->second.A) vol_matrix_destroy(victim->second.A);
    if (victim->second.mesh) vol_mesh_destroy(victim->second.mesh);
    S->inUse -= victim->second.bytes;
    w.cache.erase(victim);
}

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
        evict(S, w, victim);
    }
    w.cache[mid].bytes = need;
    S->inUse += need;
    return true;
}

// Finishes a group of requests (under the lock). grown: device bytes the entry's
// allocations added while the group ran. When the measured size pushes the total over
// the budget, idle entries of the worker go, least recently used first.
void finish(vol_server S, Worker& w, CacheEntry* e, const std::vector<Req*>& group, int rc, const std::string& err,
            int64_t grown) {
    std::lock_guard<std::mutex> lk(S->mu);
    const uint64_t was = e->bytes;
    if (!e->measured) {  // the reservation becomes the measured size
        e->bytes = uint64_t(std::max<int64_t>(grown, 0));
        e->measured = true;
    } else if (grown > 0) {
        e->bytes += uint64_t(grown);
    }
    S->inUse = S->inUse - was + e->bytes;
    while (S->inUse > S->budget) {
        auto victim = w.cache.end();
        for (auto it = w.cache.begin(); it != w.cache.end(); ++it)
            if (&it->second != e && !it->second.busy &&
                (victim == w.cache.end() || it->second.lastUse < victim->second.lastUse))
                victim = it;
        if (victim == w.cache.end()) break;
        evict(S, w, victim);
    }
    e->busy = false;
    e->lastUse = ++S->clock;
    for (Req* q : group) {
        q->status = rc;
        q->error = err;
        q->tDone = S->now();
        q->done = true;
    }
    S->cvDone.notify_all();
    S->cvWork.notify_all();
}

// The right-hand sides of the group (pressure requests on one mesh, same controls, at
// most BATCH_SYSTEMS in all) as systems of the entry's batch, in group order.
void runBatched(Worker& w, CacheEntry* e, const std::vector<Req*>& group) {
    if (!e->batch) e->batch = batchOfCopies(w.ctx, e->mesh, BATCH_SYSTEMS);
    const MeshView v = meshView(e->mesh);
    const size_t n = size_t(v.n);
    std::vector<const double*> diag, upper;
    std::vector<std::pair<Req*, int>> sys;
    for (Req* q : group)
        for (int k = 0; k < q->r.nRhs; ++k) sys.push_back({q, k});
    for (int s = 0; s < BATCH_SYSTEMS; ++s) {
        const Req* q = sys[std::min<size_t>(s, sys.size() - 1)].first;  // spare systems: any valid matrix
        diag.push_back(q->r.diag);
        upper.push_back(q->r.upper);
    }
    batchSetSystems(e->batch, diag, upper);
    hipStream_t st = w.ctx->stream;
    e->bb.alloc(n * BATCH_SYSTEMS);
    e->bx.alloc(n * BATCH_SYSTEMS);
    HIPCHECK(hipMemsetAsync(e->bb.p, 0, sizeof(double) * n * BATCH_SYSTEMS, st));
    HIPCHECK(hipMemsetAsync(e->bx.p, 0, sizeof(double) * n * BATCH_SYSTEMS, st));
    for (size_t s = 0; s < sys.size(); ++s) {
        Req* q = sys[s].first;
        const int k = sys[s].second;
        HIPCHECK(hipMemcpyAsync(e->bb.p + s * n, q->b[k], sizeof(double) * n, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(e->bx.p + s * n, q->x[k], sizeof(double) * n, hipMemcpyHostToDevice, st));
    }
    std::vector<vol_solver_performance> perf(BATCH_SYSTEMS);
    const int rc = vol_batch_solve(e->batch, e->bb.p, e->bx.p, &group[0]->r.controls, perf.data(), VOL_DEVICE);
    if (rc != VOL_OK) throw VolError(rc, vol_last_error());
    for (size_t s = 0; s < sys.size(); ++s) {
        Req* q = sys[s].first;
        const int k = sys[s].second;
        HIPCHECK(hipMemcpyAsync(q->x[k], e->bx.p + s * n, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        if (q->r.perf) q->r.perf[k] = perf[s];
    }
    HIPCHECK(hipStreamSynchronize(st));
}

// A worker's resident copy of its mesh: the copy, matrix storage and, for a mesh whose
// pressure requests batch, the batch of copies they run in.
void prefetch(Worker& w, CacheEntry* e, bool batchMesh, int n) {
    if (!e->A) {
        const int rc = vol_matrix_create(e->mesh, &e->A);
        if (rc != VOL_OK) throw VolError(rc, vol_last_error());
    }
    if (batchMesh && !e->batch) {
        e->batch = batchOfCopies(w.ctx, e->mesh, BATCH_SYSTEMS);
        e->bb.alloc(size_t(n) * BATCH_SYSTEMS);
        e->bx.alloc(size_t(n) * BATCH_SYSTEMS);
    }
}

void runRequest(vol_server S, Worker& w, const std::vector<Req*>& group, bool batched) {
    Req* q = group[0];
    const int64_t bytes0 = tlDeviceBytes;
    bool batchMesh;
    int nCells;
    CacheEntry* e;
    vol_mesh proto;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        e = &w.cache[q->r.mesh];
        proto = S->meshes[q->r.mesh].mesh;
        batchMesh = S->meshes[q->r.mesh].batched;
        nCells = S->meshes[q->r.mesh].n;
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
    try {
        if (!e->mesh) e->mesh = meshCloneTo(proto, w.ctx);
        if (q->prefetchWorker >= 0) {
            prefetch(w, e, batchMesh, nCells);
        } else if (batched) {
            runBatched(w, e, group);
        } else {
            if (!e->A) chk(vol_matrix_create(e->mesh, &e->A));
            if (rc == VOL_OK && chk(vol_matrix_set(e->A, q->r.diag, q->r.upper, q->r.lower, VOL_HOST))) {
                for (int k = 0; k < q->r.nRhs && rc == VOL_OK; ++k)
                    chk(vol_solve(e->A, q->b[k], q->x[k], &q->r.controls, q->r.perf ? &q->r.perf[k] : nullptr,
                                  VOL_HOST));
            }
        }
    } catch (const VolError& ex) {
        if (rc == VOL_OK) {
            rc = ex.code;
            err = ex.what();
        }
    } catch (const std::exception& ex) {
        if (rc == VOL_OK) {
            rc = VOL_ERR_HIP;
            err = ex.what();
        }
    }
    finish(S, w, e, group, rc, err, tlDeviceBytes - bytes0);
}

void workerLoop(vol_server S, int wi) {
    Worker& w = *S->workers[wi];
    for (;;) {
        std::vector<Req*> group;
        bool batched = false;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop) return;
                for (auto it = S->queue.begin(); it != S->queue.end(); ++it)
                    if (((*it)->prefetchWorker < 0 || (*it)->prefetchWorker == wi) && admit(S, w, *it)) {
                        group.push_back(*it);
                        S->queue.erase(it);
                        break;
                    }
                if (!group.empty()) break;
                S->cvWork.wait(lk);
            }
            Req* q = group[0];
            batched = q->prefetchWorker < 0 && batchable(q, S->meshes[q->r.mesh].batched) && q->r.nRhs <= BATCH_SYSTEMS;
            if (batched) {  // the requests queued behind it that share its solve join the batch
                int used = q->r.nRhs;
                for (auto it = S->queue.begin(); it != S->queue.end();) {
                    if ((*it)->prefetchWorker < 0 && used + (*it)->r.nRhs <= BATCH_SYSTEMS && sameSolve(q, *it)) {
                        used += (*it)->r.nRhs;
                        group.push_back(*it);
                        it = S->queue.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
            w.cache[q->r.mesh].busy = true;
            for (Req* g : group) g->tStart = S->now();
        }
        runRequest(S, w, group, batched);
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
    p.batched = gamgWeights != nullptr && nCells <= BATCH_MAX_CELLS;
    // one resident copy + matrix, and for pressure requests the batch of copies
    p.bytes = estimateBytes(nCells, nInternalFaces) * (p.batched ? 1 + BATCH_SYSTEMS : 1);
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

int vol_server_prefetch(vol_server S, int meshId, double timeoutSeconds) {
    API_BEGIN
    if (!S) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null server");
    std::vector<Req*> mine;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (meshId < 0 || meshId >= int(S->meshes.size())) throw VolError(VOL_ERR_INVALID_ARGUMENT, "unknown mesh id");
        if (S->meshes[meshId].bytes + S->registered > S->budget)
            throw VolError(VOL_ERR_INVALID_ARGUMENT, "the mesh can never fit the HBM budget");
        for (int wi = 0; wi < int(S->workers.size()); ++wi) {
            auto q = std::make_unique<Req>();
            q->r.mesh = meshId;
            q->prefetchWorker = wi;
            q->id = S->nextId++;
            q->tQueued = S->now();
            mine.push_back(q.get());
            S->queue.push_back(q.get());
            S->reqs[q->id] = std::move(q);
        }
        S->cvWork.notify_all();
    }
    std::unique_lock<std::mutex> lk(S->mu);
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeoutSeconds);
    for (Req* q : mine)
        while (!q->done)
            if (S->cvDone.wait_until(lk, until) == std::cv_status::timeout && !q->done)
                throw VolError(VOL_ERR_TIMEOUT, "prefetch not finished within the timeout");
    int st = VOL_OK;
    std::string err;
    for (Req* q : mine) {
        if (q->status != VOL_OK && st == VOL_OK) {
            st = q->status;
            err = q->error;
        }
        S->reqs.erase(q->id);
    }
    if (st != VOL_OK) throw VolError(st, "prefetch failed: " + err);
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
            if (e.batch) vol_batch_destroy(e.batch);
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
