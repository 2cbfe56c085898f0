// Serving engine: continuous batching of sound-soft scattering solves on one GPU.
//
// A request is one scattering problem (P0 geometry, quadrature option, k, eta, nrhs plane-wave
// directions, far-field angles).  The engine keeps a FIFO queue and admits the head request when
// its device footprint fits into the free part of one preallocated HBM arena (and fewer than
// max_active requests are in flight); all of its device data -- mesh, A64, right-hand sides,
// solution, FP32 factors, refinement workspace, far field -- is carved from that one arena block,
// so admission never calls hipMalloc and retirement never calls hipFree (which would synchronise
// the device).  Every engine step advances every admitted request by one stage, with one batched
// launch per kernel over all requests in that stage:
//
//   setup      mesh tables and right-hand sides                    (setup stream)
//   assemble   A = M + 2K for all newly admitted requests          (assembly stream, batched)
//   factor     FP32 LU: cast + diagonal block 0, then one panel step per engine step for every
//              request in the stage, each at its own step           (LU stream, batched)
//              -- or, on the FP64 path, rocSOLVER zgetrf + zgetrs per request
//   refine     FP64 refinement steps of all factorised requests     (refinement stream, batched)
//   finish     far field and the copies of x and u_inf to pinned host memory
//
// so requests that arrive while others are on the GPU join the next batch of their stage, and the
// stages of different requests overlap on the device (four streams, default priority).  Progress is
// observed through events; the host never waits on a stream inside step() except for the setup
// stream after a mesh upload.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/bem2d_hip.h"
#include "internal.h"

namespace {

constexpr size_t kAlign = 256;
constexpr int kAdmitPerStep = 8; // requests admitted (meshes built) per engine step at most
size_t round_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation, with coalescing of free neighbours.
class Arena : public DevArena {
public:
    Arena(char* base, size_t size) : base_(base), size_(size) { free_[0] = size; }
    void* alloc(size_t bytes) override
    {
        bytes = round_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_[off] = bytes;
            in_use_ += bytes;
            return base_ + off;
        }
        return nullptr;
    }
    void release(void* p) override
    {
        if (!p) return;
        const size_t off = (size_t)((char*)p - base_);
        auto u = used_.find(off);
        if (u == used_.end()) return;
        size_t len = u->second;
        used_.erase(u);
        in_use_ -= len;
        size_t start = off;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first == off + len) {
            len += next->second;
            free_.erase(next);
        }
        auto prev = free_.lower_bound(off);
        if (prev != free_.begin()) {
            --prev;
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
    }
    size_t capacity() const { return size_; }
    size_t in_use() const { return in_use_; }

private:
    char* base_;
    size_t size_;
    size_t in_use_ = 0;
    std::map<size_t, size_t> free_; // offset -> length
    std::unordered_map<size_t, size_t> used_;
};

// Carves consecutive pieces out of one block reserved for a request; release is a no-op (the block
// returns to the arena as a whole when the request retires).
class BumpArena : public DevArena {
public:
    void reset(char* base, size_t size)
    {
        base_ = base;
        size_ = size;
        off_ = 0;
    }
    void* alloc(size_t bytes) override
    {
        bytes = round_up(bytes);
        if (off_ + bytes > size_) return nullptr;
        void* p = base_ + off_;
        off_ += bytes;
        return p;
    }
    void release(void*) override {}

private:
    char* base_ = nullptr;
    size_t size_ = 0, off_ = 0;
};

enum Stage { QUEUED, ASSEMBLING, FACTORING, FACTORED, REFINING, FINISHING, DONE };

struct Request {
    int64_t ticket = 0;
    // copies of the request's host data
    std::vector<int32_t> type;
    std::vector<double> param;
    std::vector<int64_t> index, next, prev;
    std::vector<double> dirs, theta;
    bem2d_quad_option opts{};
    double k = 0, eta = 0, tol = 1e-12;
    int nrhs = 0, precision = BEM2D_PRECISION_MIXED;
    int64_t n = 0, nang = 0;
    // timing (engine clock, seconds)
    double t_arrival = 0, t_admit = 0, t_done = 0;
    Stage stage = QUEUED;
    int status = BEM2D_OK;
    // device state
    size_t footprint = 0;
    char* block = nullptr;
    BumpArena bump;
    bem2d_mesh* mesh = nullptr;
    double2* A = nullptr;
    double2* B = nullptr;
    double2* X = nullptr;
    double* d_dirs = nullptr;
    double* d_theta = nullptr;
    double2* ff = nullptr;
    double2* sig = nullptr;
    bem2d_lu32* lu = nullptr;
    int* piv = nullptr;     // pivots + info of the FP64 LU (FP64 requests and the fallback)
    bool fp64_ran = false;  // the FP64 LU solved this request (its info decides the status)
    hipEvent_t ev = nullptr;
    // results, in one pinned block of the engine's pool
    char* hblock = nullptr;
    size_t hblock_bytes = 0;
    double2* hx = nullptr;
    double2* hff = nullptr;
    int* hinfo = nullptr;   // copy of the FP64 LU's info
    bem2d_ir_report rep{};
};

} // namespace

struct bem2d_engine {
    bem2d_ctx* ctx = nullptr;
    char* pool = nullptr;
    std::unique_ptr<Arena> arena;
    hipStream_t s_setup = nullptr, s_asm = nullptr, s_lu = nullptr, s_ref = nullptr;
    hipEvent_t ev_setup = nullptr;
    rocblas_handle blas = nullptr;
    void* blas_ws = nullptr;
    size_t blas_ws_bytes = 0;
    int max_active = 64;
    int64_t next_ticket = 1;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    std::unordered_map<int64_t, Request*> all;
    std::deque<hipEvent_t> lu_inflight;   // one event per launched LU step (throttle)
    std::vector<hipEvent_t> event_pool;
    RefJob* job = nullptr;
    std::vector<Request*> job_reqs;
    // Pinned host memory, allocated once and reused: hipHostMalloc costs tens to hundreds of
    // microseconds a call, which would stall every admission.  Result blocks are kept by size class
    // (powers of two) for later requests; staging slots carry the uploads of a step's new meshes
    // and are reused once the setup stream has passed the event recorded after their copies.
    std::multimap<size_t, char*> pinned_free;
    struct StageSlot {
        char* ptr = nullptr;
        size_t bytes = 0;
        hipEvent_t ev = nullptr;
        bool busy = false;                // handed out in the current step, event not yet recorded
    };
    std::vector<StageSlot> stage_slots;
    unsigned long long* pinned = nullptr; // refinement read-back buffer, kept for the engine's life
    unsigned long long* dnorms = nullptr; // its device side (outside the arena, grown with it)
    size_t pinned_units = 0;
    bool launched = false;                // the last step launched work
    std::chrono::steady_clock::time_point t0;
    bem2d_engine_stats stats{};
};

namespace {

double engine_now(const bem2d_engine* e)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - e->t0).count();
}

hipEvent_t take_event(bem2d_engine* e)
{
    if (!e->event_pool.empty()) {
        hipEvent_t v = e->event_pool.back();
        e->event_pool.pop_back();
        return v;
    }
    hipEvent_t v = nullptr;
    (void)hipEventCreateWithFlags(&v, hipEventDisableTiming);
    return v;
}

void give_event(bem2d_engine* e, hipEvent_t v)
{
    if (v) e->event_pool.push_back(v);
}

bool event_done(hipEvent_t v) { return !v || hipEventQuery(v) == hipSuccess; }

// A pinned block of at least `bytes` (size class: the next power of two from 4 KiB), or null.
char* take_pinned(bem2d_engine* e, size_t bytes, size_t* cls)
{
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    *cls = c;
    auto it = e->pinned_free.find(c);
    if (it != e->pinned_free.end()) {
        char* p = it->second;
        e->pinned_free.erase(it);
        return p;
    }
    char* p = nullptr;
    return hipHostMalloc((void**)&p, c) == hipSuccess ? p : nullptr;
}

void give_pinned(bem2d_engine* e, char* p, size_t cls)
{
    if (p) e->pinned_free.emplace(cls, p);
}

// Staging of the step's mesh uploads: a free slot (never handed out, or its copies done) of at
// least `bytes`, else a new one.
struct EngineStager : HostStager {
    bem2d_engine* e;
    explicit EngineStager(bem2d_engine* eng) : e(eng) {}
    void* get(size_t bytes) override
    {
        for (auto& sl : e->stage_slots)
            if (!sl.busy && sl.bytes >= bytes && event_done(sl.ev)) {
                sl.busy = true;
                return sl.ptr;
            }
        bem2d_engine::StageSlot sl;
        sl.bytes = std::max<size_t>(bytes, (size_t)1 << 20);
        if (hipHostMalloc((void**)&sl.ptr, sl.bytes) != hipSuccess) return nullptr;
        if (hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipHostFree(sl.ptr);
            return nullptr;
        }
        sl.busy = true;
        e->stage_slots.push_back(sl);
        return sl.ptr;
    }
    // the copies out of every slot handed out so far are enqueued on s: arm their events
    int arm(hipStream_t s)
    {
        int rc = BEM2D_OK;
        for (auto& sl : e->stage_slots)
            if (sl.busy) {
                if (hipEventRecord(sl.ev, s) != hipSuccess) rc = BEM2D_ERR_DEVICE;
                sl.busy = false;
            }
        return rc;
    }
};

} // namespace

extern "C" int64_t bem2d_engine_footprint(int64_t n, int nrhs, int64_t nang, int precision,
                                          bem2d_quad_option opts)
{
    if (n < 1 || nrhs < 1 || nang < 0) return -1;
    size_t b = bem2d_internal_mesh_bytes(n, opts);
    b += round_up(sizeof(double2) * n * n);                   // A64
    b += 2 * round_up(sizeof(double2) * n * nrhs);            // B, X
    b += round_up(sizeof(double) * 2 * nrhs) + round_up(sizeof(double) * (nang + 1));
    b += round_up(sizeof(double2) * (nang + 1) * nrhs);       // far field
    b += round_up(sizeof(double2) * (nang + 1) * n);          // far-field signatures
    b += round_up(sizeof(int) * (n + 1)); // FP64 LU pivots (also the mixed path's fallback)
    if (precision == BEM2D_PRECISION_MIXED) {
        b += bem2d_internal_lu32_bytes(n) + 8 * kAlign;
        b += bem2d_internal_ref_bytes(n, nrhs);
    }
    return (int64_t)round_up(b + 16 * kAlign); // the arena hands out whole kAlign units
}

extern "C" int bem2d_engine_create(bem2d_ctx* c, int64_t hbm_bytes, int max_active,
                                   bem2d_engine** out)
{
    if (!c || !out || hbm_bytes < (1 << 20) || max_active < 1)
        return FAIL(BEM2D_ERR_ARG, "bem2d_engine_create: bad argument");
    HIPCHK(hipSetDevice(c->device));
    bem2d_engine* e = new bem2d_engine();
    e->ctx = c;
    e->max_active = max_active;
    e->t0 = std::chrono::steady_clock::now();
    auto fail = [&](int code, const char* what) {
        bem2d_engine_destroy(e);
        return FAIL(code, "bem2d_engine_create: %s", what);
    };
    if (hipMalloc((void**)&e->pool, (size_t)hbm_bytes) != hipSuccess)
        return fail(BEM2D_ERR_NOMEM, "arena allocation failed");
    e->arena.reset(new Arena(e->pool, (size_t)hbm_bytes));
    if (hipStreamCreateWithFlags(&e->s_setup, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->s_asm, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->s_lu, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->s_ref, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_setup, hipEventDisableTiming) != hipSuccess)
        return fail(BEM2D_ERR_DEVICE, "stream creation failed");
    if (rocblas_create_handle(&e->blas) != rocblas_status_success ||
        rocblas_set_stream(e->blas, e->s_lu) != rocblas_status_success)
        return fail(BEM2D_ERR_DEVICE, "rocBLAS handle creation failed");
    *out = e;
    return BEM2D_OK;
}

static void release_request_device(bem2d_engine* e, Request* r)
{
    if (r->lu) {
        bem2d_lu32_destroy(r->lu); // bump-arena memory: nothing is freed here
        r->lu = nullptr;
    }
    if (r->mesh) {
        bem2d_mesh_destroy(r->mesh);
        r->mesh = nullptr;
    }
    if (r->block) {
        e->arena->release(r->block);
        r->block = nullptr;
    }
    give_event(e, r->ev);
    r->ev = nullptr;
}

static void free_request(bem2d_engine* e, Request* r)
{
    release_request_device(e, r);
    give_pinned(e, r->hblock, r->hblock_bytes);
    delete r;
}

extern "C" void bem2d_engine_destroy(bem2d_engine* e)
{
    if (!e) return;
    (void)hipSetDevice(e->ctx->device);
    for (hipStream_t s : {e->s_setup, e->s_asm, e->s_lu, e->s_ref})
        if (s) (void)hipStreamSynchronize(s);
    bem2d_internal_ref_destroy(e->job);
    for (auto& kv : e->all) free_request(e, kv.second);
    for (hipEvent_t v : e->lu_inflight) (void)hipEventDestroy(v);
    for (hipEvent_t v : e->event_pool) (void)hipEventDestroy(v);
    if (e->ev_setup) (void)hipEventDestroy(e->ev_setup);
    if (e->blas) (void)rocblas_destroy_handle(e->blas);
    if (e->blas_ws) (void)hipFree(e->blas_ws);
    if (e->pinned) (void)hipHostFree(e->pinned);
    if (e->dnorms) (void)hipFree(e->dnorms);
    for (auto& kv : e->pinned_free) (void)hipHostFree(kv.second);
    for (auto& sl : e->stage_slots) {
        (void)hipHostFree(sl.ptr);
        (void)hipEventDestroy(sl.ev);
    }
    for (hipStream_t s : {e->s_setup, e->s_asm, e->s_lu, e->s_ref})
        if (s) (void)hipStreamDestroy(s);
    e->arena.reset();
    if (e->pool) (void)hipFree(e->pool);
    delete e;
}

extern "C" double bem2d_engine_now(const bem2d_engine* e) { return e ? engine_now(e) : 0.0; }

extern "C" int bem2d_engine_submit(bem2d_engine* e, const bem2d_request* q, double t_arrival,
                                   int64_t* ticket)
{
    if (!e || !q || !ticket) return FAIL(BEM2D_ERR_ARG, "bem2d_engine_submit: null argument");
    const bem2d_geometry& g = q->geom;
    if (g.n < 1 || !g.type || !g.param || !g.index || !g.next || !g.prev)
        return FAIL(BEM2D_ERR_ARG, "bem2d_engine_submit: bad geometry");
    if (q->nrhs < 1 || q->nrhs > 65535 || !q->dirs || q->nang < 0 || q->nang >= (1 << 24) ||
        (q->nang > 0 && !q->theta) ||
        !(q->k > 0) || (q->precision != BEM2D_PRECISION_MIXED && q->precision != BEM2D_PRECISION_FP64))
        return FAIL(BEM2D_ERR_ARG, "bem2d_engine_submit: bad request");
    if (q->opts.N < 1 || q->opts.N > kMaxNq || q->opts.L < 0 || q->opts.L > 12 ||
        !(q->opts.sigma > 0 && q->opts.sigma < 1))
        return FAIL(BEM2D_ERR_ARG, "bem2d_engine_submit: bad quadrature option");
    Request* r = new Request();
    const int64_t n = g.n;
    r->ticket = e->next_ticket++;
    r->type.assign(g.type, g.type + n);
    r->param.assign(g.param, g.param + 8 * n);
    r->index.assign(g.index, g.index + n);
    r->next.assign(g.next, g.next + n);
    r->prev.assign(g.prev, g.prev + n);
    r->dirs.assign(q->dirs, q->dirs + 2 * q->nrhs);
    if (q->nang > 0) r->theta.assign(q->theta, q->theta + q->nang);
    r->opts = q->opts;
    r->k = q->k;
    r->eta = q->eta;
    r->tol = q->tol > 0 ? q->tol : 1e-12;
    r->nrhs = q->nrhs;
    r->precision = q->precision;
    r->n = n;
    r->nang = q->nang;
    r->t_arrival = t_arrival >= 0 ? t_arrival : engine_now(e);
    r->footprint = (size_t)bem2d_engine_footprint(n, r->nrhs, r->nang, r->precision, r->opts);
    e->queue.push_back(r);
    e->all[r->ticket] = r;
    e->stats.submitted++;
    *ticket = r->ticket;
    return BEM2D_OK;
}

// Reserves the request's block and builds its mesh and right-hand sides on the setup stream; the
// uploads leave from the stager's pinned slots, so nothing here waits for the device.
static int admit(bem2d_engine* e, Request* r, EngineStager& stager)
{
    r->block = (char*)e->arena->alloc(r->footprint);
    if (!r->block) return BEM2D_ERR_NOMEM;
    r->bump.reset(r->block, round_up(r->footprint));
    const int64_t n = r->n;
    BumpArena& b = r->bump;
    r->A = (double2*)b.alloc(sizeof(double2) * n * n);
    r->B = (double2*)b.alloc(sizeof(double2) * n * r->nrhs);
    r->X = (double2*)b.alloc(sizeof(double2) * n * r->nrhs);
    r->d_dirs = (double*)b.alloc(sizeof(double) * (2 * r->nrhs + r->nang + 1));
    r->d_theta = r->d_dirs + 2 * r->nrhs;
    r->ff = (double2*)b.alloc(sizeof(double2) * (r->nang + 1) * r->nrhs);
    r->sig = (double2*)b.alloc(sizeof(double2) * (r->nang + 1) * n);
    r->piv = (int*)b.alloc(sizeof(int) * (n + 1));
    if (!r->A || !r->B || !r->X || !r->d_dirs || !r->ff || !r->sig || !r->piv)
        return FAIL(BEM2D_ERR_NOMEM, "request footprint too small");
    const size_t bx = round_up(sizeof(double2) * n * r->nrhs);
    const size_t bff = round_up(sizeof(double2) * (r->nang + 1) * r->nrhs);
    r->hblock = take_pinned(e, bx + bff + sizeof(int), &r->hblock_bytes);
    if (!r->hblock) return FAIL(BEM2D_ERR_NOMEM, "pinned result buffers");
    r->hx = (double2*)r->hblock;
    r->hff = (double2*)(r->hblock + bx);
    r->hinfo = (int*)(r->hblock + bx + bff);
    *r->hinfo = 0;
    const size_t bdir = sizeof(double) * (2 * r->nrhs + r->nang);
    double* hd = (double*)stager.get(bdir);
    if (!hd) return FAIL(BEM2D_ERR_NOMEM, "pinned staging");
    std::memcpy(hd, r->dirs.data(), sizeof(double) * 2 * r->nrhs);
    if (r->nang > 0) std::memcpy(hd + 2 * r->nrhs, r->theta.data(), sizeof(double) * r->nang);
    HIPCHK(hipMemcpyAsync(r->d_dirs, hd, bdir, hipMemcpyHostToDevice, e->s_setup));
    bem2d_geometry g{n, r->type.data(), r->param.data(), r->index.data(), r->next.data(),
                     r->prev.data()};
    int rc = bem2d_internal_mesh_create(e->ctx, &g, r->opts, &r->bump, e->s_setup, &stager,
                                        &r->mesh);
    if (rc) return rc;
    rc = bem2d_internal_rhs(r->mesh, BEM2D_FIELD_COMBINED_PLANE, r->k, r->eta, r->d_dirs, r->nrhs,
                            2.0, r->B, n, e->s_setup);
    if (rc) return rc;
    if (r->precision == BEM2D_PRECISION_MIXED) {
        rc = bem2d_internal_lu32_create(e->ctx, n, &r->bump, &r->lu);
        if (rc) return rc;
    }
    r->ev = take_event(e);
    return BEM2D_OK;
}

static void fail_request(bem2d_engine* e, Request* r, int code)
{
    r->status = code;
    r->stage = DONE;
    r->t_done = engine_now(e);
    release_request_device(e, r);
    e->stats.failed++;
}

// far field + copies to pinned host memory on stream s; the request then waits for its event
static void finish(bem2d_engine* e, Request* r, hipStream_t s)
{
    const int64_t n = r->n;
    if (r->nang > 0) {
        const int rc = bem2d_internal_farfield(r->mesh, r->k, r->X, n, r->nrhs, r->d_theta,
                                               r->nang, r->ff, r->sig, s);
        if (rc && r->status == BEM2D_OK) r->status = rc;
    }
    bool ok = hipMemcpyAsync(r->hx, r->X, sizeof(double2) * n * r->nrhs, hipMemcpyDeviceToHost,
                             s) == hipSuccess;
    if (r->nang > 0)
        ok &= hipMemcpyAsync(r->hff, r->ff, sizeof(double2) * r->nang * r->nrhs,
                             hipMemcpyDeviceToHost, s) == hipSuccess;
    if (r->fp64_ran)
        ok &= hipMemcpyAsync(r->hinfo, r->piv + n, sizeof(int), hipMemcpyDeviceToHost, s) ==
              hipSuccess;
    ok &= hipEventRecord(r->ev, s) == hipSuccess;
    if (!ok && r->status == BEM2D_OK) r->status = BEM2D_ERR_DEVICE;
    r->stage = FINISHING;
}

// rocSOLVER workspace for an n x n zgetrf, sized once per new largest n (outside the arena).
static int ensure_blas_workspace(bem2d_engine* e, int64_t n, int nrhs)
{
    size_t need = 0;
    rocblas_start_device_memory_size_query(e->blas);
    rocsolver_zgetrf(e->blas, (rocblas_int)n, (rocblas_int)n, nullptr, (rocblas_int)n, nullptr,
                     nullptr);
    rocsolver_zgetrs(e->blas, rocblas_operation_none, (rocblas_int)n, (rocblas_int)nrhs, nullptr,
                     (rocblas_int)n, nullptr, nullptr, (rocblas_int)n);
    rocblas_stop_device_memory_size_query(e->blas, &need);
    if (need <= e->blas_ws_bytes) return BEM2D_OK;
    for (hipStream_t s : {e->s_lu}) (void)hipStreamSynchronize(s);
    if (e->blas_ws) (void)hipFree(e->blas_ws);
    e->blas_ws = nullptr;
    e->blas_ws_bytes = 0;
    if (hipMalloc(&e->blas_ws, need) != hipSuccess) return BEM2D_ERR_NOMEM;
    e->blas_ws_bytes = need;
    return rocblas_set_workspace(e->blas, e->blas_ws, need) == rocblas_status_success
               ? BEM2D_OK
               : BEM2D_ERR_DEVICE;
}

// The plain FP64 solve of a request on the LU stream (rocSOLVER zgetrf + zgetrs, the reference's
// SolveSystem), then finish.  A is factorised in place.  Serves FP64 requests and the fallback of
// mixed requests whose FP32 factors could not reach the tolerance.
static void solve_fp64(bem2d_engine* e, Request* r)
{
    int rc = ensure_blas_workspace(e, r->n, r->nrhs);
    if (rc) {
        fail_request(e, r, rc);
        return;
    }
    (void)hipMemcpyAsync(r->X, r->B, sizeof(double2) * r->n * r->nrhs, hipMemcpyDeviceToDevice,
                         e->s_lu);
    rocblas_status st = rocsolver_zgetrf(e->blas, (rocblas_int)r->n, (rocblas_int)r->n,
                                         (rocblas_double_complex*)r->A, (rocblas_int)r->n, r->piv,
                                         r->piv + r->n);
    if (st == rocblas_status_success)
        st = rocsolver_zgetrs(e->blas, rocblas_operation_none, (rocblas_int)r->n,
                              (rocblas_int)r->nrhs, (rocblas_double_complex*)r->A,
                              (rocblas_int)r->n, r->piv, (rocblas_double_complex*)r->X,
                              (rocblas_int)r->n);
    if (st != rocblas_status_success) {
        fail_request(e, r, BEM2D_ERR_DEVICE);
        return;
    }
    r->fp64_ran = true;
    finish(e, r, e->s_lu);
}

extern "C" int bem2d_engine_step(bem2d_engine* e, int* unfinished)
{
    if (!e) return FAIL(BEM2D_ERR_ARG, "bem2d_engine_step: null engine");
    HIPCHK(hipSetDevice(e->ctx->device));
    const double now = engine_now(e);
    e->launched = false;
    // ---- retire requests whose results have reached the host ----
    for (Request* r : e->active)
        if (r->stage == FINISHING && event_done(r->ev)) {
            r->stage = DONE;
            r->t_done = now;
            if (r->fp64_ran && *r->hinfo > 0 && r->status == BEM2D_OK)
                r->status = BEM2D_ERR_SINGULAR; // zgetrf info > 0: LapackError
            release_request_device(e, r);
            if (r->status == BEM2D_OK)
                e->stats.completed++;
            else
                e->stats.failed++;
        }
    e->active.erase(std::remove_if(e->active.begin(), e->active.end(),
                                   [](Request* r) { return r->stage == DONE; }),
                    e->active.end());
    // ---- admission (FIFO, by footprint), at most kAdmitPerStep per step: a mesh is built on the
    // host (rules, singular-pair list, upload), so admitting a long queue at once would hold the
    // first launches back until every mesh is ready; in slices the device starts on the first
    // requests while the host prepares the next ----
    std::vector<Request*> fresh;
    EngineStager stager(e);
    while (!e->queue.empty() && (int)e->active.size() < e->max_active &&
           (int)fresh.size() < kAdmitPerStep) {
        Request* r = e->queue.front();
        if (r->footprint > e->arena->capacity()) {
            e->queue.pop_front();
            fail_request(e, r, BEM2D_ERR_NOMEM);
            continue;
        }
        const int rc = admit(e, r, stager);
        if (rc == BEM2D_ERR_NOMEM && r->block == nullptr) break; // wait for memory
        e->queue.pop_front();
        if (rc) {
            fail_request(e, r, rc);
            continue;
        }
        r->t_admit = now;
        r->stage = ASSEMBLING;
        e->active.push_back(r);
        fresh.push_back(r);
        e->stats.admitted++;
    }
    {
        const int rc = stager.arm(e->s_setup);
        if (rc) return FAIL(rc, "staging event");
    }
    // ---- batched assembly of the new requests ----
    if (!fresh.empty()) {
        HIPCHK(hipEventRecord(e->ev_setup, e->s_setup));
        HIPCHK(hipStreamWaitEvent(e->s_asm, e->ev_setup, 0));
        std::vector<const bem2d_mesh*> meshes;
        std::vector<double> ks, etas;
        std::vector<void*> As;
        std::vector<int64_t> ldas;
        for (Request* r : fresh) {
            meshes.push_back(r->mesh);
            ks.push_back(r->k);
            etas.push_back(r->eta);
            As.push_back(r->A);
            ldas.push_back(r->n);
        }
        const int rc = bem2d_internal_assemble_many(e->ctx, e->s_asm, BEM2D_KERNEL_COMBINED,
                                                    meshes.data(), ks.data(), etas.data(), 1.0,
                                                    2.0, As.data(), ldas.data(), (int)fresh.size());
        if (rc) return rc;
        for (Request* r : fresh) HIPCHK(hipEventRecord(r->ev, e->s_asm));
        e->stats.assembly_launches++;
        e->launched = true;
    }
    // ---- assembled -> factorisation (FP32 LU begins; FP64 path: zgetrf + zgetrs + finish) ----
    std::vector<bem2d_lu32*> begin_lu;
    std::vector<const void*> begin_A;
    std::vector<int64_t> begin_ld;
    for (Request* r : e->active) {
        if (r->stage != ASSEMBLING || !event_done(r->ev)) continue;
        if (r->precision == BEM2D_PRECISION_MIXED) {
            begin_lu.push_back(r->lu);
            begin_A.push_back(r->A);
            begin_ld.push_back(r->n);
            r->stage = FACTORING;
        } else {
            solve_fp64(e, r);
        }
    }
    if (!begin_lu.empty()) {
        const int rc = bem2d_internal_lu_begin(e->s_lu, begin_lu.data(), begin_A.data(),
                                               begin_ld.data(), (int)begin_lu.size());
        if (rc) return rc;
        e->launched = true;
    }
    // ---- one panel step for every request in the factorisation stage (throttled to two steps
    //      in flight, so that late arrivals join within a step or two) ----
    while (!e->lu_inflight.empty() && event_done(e->lu_inflight.front())) {
        give_event(e, e->lu_inflight.front());
        e->lu_inflight.pop_front();
    }
    std::vector<bem2d_lu32*> stepping;
    for (Request* r : e->active)
        if (r->stage == FACTORING && !bem2d_internal_lu_done(r->lu)) stepping.push_back(r->lu);
    if (!stepping.empty() && e->lu_inflight.size() < 2) {
        // every update split from its diagonal block: the fused update holds a whole compute unit
        // (all registers of its SIMDs, 110 KB of LDS), the split one leaves room beside it for the
        // assembly wavefronts of other requests (400 served requests: 3.34 -> 2.83 s)
        const int rc = bem2d_internal_lu_advance(e->s_lu, stepping.data(), (int)stepping.size(),
                                                 1);
        if (rc) return rc;
        hipEvent_t v = take_event(e);
        HIPCHK(hipEventRecord(v, e->s_lu));
        e->lu_inflight.push_back(v);
        e->stats.lu_step_launches++;
        e->launched = true;
    }
    for (Request* r : e->active)
        if (r->stage == FACTORING && bem2d_internal_lu_done(r->lu)) {
            HIPCHK(hipEventRecord(r->ev, e->s_lu));
            r->stage = FACTORED;
        }
    // ---- refinement: one job at a time over all factorised requests ----
    if (e->job && bem2d_internal_ref_ready(e->job)) {
        bool all_done = false;
        const int rc = bem2d_internal_ref_collect(e->job, &all_done);
        for (size_t q = 0; q < e->job_reqs.size(); ++q) {
            Request* r = e->job_reqs[q];
            if (r->stage != REFINING || bem2d_internal_ref_spec_active(e->job, (int)q)) continue;
            bem2d_internal_ref_report(e->job, (int)q, &r->rep);
            if (rc) {
                r->status = BEM2D_ERR_DEVICE;
                finish(e, r, e->s_ref);
            } else if (bem2d_internal_ref_bad_pivot(e->job, (int)q) || !r->rep.converged) {
                // FP32 factors unusable (zero pivot, growth) or refinement stalled: the pivoted
                // FP64 LU solves the request.  Its refinement work has completed (the read-back
                // that retired it came after), and the LU stream waits for the refinement stream.
                r->rep.fallback_fp64 = 1;
                if (hipEventRecord(r->ev, e->s_ref) != hipSuccess ||
                    hipStreamWaitEvent(e->s_lu, r->ev, 0) != hipSuccess) {
                    fail_request(e, r, BEM2D_ERR_DEVICE);
                    continue;
                }
                solve_fp64(e, r);
            } else {
                finish(e, r, e->s_ref);
            }
        }
        if (all_done) {
            bem2d_internal_ref_destroy(e->job); // its workspace lives in the requests' blocks
            e->job = nullptr;
            e->job_reqs.clear();
        } else {
            const int rc2 = bem2d_internal_ref_step(e->job, e->s_ref);
            if (rc2) return rc2;
            e->stats.refine_step_launches++;
            e->launched = true;
        }
    }
    if (!e->job) {
        std::vector<RefSpec> specs;
        std::vector<Request*> reqs;
        size_t units = 0;
        for (Request* r : e->active)
            if (r->stage == FACTORED && event_done(r->ev)) {
                specs.push_back(RefSpec{r->lu, r->A, r->n, r->B, r->n, r->X, r->n, r->nrhs, r->tol,
                                        30, &r->bump});
                reqs.push_back(r);
                units += bem2d_internal_ref_units(r->nrhs);
            }
        if (!specs.empty()) {
            if (units > e->pinned_units) {
                // no job is in flight here (a new one is about to start), so both buffers are free
                const size_t want = std::max<size_t>(units, 2 * e->pinned_units + 64);
                if (e->pinned) (void)hipHostFree(e->pinned);
                if (e->dnorms) (void)hipFree(e->dnorms);
                e->pinned = nullptr;
                e->dnorms = nullptr;
                e->pinned_units = 0;
                if (hipHostMalloc((void**)&e->pinned, sizeof(unsigned long long) * kRefNormSlots *
                                                          want) != hipSuccess ||
                    hipMalloc((void**)&e->dnorms, sizeof(unsigned long long) * kRefNormSlots *
                                                      want) != hipSuccess)
                    return FAIL(BEM2D_ERR_NOMEM, "refinement read-back buffers");
                e->pinned_units = want;
            }
            RefJob* j = nullptr;
            const int rc = bem2d_internal_ref_create(e->ctx, e->arena.get(), specs.data(),
                                                     (int)specs.size(), e->pinned, e->dnorms, &j);
            if (rc) {
                for (Request* r : reqs) fail_request(e, r, rc);
            } else {
                e->job = j;
                e->job_reqs = reqs;
                for (Request* r : reqs) r->stage = REFINING;
                const int rc2 = bem2d_internal_ref_step(j, e->s_ref);
                if (rc2) return rc2;
                e->stats.refine_step_launches++;
                e->launched = true;
            }
        }
    }
    e->active.erase(std::remove_if(e->active.begin(), e->active.end(),
                                   [](Request* r) { return r->stage == DONE; }),
                    e->active.end());
    if (unfinished) *unfinished = (int)(e->queue.size() + e->active.size());
    return BEM2D_OK;
}

extern "C" int bem2d_engine_poll(bem2d_engine* e, int64_t ticket, int* done)
{
    if (!e || !done) return FAIL(BEM2D_ERR_ARG, "bem2d_engine_poll: null argument");
    auto it = e->all.find(ticket);
    if (it == e->all.end()) return FAIL(BEM2D_ERR_ARG, "unknown ticket %lld", (long long)ticket);
    *done = it->second->stage == DONE;
    return BEM2D_OK;
}

extern "C" int bem2d_engine_result(bem2d_engine* e, int64_t ticket, void* x, void* farfield,
                                   bem2d_result_info* info)
{
    if (!e) return FAIL(BEM2D_ERR_ARG, "bem2d_engine_result: null engine");
    auto it = e->all.find(ticket);
    if (it == e->all.end()) return FAIL(BEM2D_ERR_ARG, "unknown ticket %lld", (long long)ticket);
    Request* r = it->second;
    if (r->stage != DONE) return FAIL(BEM2D_ERR_ARG, "request %lld not finished", (long long)ticket);
    if (info) {
        info->status = r->status;
        info->n = r->n;
        info->nrhs = r->nrhs;
        info->precision = r->precision;
        info->iterations = r->rep.iterations;
        info->backward_error = r->rep.backward_error;
        info->fallback_fp64 = r->rep.fallback_fp64;
        info->t_arrival = r->t_arrival;
        info->t_admit = r->t_admit;
        info->t_done = r->t_done;
    }
    if (r->status == BEM2D_OK) {
        if (x && r->hx) std::memcpy(x, r->hx, sizeof(double2) * r->n * r->nrhs);
        if (farfield && r->hff && r->nang > 0)
            std::memcpy(farfield, r->hff, sizeof(double2) * r->nang * r->nrhs);
    }
    const int st = r->status;
    e->all.erase(it);
    free_request(e, r);
    return st == BEM2D_OK ? BEM2D_OK : FAIL(st, "request %lld failed", (long long)ticket);
}

extern "C" int bem2d_engine_get_stats(const bem2d_engine* e, bem2d_engine_stats* out)
{
    if (!e || !out) return FAIL(BEM2D_ERR_ARG, "bem2d_engine_get_stats: null argument");
    *out = e->stats;
    out->arena_bytes = (int64_t)e->arena->capacity();
    out->arena_in_use = (int64_t)e->arena->in_use();
    out->queued = (int64_t)e->queue.size();
    out->active = (int64_t)e->active.size();
    return BEM2D_OK;
}

// Runs a request trace: request q is submitted once the engine clock passes t0 + arrival[q], and
// the engine steps until every request is done (sleeping briefly when a step found nothing new to
// do).  tickets[q] receives the request's ticket for bem2d_engine_result.
extern "C" int bem2d_engine_run_trace(bem2d_engine* e, const bem2d_request* reqs,
                                      const double* arrival, int count, int64_t* tickets)
{
    if (!e || (!reqs && count > 0) || (!arrival && count > 0) || (!tickets && count > 0))
        return FAIL(BEM2D_ERR_ARG, "bem2d_engine_run_trace: null argument");
    const double t0 = engine_now(e);
    int next = 0, unfinished = 1;
    while (next < count || unfinished > 0) {
        const double now = engine_now(e);
        while (next < count && t0 + arrival[next] <= now) {
            const int rc = bem2d_engine_submit(e, &reqs[next], t0 + arrival[next], &tickets[next]);
            if (rc) return rc;
            ++next;
        }
        const int rc = bem2d_engine_step(e, &unfinished);
        if (rc) return rc;
        if (!e->launched && (unfinished > 0 || next < count))
            std::this_thread::sleep_for(std::chrono::microseconds(10));
    }
    return BEM2D_OK;
}
