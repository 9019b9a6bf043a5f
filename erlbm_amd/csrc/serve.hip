// serve.hip -- continuous-batching request loop for uneven solves.
//
// Requests of four kinds arrive over time: Taylor-Green map points
// (D3Q19 16x16x1), obstacle map points (D3Q19 22x12x1), D2Q9 Taylor-Green
// accuracy runs and D2Q9 double-shear-layer runs (the reference's four
// drivers, src/main_*.cpp).  The server keeps every admitted request's
// lattice resident in HBM until its stop test or step budget retires it:
//  * map points live in slots of per-(geometry, precision, collision) pools;
//    one chunk of the persistent LDS-resident kernel (maps.hip) advances
//    every active slot of a pool by up to `chunk_steps` steps, replayed from
//    a captured HIP graph (queue reset + kernel), the active-slot list being
//    rebuilt between chunks (retired slots drop out, new ones join);
//  * D2Q9 runs are plain lattices (lattice.hip) on streams of their own,
//    advanced by a time slice per round, concurrently with the pools.
// Admission is by HBM budget: a pool reserves its slot array the first time
// it is used, a plain lattice needs its two population buffers; a request
// waits until its bytes fit.  Per-request arithmetic does not depend on what
// else is in flight, so a result is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "maps.hpp"

using namespace erlbm;

namespace {

constexpr int Q19 = 19;

struct Pool {
    erlbm_batch *geom = nullptr;   // shared geometry (owned by the server)
    int precision = ERLBM_FP64, collision = ERLBM_COL_EXT_REG;
    int cap = 0;
    size_t bytes = 0;
    MapJob *d_jobs = nullptr;
    MapResult *d_results = nullptr;
    void *d_saved = nullptr;
    int *d_idx = nullptr;
    int *d_n = nullptr;
    unsigned int *d_counter = nullptr;
    std::vector<MapJob> h_jobs;
    std::vector<MapResult> h_results;
    std::vector<long long> slot_req;   // request id per slot, -1 = free
    std::vector<int> free_slots;
    std::vector<int> active;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
};

struct Big {
    long long id = 0;
    erlbm_request req;
    erlbm_lattice *lat = nullptr;
    hipStream_t stream = nullptr;
    long long target = 0, done = 0;
    size_t bytes = 0;
    double u0 = 0, visc = 0;
    double step_us = 10.0;
};

struct Pending {
    long long id;
    erlbm_request req;
    double t_submit;
};

double now_s(const std::chrono::steady_clock::time_point &t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct erlbm_server {
    size_t budget = 0, used = 0;
    int chunk_steps = 1000;
    int tiny_slots = 4096;
    double slice_us = 5000.0;
    bool use_graphs = true;
    hipStream_t s_tiny = nullptr;
    erlbm_batch *tg_geom = nullptr, *obst_geom = nullptr;
    std::map<int, std::unique_ptr<Pool>> pools;   // key = geom * 64 + precision * 8 + collision
    std::deque<Pending> queue;
    std::vector<std::unique_ptr<Big>> bigs;
    std::vector<erlbm_response> finished;
    std::map<long long, double> admit_time, submit_time;
    long long next_id = 1;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
};

namespace {

size_t pool_bytes(const erlbm_server *s, const erlbm_batch *geom, int precision);
int pool_key(int geom_kind, int precision, int collision);

int make_pool(erlbm_server *s, int geom_kind, int precision, int collision, Pool **out)
{
    const int key = pool_key(geom_kind, precision, collision);
    auto it = s->pools.find(key);
    if (it != s->pools.end()) {
        *out = it->second.get();
        return ERLBM_OK;
    }
    auto p = std::make_unique<Pool>();
    p->geom = geom_kind == 0 ? s->tg_geom : s->obst_geom;
    p->precision = precision;
    p->collision = collision;
    p->cap = s->tiny_slots;
    const size_t sb = tiny::state_bytes(p->geom, precision);
    p->bytes = pool_bytes(s, p->geom, precision);
    if (s->used + p->bytes > s->budget) return ERLBM_ERR_NOMEM;
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t r) {
        if (r != hipSuccess && e == hipSuccess) e = r;
    };
    ok(hipMalloc(&p->d_jobs, sizeof(MapJob) * p->cap));
    ok(hipMalloc(&p->d_results, sizeof(MapResult) * p->cap));
    ok(hipMalloc(&p->d_saved, sb * p->cap));
    ok(hipMalloc(&p->d_idx, sizeof(int) * p->cap));
    ok(hipMalloc(&p->d_n, sizeof(int)));
    ok(hipMalloc(&p->d_counter, sizeof(unsigned int)));
    if (e != hipSuccess) {
        (void)hipFree(p->d_jobs);
        (void)hipFree(p->d_results);
        (void)hipFree(p->d_saved);
        (void)hipFree(p->d_idx);
        (void)hipFree(p->d_n);
        (void)hipFree(p->d_counter);
        set_last_error("erlbm_server: pool allocation", e);
        return ERLBM_ERR_NOMEM;
    }
    p->h_jobs.resize(p->cap);
    p->h_results.assign(p->cap, MapResult{});
    p->slot_req.assign(p->cap, -1);
    for (int k = p->cap - 1; k >= 0; k--) p->free_slots.push_back(k);
    s->used += p->bytes;
    *out = p.get();
    s->pools[key] = std::move(p);
    return ERLBM_OK;
}

void free_pool(Pool *p)
{
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    if (p->graph_def) (void)hipGraphDestroy(p->graph_def);
    (void)hipFree(p->d_jobs);
    (void)hipFree(p->d_results);
    (void)hipFree(p->d_saved);
    (void)hipFree(p->d_idx);
    (void)hipFree(p->d_n);
    (void)hipFree(p->d_counter);
}

// One chunk of a pool: the captured graph (queue reset + persistent kernel).
int run_pool_chunk(erlbm_server *s, Pool *p)
{
    const int n = (int)p->active.size();
    ERLBM_HIP_TRY(hipMemcpyAsync(p->d_idx, p->active.data(), sizeof(int) * n, hipMemcpyHostToDevice, s->s_tiny));
    ERLBM_HIP_TRY(hipMemcpyAsync(p->d_n, &n, sizeof(int), hipMemcpyHostToDevice, s->s_tiny));
    if (s->use_graphs) {
        if (!p->graph) {
            ERLBM_HIP_TRY(hipStreamBeginCapture(s->s_tiny, hipStreamCaptureModeThreadLocal));
            int rc = tiny::launch(p->geom, p->precision, p->collision, p->d_jobs, p->d_results, p->d_idx, p->d_n,
                                  p->cap, p->d_counter, s->chunk_steps, p->d_saved, s->s_tiny);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s->s_tiny, &g);
            if (rc) return rc;
            if (e != hipSuccess) {
                set_last_error("erlbm_server: graph capture", e);
                return ERLBM_ERR_HIP;
            }
            p->graph_def = g;
            ERLBM_HIP_TRY(hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0));
        }
        ERLBM_HIP_TRY(hipGraphLaunch(p->graph, s->s_tiny));
        return ERLBM_OK;
    }
    return tiny::launch(p->geom, p->precision, p->collision, p->d_jobs, p->d_results, p->d_idx, p->d_n, p->cap,
                        p->d_counter, s->chunk_steps, p->d_saved, s->s_tiny);
}

int resolve_precision(int precision, double tau)
{
    if (precision == ERLBM_FP_AUTO) return (tau - 0.5 < 1e-3) ? ERLBM_FP64 : ERLBM_FP32_SHIFTED;
    return precision;
}

bool is_map(const erlbm_request &r)
{
    return r.kind == ERLBM_REQ_TG_MAP_POINT || r.kind == ERLBM_REQ_OBSTACLE_MAP_POINT;
}

// The map driver's job for a map-point request, precision resolved.
int map_job(const erlbm_request &r, MapJob *J)
{
    int rc = r.kind == ERLBM_REQ_TG_MAP_POINT
                 ? erlbm_tg_map_jobs(r.n_i, r.n_j, r.collision, r.tau_nh, r.precision, r.i, r.i + 1, r.j, r.j + 1, J)
                 : erlbm_obstacle_map_jobs(r.n_i, r.n_j, 0.1, r.precision, r.i, r.i + 1, r.j, r.j + 1, J);
    if (rc) return rc;
    J->precision = resolve_precision(J->precision, J->tau);
    return ERLBM_OK;
}

int pool_key(int geom_kind, int precision, int collision) { return geom_kind * 64 + precision * 8 + collision; }

size_t pool_bytes(const erlbm_server *s, const erlbm_batch *geom, int precision)
{
    const size_t sb = tiny::state_bytes(geom, precision);
    return (size_t)s->tiny_slots * (sb + sizeof(MapJob) + sizeof(MapResult) + sizeof(int));
}

size_t lattice_bytes(const erlbm_request &r)
{
    const int prec = r.precision == ERLBM_FP_AUTO ? ERLBM_FP64 : r.precision;
    const size_t elem = prec == ERLBM_FP64 ? 8 : (prec == ERLBM_FP16_SHIFTED ? 2 : 4);
    return 2 * elem * 9 * (size_t)r.L * r.L + 8 * 3 * (size_t)r.L * r.L;   // populations + monitors
}

// Bytes of budget a request would newly reserve: its pool's slot array if
// that pool does not exist yet (map points), its lattice (D2Q9 runs); 0 when
// it only needs a free slot of an existing pool or is malformed (admit()
// then answers it).  *pool: the existing pool it would use, if any.
size_t request_bytes(const erlbm_server *s, const erlbm_request &r, const Pool **pool)
{
    *pool = nullptr;
    if (is_map(r)) {
        MapJob J;
        if (map_job(r, &J)) return 0;
        const int kind = r.kind == ERLBM_REQ_TG_MAP_POINT ? 0 : 1;
        auto it = s->pools.find(pool_key(kind, J.precision, J.collision));
        if (it != s->pools.end()) {
            *pool = it->second.get();
            return 0;
        }
        return pool_bytes(s, kind == 0 ? s->tg_geom : s->obst_geom, J.precision);
    }
    if ((r.kind != ERLBM_REQ_TAYLOR_GREEN && r.kind != ERLBM_REQ_DOUBLE_SHEAR) || r.L < 4 || r.L > 65536) return 0;
    return lattice_bytes(r);
}

// Try to admit one request; ERLBM_ERR_NOMEM = does not fit yet.
int admit(erlbm_server *s, const Pending &pd)
{
    const erlbm_request &r = pd.req;
    if (is_map(r)) {
        MapJob J;
        const int kind = r.kind == ERLBM_REQ_TG_MAP_POINT ? 0 : 1;
        int rc = map_job(r, &J);
        if (rc) return rc;
        Pool *p = nullptr;
        rc = make_pool(s, kind, J.precision, J.collision, &p);
        if (rc) return rc;
        if (p->free_slots.empty()) return ERLBM_ERR_NOMEM;
        const int slot = p->free_slots.back();
        p->free_slots.pop_back();
        p->slot_req[slot] = pd.id;
        p->h_jobs[slot] = J;
        p->h_results[slot] = MapResult{};
        ERLBM_HIP_TRY(hipMemcpyAsync(p->d_jobs + slot, &J, sizeof(MapJob), hipMemcpyHostToDevice, s->s_tiny));
        ERLBM_HIP_TRY(hipMemcpyAsync(p->d_results + slot, &p->h_results[slot], sizeof(MapResult),
                                     hipMemcpyHostToDevice, s->s_tiny));
        p->active.push_back(slot);
        return ERLBM_OK;
    }
    if (r.kind != ERLBM_REQ_TAYLOR_GREEN && r.kind != ERLBM_REQ_DOUBLE_SHEAR) return ERLBM_ERR_ARG;
    const int L = r.L;
    if (L < 4 || L > 65536) return ERLBM_ERR_ARG;
    const int prec = r.precision == ERLBM_FP_AUTO ? ERLBM_FP64 : r.precision;
    const size_t elem = prec == ERLBM_FP64 ? 8 : (prec == ERLBM_FP16_SHIFTED ? 2 : 4);
    const size_t bytes = lattice_bytes(r);
    if (s->used + bytes > s->budget) return ERLBM_ERR_NOMEM;
    auto b = std::make_unique<Big>();
    b->id = pd.id;
    b->req = r;
    b->bytes = bytes;
    ERLBM_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    int rc = erlbm_lattice_create(9, L, L, 1, nullptr, prec, b->stream, &b->lat);
    if (rc) {
        (void)hipStreamDestroy(b->stream);
        return rc;
    }
    std::vector<double> f((size_t)L * L * 9);
    double tau, tnh;
    if (r.kind == ERLBM_REQ_TAYLOR_GREEN) {
        // main_taylor-green.cpp:105-158, Fig. 3 mode: U0 = 2/L, Re = 50
        b->u0 = 2. / L;
        b->visc = b->u0 * L / 50.;
        tau = 0.5 + 3. * b->visc;
        tnh = r.collision == ERLBM_COL_REG ? 1.0 : (r.collision == ERLBM_COL_BGK ? tau : r.tau_nh);
        erlbm_init_taylor_green(9, L, b->u0, tau, 1, 1, f.data());
        b->target = (long long)round(L / b->u0) + 1;
    } else {
        // main_double-shear-layer.cpp:83-124: Re = 80000, U0 = 1/32
        const double U0 = 1. / 32.;
        tau = 0.5 + 3. * U0 * L / 80000.;
        tnh = r.collision == ERLBM_COL_REG ? 1.0 : r.tau_nh;
        erlbm_init_double_shear(L, 1.0, f.data());
        const long long per_t = (long long)round(L / U0);
        b->target = (long long)round(r.t_final * per_t) + 1;
    }
    rc = erlbm_set_params(b->lat, tau, tnh, 1.0, r.collision);
    if (!rc) rc = erlbm_upload_aos(b->lat, f.data());
    if (rc) {
        erlbm_lattice_destroy(b->lat);
        (void)hipStreamDestroy(b->stream);
        return rc;
    }
    // time per step: HBM-bound above ~64k sites, launch-bound below
    b->step_us = std::max(4.0, (double)L * L * 2.0 * elem * 9 / 5.0e6);
    s->used += bytes;
    s->bigs.push_back(std::move(b));
    return ERLBM_OK;
}

// Bytes held by pools with no live request other than `keep`.
size_t idle_pool_bytes(const erlbm_server *s, const Pool *keep)
{
    size_t b = 0;
    for (const auto &kv : s->pools)
        if (kv.second.get() != keep && kv.second->active.empty()) b += kv.second->bytes;
    return b;
}

// Frees the pools that hold no live request, `keep` aside (a pool's slot
// array counts against the budget for as long as the pool exists).
void release_idle_pools(erlbm_server *s, const Pool *keep)
{
    for (auto it = s->pools.begin(); it != s->pools.end();) {
        Pool *p = it->second.get();
        if (p == keep || !p->active.empty()) {
            ++it;
            continue;
        }
        free_pool(p);
        s->used -= p->bytes;
        it = s->pools.erase(it);
    }
}

void respond(erlbm_server *s, long long id, int kind, int unstable, long long steps, const double *v, double t_done)
{
    erlbm_response o{};
    o.id = id;
    o.kind = kind;
    o.status = ERLBM_OK;
    o.unstable = unstable;
    o.steps = steps;
    for (int k = 0; k < 3; k++) o.value[k] = v[k];
    o.t_admit = s->admit_time.count(id) ? s->admit_time[id] : 0.0;
    o.t_submit = s->submit_time.count(id) ? s->submit_time[id] : 0.0;
    o.t_done = t_done;
    s->finished.push_back(o);
    s->admit_time.erase(id);
    s->submit_time.erase(id);
}

}  // namespace

extern "C" int erlbm_server_create(long long hbm_budget_bytes, int chunk_steps, int tiny_slots, void *stream,
                                   erlbm_server **out)
{
    if (!out || hbm_budget_bytes <= 0 || chunk_steps < 1 || tiny_slots < 1) return ERLBM_ERR_ARG;
    auto *s = new erlbm_server();
    s->budget = (size_t)hbm_budget_bytes;
    s->chunk_steps = chunk_steps;
    s->tiny_slots = tiny_slots;
    const char *g = getenv("ERLBM_SERVER_GRAPHS");
    if (g && *g == '0') s->use_graphs = false;
    int rc = ERLBM_OK;
    if (stream) s->s_tiny = (hipStream_t)stream;
    else if (hipStreamCreateWithFlags(&s->s_tiny, hipStreamNonBlocking) != hipSuccess) rc = ERLBM_ERR_HIP;
    int mask[22 * 12];
    if (!rc) rc = erlbm_batch_create(16, 16, nullptr, -1, -1, s->s_tiny, &s->tg_geom);
    if (!rc) rc = erlbm_obstacle_mask(mask);
    if (!rc) rc = erlbm_batch_create(22, 12, mask, 1, 20, s->s_tiny, &s->obst_geom);
    if (rc) {
        erlbm_server_destroy(s);
        return rc;
    }
    *out = s;
    return ERLBM_OK;
}

extern "C" int erlbm_server_destroy(erlbm_server *s)
{
    if (!s) return ERLBM_OK;
    if (s->s_tiny) (void)hipStreamSynchronize(s->s_tiny);
    for (auto &b : s->bigs) {
        erlbm_lattice_destroy(b->lat);
        (void)hipStreamDestroy(b->stream);
    }
    for (auto &kv : s->pools) free_pool(kv.second.get());
    erlbm_batch_destroy(s->tg_geom);
    erlbm_batch_destroy(s->obst_geom);
    delete s;
    return ERLBM_OK;
}

extern "C" int erlbm_server_submit(erlbm_server *s, const erlbm_request *req, long long *id)
{
    if (!s || !req) return ERLBM_ERR_ARG;
    Pending p;
    p.id = s->next_id++;
    p.req = *req;
    p.t_submit = now_s(s->t0);
    s->submit_time[p.id] = p.t_submit;
    s->queue.push_back(p);
    if (id) *id = p.id;
    return ERLBM_OK;
}

extern "C" int erlbm_server_clock(erlbm_server *s, double *t)
{
    if (!s || !t) return ERLBM_ERR_ARG;
    *t = now_s(s->t0);
    return ERLBM_OK;
}

extern "C" int erlbm_server_pump(erlbm_server *s, int *in_flight, int *queued)
{
    if (!s) return ERLBM_ERR_ARG;
    // 1. admission, first fit in arrival order.  A request that needs more
    // than the whole budget never fits and is answered ERLBM_ERR_NOMEM; one
    // that would fit without the idle pools' reservations frees them; any
    // other waits for in-flight requests to retire (they all do: every
    // request has a step budget)
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        int rc = admit(s, *it);
        if (rc == ERLBM_ERR_NOMEM) {
            const Pool *keep = nullptr;
            const size_t need = request_bytes(s, it->req, &keep);
            const size_t idle = idle_pool_bytes(s, keep);
            if (need > s->budget) {
                set_last_error_msg("erlbm_server: request needs more than the HBM budget");
            } else if (idle > 0 && s->used - idle + need <= s->budget) {
                release_idle_pools(s, keep);
                rc = admit(s, *it);
            }
            if (rc == ERLBM_ERR_NOMEM && need <= s->budget) {
                ++it;
                continue;
            }
        }
        if (rc == ERLBM_OK) {
            s->admit_time[it->id] = now_s(s->t0);
            it = s->queue.erase(it);
        } else {
            erlbm_response o{};
            o.id = it->id;
            o.kind = it->req.kind;
            o.status = rc;
            o.t_submit = it->t_submit;
            o.t_done = now_s(s->t0);
            s->finished.push_back(o);
            s->submit_time.erase(it->id);
            it = s->queue.erase(it);
        }
    }
    // 2. one chunk of every pool with work, one time slice of every plain run
    for (auto &kv : s->pools) {
        Pool *p = kv.second.get();
        if (p->active.empty()) continue;
        int rc = run_pool_chunk(s, p);
        if (rc) return rc;
    }
    for (auto &b : s->bigs) {
        long long n = std::max(1LL, (long long)(s->slice_us / b->step_us));
        n = std::min(n, b->target - b->done);
        if (n > 0) {
            int rc = erlbm_step(b->lat, (int)n, 0, 0.0);
            if (rc) return rc;
            b->done += n;
        }
    }
    ERLBM_HIP_TRY(hipStreamSynchronize(s->s_tiny));
    // 3. retire finished map points
    for (auto &kv : s->pools) {
        Pool *p = kv.second.get();
        if (p->active.empty()) continue;
        ERLBM_HIP_TRY(hipMemcpy(p->h_results.data(), p->d_results, sizeof(MapResult) * p->cap,
                                hipMemcpyDeviceToHost));
        const double t = now_s(s->t0);
        std::vector<int> still;
        for (int slot : p->active) {
            const MapResult &r = p->h_results[slot];
            if (r.done) {
                const double v[3] = {r.margin, r.e0, r.energy};
                respond(s, p->slot_req[slot], p->geom == s->tg_geom ? ERLBM_REQ_TG_MAP_POINT
                                                                     : ERLBM_REQ_OBSTACLE_MAP_POINT,
                        r.unstable, r.steps, v, t);
                s->finished.back().precision = r.precision;
                p->slot_req[slot] = -1;
                p->free_slots.push_back(slot);
            } else {
                still.push_back(slot);
            }
        }
        p->active.swap(still);
    }
    // 4. retire finished plain runs
    for (auto it = s->bigs.begin(); it != s->bigs.end();) {
        Big *b = it->get();
        int rc = erlbm_synchronize(b->lat);
        if (rc) return rc;
        if (b->done < b->target) {
            ++it;
            continue;
        }
        double v[3] = {0, 0, 0};
        long long steps = 0;
        int reason = 0;
        erlbm_status(b->lat, &steps, &reason);
        if (b->req.kind == ERLBM_REQ_TAYLOR_GREEN) {
            const int L = b->req.L;
            std::vector<double> f((size_t)L * L * 9);
            rc = erlbm_download_aos(b->lat, ERLBM_POST_COLLISION, f.data());
            if (!rc) rc = erlbm_taylor_green_error(9, L, b->u0, b->visc, steps - 1, f.data(), v);
        } else {
            rc = erlbm_kinetic_energy(b->lat, ERLBM_POST_COLLISION, 0, &v[0]);
        }
        if (rc) return rc;
        respond(s, b->id, b->req.kind, reason != 0, steps, v, now_s(s->t0));
        s->finished.back().precision = b->req.precision == ERLBM_FP_AUTO ? ERLBM_FP64 : b->req.precision;
        erlbm_lattice_destroy(b->lat);
        (void)hipStreamDestroy(b->stream);
        s->used -= b->bytes;
        it = s->bigs.erase(it);
    }
    if (in_flight) {
        int n = (int)s->bigs.size();
        for (auto &kv : s->pools) n += (int)kv.second->active.size();
        *in_flight = n;
    }
    if (queued) *queued = (int)s->queue.size();
    return ERLBM_OK;
}

extern "C" int erlbm_server_poll(erlbm_server *s, erlbm_response *out, int max_out, int *n_out)
{
    if (!s || !n_out || (max_out > 0 && !out)) return ERLBM_ERR_ARG;
    const int n = std::min(max_out, (int)s->finished.size());
    for (int k = 0; k < n; k++) out[k] = s->finished[k];
    s->finished.erase(s->finished.begin(), s->finished.begin() + n);
    *n_out = n;
    return ERLBM_OK;
}

extern "C" int erlbm_server_memory(erlbm_server *s, long long *used, long long *budget)
{
    if (!s) return ERLBM_ERR_ARG;
    if (used) *used = (long long)s->used;
    if (budget) *budget = (long long)s->budget;
    return ERLBM_OK;
}
