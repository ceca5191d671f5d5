// Serving: a request loop over one batch (lcs_server_*).  Runs arrive as
// requests (size, particles, Stokes range, forcing, step count, seed); each
// is admitted when a slot and its memory footprint fit, gets its own box
// (fields and particles resident in HBM from admission to release), and joins
// the batch, so every server step advances all live runs with one launch per
// phase.  A run that reaches its step count leaves the batch at once and its
// slot goes to the next request (continuous batching); its box stays readable
// until the caller releases it.  A run's result does not depend on which
// other runs shared the batch (per-box reductions, see lcs_batch.hip).
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <string>
#include <algorithm>
#include <map>
#include <vector>

#include "lcs_internal.h"
#include "lcs_hsmac_xface.h"

namespace lcs {

struct Run {
    int64_t id;
    lcs_request req;
    lcs_box *box = nullptr;
    int status = LCS_RUN_QUEUED;
    double submit_s = 0, start_s = 0, finish_s = 0;
    size_t bytes = 0;
    int bypassed = 0;  // admissions of later requests while this one waited for memory
};

// one MI355X's HBM, the most a server's memory budget defaults to
static constexpr size_t kMaxServeBudget = size_t(288) * 1000 * 1000 * 1000;

struct Server {
    lcs_ctx *ctx = nullptr;
    lcs_server_config cfg{};
    lcs_batch *batch = nullptr;
    std::map<int64_t, Run> runs;
    std::deque<int64_t> queue;
    std::vector<int64_t> live;
    int64_t next_id = 1;
    size_t mem_used = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Device bytes a request will hold: 11 halo arrays (Us, Us2, P and the four
// HSMAC partners), the compact buffer of the x-slab HSMAC pair, the reduction
// partials, and per particle 17 properties plus staging, map and sort buffers.
static size_t footprint(const lcs_request &r)
{
    const size_t es = r.dtype == LCS_F64 ? 8 : 4;
    const size_t m = (size_t)(r.n + 6);
    size_t bytes = 11 * m * m * m * es + sizeof(double) * (size_t)kPartSlots + (1u << 20);
    bytes += (size_t)kXfVals * 2 * (size_t)(r.n - 1) * (size_t)(r.n - 1) * es;
    const size_t np = (size_t)std::max<int64_t>(r.particles, 0);
    bytes += np * (16 * es + 8 + 8 + es + 8 + 4 + 16 + 8);
    return bytes;
}

static int check_request(const lcs_request &r)
{
    if (r.n < 8 || r.n > 1024) return fail(LCS_ERR_INVALID, "request size n must be in 8..1024");
    if (r.dtype != LCS_F32 && r.dtype != LCS_F64) return fail(LCS_ERR_INVALID, "request dtype must be F32 or F64");
    if (r.particles < 0 || r.particles >= (int64_t(1) << 31)) return fail(LCS_ERR_INVALID, "bad particle count");
    if (r.particles > 0 && !(r.st_min > 0 && r.st_max >= r.st_min)) return fail(LCS_ERR_INVALID, "need 0 < st_min <= st_max");
    if (r.steps < 1) return fail(LCS_ERR_INVALID, "a request needs at least one step");
    if (!(r.k0 > 0) || !(r.Re > 0)) return fail(LCS_ERR_INVALID, "k0 and Re must be positive");
    if (r.forcing != LCS_FORCING_NONE && r.forcing != LCS_FORCING_LINEAR_POWER && r.forcing != LCS_FORCING_RC_ENERGY)
        return fail(LCS_ERR_INVALID, "unknown forcing kind");
    if (r.forcing == LCS_FORCING_RC_ENERGY && r.n % 8 != 0)
        return fail(LCS_ERR_INVALID, "rc-energy requests need n divisible by the coarse size 8");
    return LCS_OK;
}

// builds the run's box: seeded spectrum, periodic ghosts, one particle profile
static int materialise(Server &S, Run &run)
{
    const lcs_request &r = run.req;
    lcs_box_desc d = {r.n, r.n, r.n, 3, r.dtype, 1.0, 1.0, 1.0 / r.Re};
    int rc = lcs_box_create(S.ctx, &d, &run.box);
    if (rc) return rc;
    Box &b = *reinterpret_cast<Box *>(run.box);
    if ((rc = lcs_init_spectrum(run.box, r.seed, r.k0, 1.0))) return rc;
    const int uvwp = (1 << LCS_FIELD_U) | (1 << LCS_FIELD_V) | (1 << LCS_FIELD_W) | (1 << LCS_FIELD_P);
    if ((rc = lcs_fill_ghost(run.box, uvwp, 3, 1))) return rc;
    if (r.particles > 0) {
        lcs_particle_desc pd = {r.particles, r.particles, 0.0, 843.0, LCS_DRAG_LINEAR, 4};
        int prof = 0;
        if ((rc = lcs_particles_add(run.box, &pd, &prof))) return rc;
        if ((rc = lcs_init_particles(run.box, prof, r.seed ^ 0x9E3779B97F4A7C15ull, r.st_min, r.st_max, 0.5, 1)))
            return rc;
    }
    LCS_HIP(hipStreamSynchronize(b.stream));
    run.bytes = b.bytes;
    return LCS_OK;
}

static lcs_step_config run_config(const Server &S, const lcs_request &r)
{
    lcs_step_config c;
    memset(&c, 0, sizeof(c));
    c.cfl = S.cfg.cfl;
    c.dtmax = S.cfg.dtmax;
    c.hsmac = S.cfg.hsmac;
    c.forcing.kind = r.forcing;
    c.forcing.power = r.forcing_param;
    c.forcing.energy = r.forcing_param;
    c.forcing.kmin = 1;
    c.forcing.kmax = 2;
    c.forcing.coarse_size = 8;
    c.localize_interval = S.cfg.localize_interval;
    return c;
}

// Admits queued requests while a slot is free, in arrival order, except that
// a request whose footprint does not fit the memory left lets later requests
// that do fit go first.  Once it has been passed over max_live times nothing
// passes it any more, so a large run waits at most about one generation of
// the batch.  Every queued request fits the budget on its own (submit checks
// it), so an empty server always admits the head.
static int admit(Server &S)
{
    for (auto it = S.queue.begin(); it != S.queue.end() && (int)S.live.size() < S.cfg.max_live;) {
        Run &run = S.runs[*it];
        const size_t need = footprint(run.req);
        if (S.mem_used + need > S.cfg.memory_budget) {
            if (run.bypassed >= S.cfg.max_live) break;
            ++it;
            continue;
        }
        for (auto jt = S.queue.begin(); jt != it; ++jt) S.runs[*jt].bypassed++;
        it = S.queue.erase(it);
        int rc = materialise(S, run);
        if (rc == LCS_OK) {
            const lcs_step_config c = run_config(S, run.req);
            rc = lcs_batch_add(S.batch, run.box, &c);
        }
        if (rc) {
            run.status = LCS_RUN_FAILED;
            run.finish_s = S.now();
            if (run.box) (void)lcs_box_destroy(run.box);
            run.box = nullptr;
            continue;
        }
        run.status = LCS_RUN_LIVE;
        run.start_s = S.now();
        S.mem_used += std::max(need, run.bytes);
        S.live.push_back(run.id);
    }
    return LCS_OK;
}

}  // namespace lcs

using namespace lcs;

static Server *SV(lcs_server *p) { return reinterpret_cast<Server *>(p); }

extern "C" {

int lcs_server_create(lcs_ctx *ctx, const lcs_server_config *cfg, lcs_server **out)
{
    if (!ctx || !cfg || !out) return fail(LCS_ERR_INVALID, "null argument");
    if (cfg->max_live < 1) return fail(LCS_ERR_INVALID, "max_live must be >= 1");
    lcs_server_config c = *cfg;
    if (c.memory_budget == 0) {
        // the device's free memory now, less a tenth for the batch's own tables,
        // graphs and the runtime, and at most one MI355X's 288 GB
        size_t free_b = 0, total_b = 0;
        LCS_HIP(hipSetDevice(reinterpret_cast<Ctx *>(ctx)->device));
        LCS_HIP(hipMemGetInfo(&free_b, &total_b));
        c.memory_budget = std::min<size_t>(free_b / 10 * 9, kMaxServeBudget);
    }
    Server *S = new Server;
    S->ctx = ctx;
    S->cfg = c;
    int rc = lcs_batch_create(ctx, &S->batch);
    if (rc) {
        delete S;
        return rc;
    }
    *out = reinterpret_cast<lcs_server *>(S);
    return LCS_OK;
}

int lcs_server_destroy(lcs_server *server)
{
    if (!server) return LCS_OK;
    Server *S = SV(server);
    (void)lcs_batch_destroy(S->batch);
    for (auto &kv : S->runs)
        if (kv.second.box) (void)lcs_box_destroy(kv.second.box);
    delete S;
    return LCS_OK;
}

int lcs_server_submit(lcs_server *server, const lcs_request *req, int64_t *id)
{
    if (!server || !req) return fail(LCS_ERR_INVALID, "null argument");
    int rc = check_request(*req);
    if (rc) return rc;
    Server &S = *SV(server);
    if (footprint(*req) > S.cfg.memory_budget)
        return fail(LCS_ERR_CAPACITY, "request needs " + std::to_string(footprint(*req)) +
                                          " bytes of device memory, more than the server's budget of " +
                                          std::to_string(S.cfg.memory_budget));
    Run run;
    run.id = S.next_id++;
    run.req = *req;
    run.submit_s = S.now();
    S.runs[run.id] = run;
    S.queue.push_back(run.id);
    if (id) *id = run.id;
    return LCS_OK;
}

int lcs_server_step(lcs_server *server, int *n_live, int *n_finished)
{
    if (!server) return fail(LCS_ERR_INVALID, "null argument");
    Server &S = *SV(server);
    int rc = admit(S);
    if (rc) return rc;
    int finished = 0;
    if (!S.live.empty()) {
        if ((rc = lcs_batch_step(S.batch, nullptr, 0))) return rc;
        if ((rc = lcs_batch_sync(S.batch))) return rc;
        const double t = S.now();
        std::vector<int64_t> still;
        for (int64_t id : S.live) {
            Run &run = S.runs[id];
            const Box &b = *reinterpret_cast<Box *>(run.box);
            if (b.steps_host >= run.req.steps) {
                if ((rc = lcs_batch_remove(S.batch, run.box))) return rc;
                run.status = LCS_RUN_DONE;
                run.finish_s = t;
                S.mem_used -= std::min(S.mem_used, std::max(footprint(run.req), run.bytes));
                finished++;
            } else {
                still.push_back(id);
            }
        }
        S.live.swap(still);
    }
    if (n_live) *n_live = (int)S.live.size();
    if (n_finished) *n_finished = finished;
    return LCS_OK;
}

int lcs_server_status(lcs_server *server, int64_t id, lcs_run_status *st)
{
    if (!server || !st) return fail(LCS_ERR_INVALID, "null argument");
    Server &S = *SV(server);
    auto it = S.runs.find(id);
    if (it == S.runs.end()) return fail(LCS_ERR_INVALID, "no such request");
    const Run &run = it->second;
    st->id = run.id;
    st->status = run.status;
    st->steps_done = run.box ? reinterpret_cast<Box *>(run.box)->steps_host : 0;
    st->submit_s = run.submit_s;
    st->start_s = run.start_s;
    st->finish_s = run.finish_s;
    st->bytes = run.bytes;
    return LCS_OK;
}

int lcs_server_counts(lcs_server *server, int *queued, int *live, size_t *memory_used)
{
    if (!server) return fail(LCS_ERR_INVALID, "null argument");
    Server &S = *SV(server);
    if (queued) *queued = (int)S.queue.size();
    if (live) *live = (int)S.live.size();
    if (memory_used) *memory_used = S.mem_used;
    return LCS_OK;
}

int lcs_server_box(lcs_server *server, int64_t id, lcs_box **box)
{
    if (!server || !box) return fail(LCS_ERR_INVALID, "null argument");
    Server &S = *SV(server);
    auto it = S.runs.find(id);
    if (it == S.runs.end() || !it->second.box) return fail(LCS_ERR_INVALID, "request has no box");
    *box = it->second.box;
    return LCS_OK;
}

int lcs_server_release(lcs_server *server, int64_t id)
{
    if (!server) return fail(LCS_ERR_INVALID, "null argument");
    Server &S = *SV(server);
    auto it = S.runs.find(id);
    if (it == S.runs.end()) return fail(LCS_ERR_INVALID, "no such request");
    Run &run = it->second;
    if (run.status == LCS_RUN_LIVE || run.status == LCS_RUN_QUEUED)
        return fail(LCS_ERR_STATE, "request is still queued or running");
    if (run.box) (void)lcs_box_destroy(run.box);
    S.runs.erase(it);
    return LCS_OK;
}

}  // extern "C"
