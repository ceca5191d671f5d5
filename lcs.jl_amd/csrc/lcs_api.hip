// C ABI of the MI355X backend (include/lcs_mi355x.h): handles, memory,
// per-phase entry points mirroring the reference's positional methods, and
// the whole-step driver (LCS.timestep! + LCS.evolve!).
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <cstring>
#include <string>

#include "lcs_internal.h"

namespace lcs {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }
int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

__global__ void k_set_double(double *dst, double v) { *dst = v; }

// ---- scratch of the allocation-free rc-energy forcing ------------------------
// The coarse grids, spectra and crop of lcs_force_rcf_noalloc live in this
// file-scope device array: every device the library runs on holds one copy
// until the process ends.  In host code `scratch` is only a host-side stand-in
// for the symbol; its device address comes from hipGetSymbolAddress, once per
// device (lcs_ctx_create, with that device current), and only that address is
// ever handed to kernels.
constexpr size_t kRcfScratchDoubles = size_t(1) << 17;  // 1 MiB: c <= 32 with kmax <= 8
static __device__ __align__(256) double scratch[kRcfScratchDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices] = {};

static int register_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(LCS_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return LCS_OK;
    double *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(LCS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = ptr;
    return LCS_OK;
}

__global__ void k_set_state(DevScalars *S, int64_t step, double t)
{
    S->step = step;
    S->t = t;
}

__global__ void k_advance_state(DevScalars *S)
{
    // evolve!(state): src/evolve.jl:10-13
    S->step += 1;
    S->t += S->dt;
}

ProfScope::ProfScope(Box &box, const char *name, double bytes) : b(box)
{
    if (!b.prof.on) return;
    auto &E = b.prof.entries;
    for (size_t q = 0; q < E.size(); q++)
        if (E[q].name == name) idx = (int)q;
    if (idx < 0) {
        E.emplace_back();
        E.back().name = name;
        idx = (int)E.size() - 1;
    }
    auto take = [&]() {
        hipEvent_t e = nullptr;
        if (!b.prof.pool.empty()) {
            e = b.prof.pool.back();
            b.prof.pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
        return e;
    };
    e0 = take();
    e1 = take();
    if (e0) (void)hipEventRecord(e0, b.stream);
    E[idx].bytes += bytes;
    E[idx].launches += 1;
}

ProfScope::~ProfScope()
{
    if (idx < 0) return;
    if (e1) (void)hipEventRecord(e1, b.stream);
    if (e0 && e1) b.prof.entries[idx].pending.emplace_back(e0, e1);
}

void prof_uncount(Box &b, const char *name, int64_t launches, double bytes)
{
    if (!b.prof.on) return;
    for (auto &e : b.prof.entries)
        if (e.name == name) {
            e.launches -= launches;
            e.bytes -= bytes;
        }
}

static void prof_collect(Box &b)
{
    for (auto &e : b.prof.entries) {
        for (auto &pr : e.pending) {
            float ms = 0;
            if (hipEventSynchronize(pr.second) == hipSuccess && hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess)
                e.ms += ms;
            b.prof.pool.push_back(pr.first);
            b.prof.pool.push_back(pr.second);
        }
        e.pending.clear();
    }
}

static int set_phase_dt(Box &b, double dt)
{
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(1), 0, b.stream, &b.S->dt_phase, dt);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

static int read_scalars(Box &b)
{
    LCS_HIP(hipMemcpyAsync(b.S_host, b.S, sizeof(DevScalars), hipMemcpyDeviceToHost, b.stream));
    LCS_HIP(hipStreamSynchronize(b.stream));
    return LCS_OK;
}

static int device_alloc(Box &b, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(LCS_ERR_CAPACITY, std::string("hipMalloc of ") + std::to_string(bytes) +
                                          " bytes failed: " + hipGetErrorString(e));
    }
    b.bytes += bytes;
    return LCS_OK;
}

static int check_box(const lcs_box *box)
{
    if (!box) return fail(LCS_ERR_INVALID, "null box");
    return LCS_OK;
}

static int check_stage(int stage)
{
    if (stage != 1 && stage != 2) return fail(LCS_ERR_INVALID, "stage must be 1 or 2");
    return LCS_OK;
}

static Box *B(lcs_box *b) { return reinterpret_cast<Box *>(b); }
static const Box *B(const lcs_box *b) { return reinterpret_cast<const Box *>(b); }

static int ensure_dudt(Box &b)
{
    for (int q = LCS_FIELD_DU; q <= LCS_FIELD_DW; q++)
        if (!b.f[q]) {
            int rc = device_alloc(b, &b.f[q], (size_t)b.g.cells() * b.esize);
            if (rc) return rc;
            LCS_HIP(hipMemsetAsync(b.f[q], 0, (size_t)b.g.cells() * b.esize, b.stream));
        }
    return LCS_OK;
}

static int profile_of(Box &b, int profile, ParticleProfile **pp)
{
    if (profile < 0 || profile >= (int)b.profiles.size()) return fail(LCS_ERR_STATE, "no such particle profile");
    *pp = &b.profiles[profile];
    return LCS_OK;
}

}  // namespace lcs

using namespace lcs;

extern "C" {

const char *lcs_last_error(void) { return g_error.c_str(); }
const char *lcs_version(void) { return "lcs-mi355x 0.1 (gfx950)"; }

int lcs_ctx_create(int device, lcs_ctx **out)
{
    if (!out) return fail(LCS_ERR_INVALID, "null output pointer");
    int n = 0;
    LCS_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(LCS_ERR_INVALID, "no such device");
    LCS_HIP(hipSetDevice(device));
    int rc = register_scratch(device);
    if (rc) return rc;
    Ctx *c = new Ctx;
    c->device = device;
    *out = reinterpret_cast<lcs_ctx *>(c);
    return LCS_OK;
}

int lcs_ctx_destroy(lcs_ctx *ctx)
{
    delete reinterpret_cast<Ctx *>(ctx);
    return LCS_OK;
}

int lcs_box_create(lcs_ctx *ctx, const lcs_box_desc *d, lcs_box **out)
{
    if (!ctx || !d || !out) return fail(LCS_ERR_INVALID, "null argument");
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(LCS_ERR_INVALID, "grid dims must be positive");
    if (d->halo < 1) return fail(LCS_ERR_INVALID, "halo_size must be >= 1");
    if (d->dtype != LCS_F64 && d->dtype != LCS_F32) return fail(LCS_ERR_INVALID, "dtype must be LCS_F64 or LCS_F32");
    if (!(d->NU_m2_s > 0) || !(d->U0_m_s > 0) || !(d->L0_m > 0))
        return fail(LCS_ERR_INVALID, "flow params must be positive");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    LCS_HIP(hipSetDevice(c->device));
    Box *b = new Box;
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && cus > 0)
            b->cus = cus;
    }
    b->ctx = c;
    b->desc = *d;
    b->g = make_geom(d->nx, d->ny, d->nz, d->halo);
    b->esize = d->dtype == LCS_F64 ? 8 : 4;
    int rc = LCS_OK;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(LCS_ERR_HIP, "hipStreamCreate failed");
    }
    size_t fb = (size_t)b->g.len * b->esize;
    for (int q = 0; q <= LCS_FIELD_P && rc == LCS_OK; q++) {
        rc = device_alloc(*b, &b->f[q], fb);
        if (rc == LCS_OK && hipMemsetAsync(b->f[q], 0, fb, b->stream) != hipSuccess)
            rc = fail(LCS_ERR_HIP, "hipMemset failed");
    }
    if (rc == LCS_OK) rc = device_alloc(*b, (void **)&b->S, sizeof(DevScalars));
    if (rc == LCS_OK) rc = device_alloc(*b, (void **)&b->part, sizeof(double) * kPartSlots);
    if (rc == LCS_OK && hipHostMalloc((void **)&b->S_host, sizeof(DevScalars), 0) != hipSuccess)
        rc = fail(LCS_ERR_HIP, "hipHostMalloc failed");
    if (rc == LCS_OK && hipMemsetAsync(b->S, 0, sizeof(DevScalars), b->stream) != hipSuccess)
        rc = fail(LCS_ERR_HIP, "hipMemset failed");
    if (rc == LCS_OK && hipStreamSynchronize(b->stream) != hipSuccess) rc = fail(LCS_ERR_HIP, "sync failed");
    if (rc != LCS_OK) {
        std::string msg = g_error;
        lcs_box_destroy(reinterpret_cast<lcs_box *>(b));
        set_error(msg);
        return rc;
    }
    *out = reinterpret_cast<lcs_box *>(b);
    return LCS_OK;
}

int lcs_box_destroy(lcs_box *box)
{
    if (!box) return LCS_OK;
    Box *b = B(box);
    (void)batch_detach_box(*b);
    (void)hipSetDevice(b->ctx->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (void *p : b->f)
        if (p) (void)hipFree(p);
    for (void *p : b->spare)
        if (p) (void)hipFree(p);
    for (auto &pp : b->profiles) {
        for (void *p : pp.props)
            if (p) (void)hipFree(p);
        void *extra[] = {pp.hashes,  pp.perm,     pp.starts, pp.stops,  pp.scratch,
                         pp.scratch_t, pp.sort_tmp, pp.loc,    pp.loc_tmp};
        for (void *p : extra)
            if (p) (void)hipFree(p);
    }
    if (b->S) (void)hipFree(b->S);
    if (b->part) (void)hipFree(b->part);
    if (b->xf) (void)hipFree(b->xf);
    if (b->rcf) (void)hipFree(b->rcf);
    if (b->rcf_modes) (void)hipFree(b->rcf_modes);
    if (b->S_host) (void)hipHostFree(b->S_host);
    prof_collect(*b);
    for (hipEvent_t e : b->prof.pool) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LCS_OK;
}

int lcs_box_bytes(const lcs_box *box, size_t *bytes)
{
    if (check_box(box) || !bytes) return fail(LCS_ERR_INVALID, "null argument");
    *bytes = B(box)->bytes;
    return LCS_OK;
}

int lcs_box_describe(const lcs_box *box, lcs_box_desc *desc, int *nprofiles)
{
    if (check_box(box) || !desc) return fail(LCS_ERR_INVALID, "null argument");
    const Box &b = *B(box);
    *desc = b.desc;
    if (nprofiles) *nprofiles = (int)b.profiles.size();
    return LCS_OK;
}

static int field_xfer(lcs_box *box, int field, void *host, size_t bytes, bool up)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    if (field < 0 || field > LCS_FIELD_DW) return fail(LCS_ERR_INVALID, "no such field");
    if (!host) return fail(LCS_ERR_INVALID, "null host pointer");
    size_t need = (size_t)(field >= LCS_FIELD_DU ? b.g.cells() : b.g.len) * b.esize;
    if (bytes != need)
        return fail(LCS_ERR_INVALID, "field size mismatch: expected " + std::to_string(need) + " bytes, got " +
                                         std::to_string(bytes));
    if (field >= LCS_FIELD_DU) {
        int rc = ensure_dudt(b);
        if (rc) return rc;
    }
    LCS_HIP(hipSetDevice(b.ctx->device));
    if (up)
        LCS_HIP(hipMemcpyAsync(b.f[field], host, bytes, hipMemcpyHostToDevice, b.stream));
    else
        LCS_HIP(hipMemcpyAsync(host, b.f[field], bytes, hipMemcpyDeviceToHost, b.stream));
    LCS_HIP(hipStreamSynchronize(b.stream));
    return LCS_OK;
}

int lcs_field_upload(lcs_box *box, int field, const void *host, size_t bytes)
{
    return field_xfer(box, field, const_cast<void *>(host), bytes, true);
}
int lcs_field_download(lcs_box *box, int field, void *host, size_t bytes)
{
    return field_xfer(box, field, host, bytes, false);
}

int lcs_fill_ghost(lcs_box *box, unsigned mask, int width, int full)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    void *arr[4];
    int n = 0;
    for (int q = 0; q <= LCS_FIELD_P; q++)
        if (mask & (1u << q)) {
            if (n == 4) {
                int rc = launch_ghost_fill(b, arr, n, width, full, nullptr);
                if (rc) return rc;
                n = 0;
            }
            arr[n++] = b.f[q];
        }
    if (mask >> (LCS_FIELD_P + 1)) return fail(LCS_ERR_INVALID, "ghost fill applies to halo fields only");
    if (n) return launch_ghost_fill(b, arr, n, width, full, nullptr);
    return LCS_OK;
}

int lcs_flow_timestep(lcs_box *box, double cfl, double dtmax, double *dt, double *umax3)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    int rc = launch_cfl(b, cfl, dtmax);
    if (rc || (rc = read_scalars(b))) return rc;
    if (dt) *dt = b.S_host->dt_flow;
    if (umax3)
        for (int a = 0; a < 3; a++) umax3[a] = b.S_host->umax[a];
    return LCS_OK;
}

int lcs_rhs(lcs_box *box, int stage)
{
    if (check_box(box) || check_stage(stage)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    if (b.g.h < 3) return fail(LCS_ERR_INVALID, "the RHS needs halo_size >= 3");
    int rc = ensure_dudt(b);
    if (rc) return rc;
    void *out[3] = {b.f[LCS_FIELD_DU], b.f[LCS_FIELD_DV], b.f[LCS_FIELD_DW]};
    void *base[3] = {nullptr, nullptr, nullptr};
    return launch_rhs(b, stage, false, out, base, DtRef{&b.S->dt_phase, 1.0});
}

int lcs_update(lcs_box *box, int stage, double dt_stage)
{
    if (check_box(box) || check_stage(stage)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    int rc = ensure_dudt(b);
    if (rc || (rc = set_phase_dt(b, dt_stage))) return rc;
    void *const *out = stage == 1 ? &b.f[LCS_FIELD_U2] : &b.f[LCS_FIELD_U];
    void *const *base = &b.f[LCS_FIELD_U];
    return launch_update(b, out, base, DtRef{&b.S->dt_phase, 1.0});
}

int lcs_rhs_update(lcs_box *box, int stage, double dt_stage)
{
    if (check_box(box) || check_stage(stage)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    if (b.g.h < 3) return fail(LCS_ERR_INVALID, "the RHS needs halo_size >= 3");
    int rc = set_phase_dt(b, dt_stage);
    if (rc) return rc;
    void *const *out = stage == 1 ? &b.f[LCS_FIELD_U2] : &b.f[LCS_FIELD_U];
    return launch_rhs(b, stage, true, out, &b.f[LCS_FIELD_U], DtRef{&b.S->dt_phase, 1.0});
}

int lcs_hsmac(lcs_box *box, int stage, double dt_stage, const lcs_hsmac_params *prm, int *itr, double *div)
{
    if (check_box(box) || check_stage(stage)) return LCS_ERR_INVALID;
    if (!prm) return fail(LCS_ERR_INVALID, "null params");
    Box &b = *B(box);
    if (b.batch) return fail(LCS_ERR_STATE, "box is stepped by a batch");
    int rc = set_phase_dt(b, dt_stage);
    if (rc) return rc;
    void **Us = stage == 1 ? &b.f[LCS_FIELD_U2] : &b.f[LCS_FIELD_U];
    if ((rc = launch_hsmac(b, Us, DtRef{&b.S->dt_phase, 1.0}, *prm, 0))) return rc;
    if ((rc = read_scalars(b))) return rc;
    if (itr) *itr = b.S_host->itr[0];
    if (div) *div = b.S_host->div[0];
    return LCS_OK;
}

int lcs_force_linear(lcs_box *box, double power, double dt, double *p)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    int rc = set_phase_dt(b, dt);
    if (rc || (rc = launch_force_linear(b, power, DtRef{&b.S->dt_phase, 1.0})) || (rc = read_scalars(b))) return rc;
    if (p) *p = b.S_host->p;
    return LCS_OK;
}

int lcs_force_rcf(lcs_box *box, double energy, int kmin, int kmax, int c, double dt, double *p)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    int rc = set_phase_dt(b, dt);
    if (rc || (rc = launch_force_rcf(b, energy, kmin, kmax, c, DtRef{&b.S->dt_phase, 1.0})) ||
        (rc = read_scalars(b)))
        return rc;
    if (p) *p = b.S_host->p;
    return LCS_OK;
}

int lcs_force_rcf_noalloc(lcs_box *box, double energy, int kmin, int kmax, int c, double dt, double *p)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    if (c < 1 || kmax < 0) return fail(LCS_ERR_INVALID, "coarse_size must be >= 1 and kmax >= 0");
    const size_t need = rcf_scratch_doubles(c, kmax);
    if (need > kRcfScratchDoubles)
        return fail(LCS_ERR_CAPACITY, "lcs_force_rcf_noalloc: coarse_size " + std::to_string(c) + " and kmax " +
                                          std::to_string(kmax) + " need " + std::to_string(need) +
                                          " scratch doubles, more than the fixed " +
                                          std::to_string(kRcfScratchDoubles) + " (use lcs_force_rcf)");
    double *ptr = g_scratch_dev[b.ctx->device];
    if (!ptr) return fail(LCS_ERR_INVALID, "scratch of this device was never registered");
    int rc = set_phase_dt(b, dt);
    if (rc || (rc = launch_force_rcf(b, energy, kmin, kmax, c, DtRef{&b.S->dt_phase, 1.0}, ptr)) ||
        (rc = read_scalars(b)))
        return rc;
    if (p) *p = b.S_host->p;
    return LCS_OK;
}

// ---- tile shapes ---------------------------------------------------------------
static bool shape_ok(int esize, int kind, int tx, int ty)
{
    int c[8][2];
    const int n = tile_candidates(esize, kind, c, 8);
    for (int i = 0; i < n; i++)
        if (c[i][0] == tx && c[i][1] == ty) return true;
    return false;
}

// One time-blocked sweep group (three fused sweeps with their fills and
// checks) and one fused stage-1 RHS on scratch box t, with shapes s.
static int time_shapes(Box &t, const TileShapes &s, int coloring, float &sweep_ms, float &rhs_ms)
{
    t.tiles = s;
    // one time-blocked group (epsp 0: exactly itrmax sweeps); itrmin 3 lets its
    // d2 and d3 sweeps run as one pass where the shapes ask for it
    const lcs_hsmac_params p = {1, 0, coloring, 1.7, 0.0, 3, 3};
    void *out[3] = {t.f[LCS_FIELD_U2], t.f[LCS_FIELD_V2], t.f[LCS_FIELD_W2]};
    void *base[3] = {t.f[LCS_FIELD_U], t.f[LCS_FIELD_V], t.f[LCS_FIELD_W]};
    hipEvent_t e[3];
    for (auto &x : e) LCS_HIP(hipEventCreate(&x));
    int rc = LCS_OK;
    sweep_ms = rhs_ms = 1e30f;
    for (int rep = 0; rep < 4 && rc == LCS_OK; rep++) {  // rep 0 sets kernel attributes; best of the rest
        if (hipEventRecord(e[0], t.stream) != hipSuccess) rc = fail(LCS_ERR_HIP, "hipEventRecord failed");
        if (rc == LCS_OK) rc = launch_hsmac(t, &t.f[LCS_FIELD_U], DtRef{&t.S->dt_phase, 1.0}, p, -1);
        if (rc == LCS_OK && hipEventRecord(e[1], t.stream) != hipSuccess) rc = fail(LCS_ERR_HIP, "hipEventRecord failed");
        if (rc == LCS_OK) rc = launch_rhs(t, 1, true, out, base, DtRef{&t.S->dt_phase, 0.5});
        if (rc == LCS_OK && hipEventRecord(e[2], t.stream) != hipSuccess) rc = fail(LCS_ERR_HIP, "hipEventRecord failed");
        if (rc == LCS_OK && hipEventSynchronize(e[2]) != hipSuccess) rc = fail(LCS_ERR_HIP, "hipEventSynchronize failed");
        float a = 0, b = 0;
        if (rc == LCS_OK && (hipEventElapsedTime(&a, e[0], e[1]) != hipSuccess ||
                             hipEventElapsedTime(&b, e[1], e[2]) != hipSuccess))
            rc = fail(LCS_ERR_HIP, "hipEventElapsedTime failed");
        if (rc == LCS_OK && rep > 0) {
            sweep_ms = std::min(sweep_ms, a);
            rhs_ms = std::min(rhs_ms, b);
        }
    }
    for (auto &x : e) (void)hipEventDestroy(x);
    return rc;
}

struct TuneKey {
    int device, nx, ny, nz, h, esize;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(device, nx, ny, nz, h, esize) < std::tie(o.device, o.nx, o.ny, o.nz, o.h, o.esize);
    }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, TileShapes> g_tune_cache;

static void to_c(const TileShapes &t, lcs_tile_shapes *o)
{
    if (!o) return;
    o->sweep_tx = t.sweep_tx;
    o->sweep_ty = t.sweep_ty;
    o->rhs_tx = t.rhs_tx;
    o->rhs_ty = t.rhs_ty;
    o->sweep2_tx = t.sweep2_tx;
    o->sweep2_ty = t.sweep2_ty;
}

int lcs_autotune(lcs_box *box, lcs_tile_shapes *chosen)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    const TuneKey key{b.ctx->device, b.g.nx, b.g.ny, b.g.nz, b.g.h, b.esize};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            b.tiles = it->second;
            to_c(b.tiles, chosen);
            return LCS_OK;
        }
    }
    if (b.g.h < 3) {  // time-blocked sweeps need a 3-wide halo: keep the rule
        b.tiles = tiles_default(b.g, b.esize);
        to_c(b.tiles, chosen);
        return LCS_OK;
    }
    // a scratch box as deep as the box, up to 256 planes (the kernels' z-chunk
    // choice depends on the depth)
    lcs_box_desc d = b.desc;
    d.nz = std::min(d.nz, 256);
    lcs_box *tb = nullptr;
    if (lcs_box_create(reinterpret_cast<lcs_ctx *>(b.ctx), &d, &tb) != LCS_OK) {
        b.tiles = tiles_default(b.g, b.esize);  // no room for a scratch box: keep the rule
        to_c(b.tiles, chosen);
        return LCS_OK;
    }
    Box &t = *B(tb);
    int rc = set_phase_dt(t, 1e-3);
    TileShapes best = tiles_default(b.g, b.esize);
    int sw[8][2], rh[8][2];
    const int nsw = tile_candidates(b.esize, 0, sw, 8), nrh = tile_candidates(b.esize, 1, rh, 8);
    float best_sweep = 1e30f, best_rhs = 1e30f;
    // sweep shapes (single sweeps), then the two-sweep pass, then RHS shapes
    best.sweep2_tx = best.sweep2_ty = 0;
    for (int i = 0; i < nsw && rc == LCS_OK; i++) {
        TileShapes s = best;
        s.sweep_tx = sw[i][0];
        s.sweep_ty = sw[i][1];
        float ms_sw = 0, ms_rh = 0;
        rc = time_shapes(t, s, LCS_COLORING_RED_BLACK_FAST, ms_sw, ms_rh);
        if (rc == LCS_OK && ms_sw < best_sweep) {
            best_sweep = ms_sw;
            best.sweep_tx = s.sweep_tx;
            best.sweep_ty = s.sweep_ty;
        }
    }
    int s2[8][2];
    const int ns2 = tile_candidates(b.esize, 2, s2, 8);
    float best_group = 1e30f;
    for (int i = 0; i < ns2 && rc == LCS_OK; i++) {  // the two-sweep pass, or none
        TileShapes s = best;
        s.sweep2_tx = s2[i][0];
        s.sweep2_ty = s2[i][1];
        float ms_sw = 0, ms_rh = 0;
        rc = time_shapes(t, s, LCS_COLORING_RED_BLACK_FAST, ms_sw, ms_rh);
        if (rc == LCS_OK && ms_sw < best_group) {
            best_group = ms_sw;
            best.sweep2_tx = s.sweep2_tx;
            best.sweep2_ty = s.sweep2_ty;
        }
    }
    for (int i = 0; i < nrh && rc == LCS_OK; i++) {
        TileShapes s = best;
        s.rhs_tx = rh[i][0];
        s.rhs_ty = rh[i][1];
        float ms_sw = 0, ms_rh = 0;
        rc = time_shapes(t, s, LCS_COLORING_RED_BLACK_FAST, ms_sw, ms_rh);
        if (rc == LCS_OK && ms_rh < best_rhs) {
            best_rhs = ms_rh;
            best.rhs_tx = s.rhs_tx;
            best.rhs_ty = s.rhs_ty;
        }
    }
    const int drc = lcs_box_destroy(tb);
    if (rc) return rc;
    if (drc) return drc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = best;
    }
    b.tiles = best;
    to_c(best, chosen);
    return LCS_OK;
}

int lcs_set_tiles(lcs_box *box, const lcs_tile_shapes *sh)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    if (!sh) {
        b.tiles = TileShapes{};
        return LCS_OK;
    }
    if (!shape_ok(b.esize, 0, sh->sweep_tx, sh->sweep_ty)) return fail(LCS_ERR_INVALID, "unsupported sweep tile shape");
    if (!shape_ok(b.esize, 1, sh->rhs_tx, sh->rhs_ty)) return fail(LCS_ERR_INVALID, "unsupported RHS tile shape");
    if ((sh->sweep2_tx || sh->sweep2_ty) && !shape_ok(b.esize, 2, sh->sweep2_tx, sh->sweep2_ty))
        return fail(LCS_ERR_INVALID, "unsupported two-sweep tile shape");
    b.tiles.sweep_tx = sh->sweep_tx;
    b.tiles.sweep_ty = sh->sweep_ty;
    b.tiles.rhs_tx = sh->rhs_tx;
    b.tiles.rhs_ty = sh->rhs_ty;
    b.tiles.sweep2_tx = sh->sweep2_tx;
    b.tiles.sweep2_ty = sh->sweep2_ty;
    return LCS_OK;
}

int lcs_get_tiles(lcs_box *box, lcs_tile_shapes *sh)
{
    if (check_box(box) || !sh) return fail(LCS_ERR_INVALID, "null argument");
    Box &b = *B(box);
    to_c(b.tiles.sweep_tx ? b.tiles : tiles_default(b.g, b.esize), sh);
    return LCS_OK;
}

int lcs_flow_stat(lcs_box *box, int integral_length, lcs_flow_stats *out)
{
    if (check_box(box) || !out) return fail(LCS_ERR_INVALID, "null argument");
    Box &b = *B(box);
    double *dev = b.part + 16L * kMaxPartBlocks;  // 15 results after the partial slots the passes use
    int rc = launch_flow_stat(b, integral_length, dev);
    if (rc) return rc;
    double r[15] = {0};
    LCS_HIP(hipMemcpyAsync(r, dev, sizeof(double) * (integral_length ? 15 : 14), hipMemcpyDeviceToHost, b.stream));
    LCS_HIP(hipStreamSynchronize(b.stream));
    // derived quantities: src/flows/stat.jl:132-154
    lcs_flow_stats &o = *out;
    o.u_mean = r[0], o.u_max = r[1], o.u_min = r[2];
    o.v_mean = r[3], o.v_max = r[4], o.v_min = r[5];
    o.w_mean = r[6], o.w_max = r[7], o.w_min = r[8];
    const double u2 = r[9], s2 = r[10], d2 = r[11], d3 = r[12], d4 = r[13], Re = b.Re();
    o.urms = std::sqrt(u2 / 3);
    o.k = 0.5 * (u2 - o.u_mean * o.u_mean - o.v_mean * o.v_mean - o.w_mean * o.w_mean);
    o.eps = s2 / Re;
    o.tau_eta = std::sqrt(1 / o.eps / Re);
    o.l_mu = std::sqrt(15 * o.urms * o.urms / o.eps / Re);
    o.l_eta = std::pow(1 / o.eps / (Re * Re * Re), 0.25);
    o.Re_lambda = o.l_mu * o.urms * Re;
    o.k_eta = (std::min(b.g.nx, std::min(b.g.ny, b.g.nz)) / 2.0) * o.l_eta;
    o.skew = -d3 / std::pow(d2, 1.5);
    o.flat = d4 / (d2 * d2);
    o.L = integral_length ? r[14] : 0.0;
    o.C_eps = o.eps * o.L / (o.urms * o.urms * o.urms);
    return LCS_OK;
}

int lcs_boxmean(lcs_box *box, int c, double *coarse3)
{
    if (check_box(box) || !coarse3) return fail(LCS_ERR_INVALID, "null argument");
    Box &b = *B(box);
    double *dev = nullptr;
    size_t bytes = sizeof(double) * 3 * (size_t)c * c * c;
    LCS_HIP(hipMalloc(&dev, bytes));
    int rc = launch_boxmean(b, c, dev);
    if (rc == LCS_OK) {
        hipError_t e = hipMemcpyAsync(coarse3, dev, bytes, hipMemcpyDeviceToHost, b.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b.stream);
        if (e != hipSuccess) rc = fail(LCS_ERR_HIP, hipGetErrorString(e));
    }
    (void)hipFree(dev);
    return rc;
}

// ---- particles ----------------------------------------------------------------
int lcs_particles_add(lcs_box *box, const lcs_particle_desc *d, int *profile)
{
    if (check_box(box) || !d) return fail(LCS_ERR_INVALID, "null argument");
    Box &b = *B(box);
    if ((int)b.profiles.size() >= kMaxProfiles) return fail(LCS_ERR_CAPACITY, "at most 8 particle profiles");
    if (d->nvalid < 0 || d->ntotal < d->nvalid) return fail(LCS_ERR_INVALID, "need 0 <= nvalid <= ntotal");
    if (d->ntotal >= (int64_t(1) << 31)) return fail(LCS_ERR_CAPACITY, "at most 2^31-1 particles per profile");
    if (d->grid_ratio < 1) return fail(LCS_ERR_INVALID, "grid_ratio must be >= 1");
    if (d->drag != LCS_DRAG_LINEAR && d->drag != LCS_DRAG_NONLINEAR) return fail(LCS_ERR_INVALID, "unknown drag");
    ParticleProfile pp;
    memset(pp.props, 0, sizeof(pp.props));
    pp.desc = *d;
    pp.nc[0] = b.g.nx / d->grid_ratio;
    pp.nc[1] = b.g.ny / d->grid_ratio;
    pp.nc[2] = b.g.nz / d->grid_ratio;
    long ncell = (long)(pp.nc[0] + 2) * (pp.nc[1] + 2) * (pp.nc[2] + 2);
    if (ncell > 4294967295L) return fail(LCS_ERR_INVALID, "cell count exceeds typemax(UInt32)");
    size_t nt = (size_t)std::max<int64_t>(1, d->ntotal);
    int rc = LCS_OK;
    for (int q = 0; q < LCS_N_PROPERTIES && rc == LCS_OK; q++) {
        size_t es = q == LCS_PROP_ID ? 8 : b.esize;
        rc = device_alloc(b, &pp.props[q], es * nt);
        if (rc == LCS_OK && q != LCS_PROP_ID && hipMemsetAsync(pp.props[q], 0, es * nt, b.stream) != hipSuccess)
            rc = fail(LCS_ERR_HIP, "hipMemset failed");
    }
    if (rc == LCS_OK) rc = device_alloc(b, &pp.hashes, 8 * nt);
    if (rc == LCS_OK) rc = device_alloc(b, &pp.perm, 4 * nt);
    if (rc == LCS_OK) rc = device_alloc(b, &pp.starts, 8 * (size_t)ncell);
    if (rc == LCS_OK) rc = device_alloc(b, &pp.stops, 8 * (size_t)ncell);
    if (rc == LCS_OK) rc = device_alloc(b, &pp.scratch, 8 * nt);
    if (rc == LCS_OK) rc = device_alloc(b, &pp.scratch_t, (size_t)b.esize * nt);
    if (rc == LCS_OK) {
        // ids = INVALID (src/particles/buffer.jl:19-20)
        std::vector<int64_t> inval(nt, kInvalidId);
        hipError_t e = hipMemcpyAsync(pp.props[LCS_PROP_ID], inval.data(), 8 * nt, hipMemcpyHostToDevice, b.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b.stream);
        if (e != hipSuccess) rc = fail(LCS_ERR_HIP, hipGetErrorString(e));
    }
    if (rc != LCS_OK) {
        for (void *p : pp.props)
            if (p) (void)hipFree(p);
        void *extra[] = {pp.hashes, pp.perm, pp.starts, pp.stops, pp.scratch, pp.scratch_t};
        for (void *p : extra)
            if (p) (void)hipFree(p);
        return rc;
    }
    b.profiles.push_back(pp);
    batch_mark_dirty(b);
    if (profile) *profile = (int)b.profiles.size() - 1;
    return LCS_OK;
}

static int prop_xfer(lcs_box *box, int profile, int prop, void *host, size_t bytes, bool up)
{
    if (check_box(box) || !host) return fail(LCS_ERR_INVALID, "null argument");
    Box &b = *B(box);
    ParticleProfile *pp;
    int rc = profile_of(b, profile, &pp);
    if (rc) return rc;
    if (prop < 0 || prop >= LCS_N_PROPERTIES) return fail(LCS_ERR_INVALID, "no such property");
    size_t es = prop == LCS_PROP_ID ? 8 : b.esize;
    size_t need = es * (size_t)pp->desc.ntotal;
    if (bytes != need)
        return fail(LCS_ERR_INVALID, "property size mismatch: expected " + std::to_string(need) + " bytes");
    if (need == 0) return LCS_OK;
    if (!pp->loc_on) {
        if (up)
            LCS_HIP(hipMemcpyAsync(pp->props[prop], host, bytes, hipMemcpyHostToDevice, b.stream));
        else
            LCS_HIP(hipMemcpyAsync(host, pp->props[prop], bytes, hipMemcpyDeviceToHost, b.stream));
    } else if (up) {  // logical order on the host, storage order on the device
        LCS_HIP(hipMemcpyAsync(pp->scratch, host, bytes, hipMemcpyHostToDevice, b.stream));
        if ((rc = launch_p_loc_move(b, *pp, pp->scratch, pp->props[prop], es, true))) return rc;
    } else {
        if ((rc = launch_p_loc_move(b, *pp, pp->props[prop], pp->scratch, es, false))) return rc;
        LCS_HIP(hipMemcpyAsync(host, pp->scratch, bytes, hipMemcpyDeviceToHost, b.stream));
    }
    LCS_HIP(hipStreamSynchronize(b.stream));
    return LCS_OK;
}

int lcs_particles_describe(const lcs_box *box, int profile, lcs_particle_desc *desc)
{
    if (check_box(box) || !desc) return fail(LCS_ERR_INVALID, "null argument");
    const Box &b = *B(box);
    if (profile < 0 || profile >= (int)b.profiles.size()) return fail(LCS_ERR_STATE, "no such particle profile");
    *desc = b.profiles[profile].desc;
    return LCS_OK;
}

int lcs_particles_upload(lcs_box *box, int profile, int prop, const void *host, size_t bytes)
{
    return prop_xfer(box, profile, prop, const_cast<void *>(host), bytes, true);
}
int lcs_particles_download(lcs_box *box, int profile, int prop, void *host, size_t bytes)
{
    return prop_xfer(box, profile, prop, host, bytes, false);
}

#define WITH_PROFILE                                            \
    if (check_box(box)) return LCS_ERR_INVALID;                 \
    Box &b = *B(box);                                           \
    ParticleProfile *pp;                                        \
    int rc = profile_of(b, profile, &pp);                       \
    if (rc) return rc;

int lcs_particles_boundary(lcs_box *box, int profile, int stage)
{
    WITH_PROFILE
    if (check_stage(stage)) return LCS_ERR_INVALID;
    return launch_p_boundary(b, *pp, stage);
}

int lcs_particles_motion(lcs_box *box, int profile, int stage)
{
    WITH_PROFILE
    if (check_stage(stage)) return LCS_ERR_INVALID;
    return launch_p_motion(b, *pp, stage);
}

int lcs_particles_update(lcs_box *box, int profile, int stage, double dt_stage)
{
    WITH_PROFILE
    if (check_stage(stage)) return LCS_ERR_INVALID;
    if ((rc = set_phase_dt(b, dt_stage))) return rc;
    return launch_p_update(b, *pp, stage, DtRef{&b.S->dt_phase, 1.0});
}

int lcs_particles_timestep(lcs_box *box, int profile, double dtmax, double *dt)
{
    WITH_PROFILE
    if ((rc = launch_p_timestep(b, *pp, dtmax, &b.S->dtp[profile])) || (rc = read_scalars(b))) return rc;
    if (dt) *dt = b.S_host->dtp[profile];
    return LCS_OK;
}

int lcs_particles_makeindex(lcs_box *box, int profile, int stage)
{
    WITH_PROFILE
    if (check_stage(stage)) return LCS_ERR_INVALID;
    batch_mark_dirty(b);
    return launch_p_makeindex(b, *pp, stage);
}

int lcs_particles_localize(lcs_box *box, int profile)
{
    WITH_PROFILE
    batch_mark_dirty(b);
    return launch_p_localize(b, *pp);
}

int lcs_particles_cellindex(lcs_box *box, int profile, int64_t *hashes, int64_t *starts, int64_t *stops,
                            int64_t *ncells)
{
    WITH_PROFILE
    if (!pp->indexed) return fail(LCS_ERR_STATE, "makeindex has not run for this profile");
    long nc = (long)(pp->nc[0] + 2) * (pp->nc[1] + 2) * (pp->nc[2] + 2);
    if (ncells) *ncells = nc;
    if (hashes && pp->desc.nvalid)
        LCS_HIP(hipMemcpyAsync(hashes, pp->hashes, 8 * (size_t)pp->desc.nvalid, hipMemcpyDeviceToHost, b.stream));
    if (starts) LCS_HIP(hipMemcpyAsync(starts, pp->starts, 8 * (size_t)nc, hipMemcpyDeviceToHost, b.stream));
    if (stops) LCS_HIP(hipMemcpyAsync(stops, pp->stops, 8 * (size_t)nc, hipMemcpyDeviceToHost, b.stream));
    LCS_HIP(hipStreamSynchronize(b.stream));
    return LCS_OK;
}

int lcs_particle_stat(lcs_box *box, int profile, const lcs_particle_stat_params *params, double tau_eta,
                      lcs_particle_stats *out, double *edges, int64_t *npairs, double *gr, double *wavenumbers,
                      double *Enp)
{
    WITH_PROFILE
    if (!params || !out) return fail(LCS_ERR_INVALID, "null argument");
    batch_mark_dirty(b);  // re-binning moves the properties
    return particle_stat(b, *pp, *params, tau_eta, *out, edges, npairs, gr, wavenumbers, Enp);
}

int lcs_particle_density(lcs_box *box, int profile, double *density)
{
    WITH_PROFILE
    if (!density && pp->desc.nvalid > 0) return fail(LCS_ERR_INVALID, "null argument");
    batch_mark_dirty(b);  // re-binning moves the properties
    return particle_density(b, *pp, density);
}

int lcs_init_spectrum(lcs_box *box, uint64_t seed, double k0, double uprime)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    if (!(k0 > 0) || !(uprime > 0)) return fail(LCS_ERR_INVALID, "k0 and uprime must be positive");
    return launch_init_spectrum(*B(box), seed, k0, uprime);
}

int lcs_init_particles(lcs_box *box, int profile, uint64_t seed, double st_min, double st_max, double eps,
                       int zero_slip)
{
    WITH_PROFILE
    if (!(st_min > 0) || !(st_max >= st_min) || !(eps > 0))
        return fail(LCS_ERR_INVALID, "need 0 < st_min <= st_max and eps > 0");
    pp->loc_on = false;  // every property is rewritten in logical order
    rc = launch_p_init(b, *pp, seed, st_min, st_max, eps, zero_slip);
    if (rc) return rc;
    LCS_HIP(hipStreamSynchronize(b.stream));
    return LCS_OK;
}

// ---- whole step -------------------------------------------------------------
static void fill_log(const DevScalars &s, lcs_step_log *log)
{
    log->step = s.step;
    log->t = s.t;
    log->dt = s.dt;
    log->itr1 = s.itr[0];
    log->itr2 = s.itr[1];
    log->div1 = s.div[0];
    log->div2 = s.div[1];
    log->p = s.p;
    log->e10 = s.e10;
}

int lcs_step(lcs_box *box, const lcs_step_config *cfg, lcs_step_log *log)
{
    if (check_box(box) || !cfg) return fail(LCS_ERR_INVALID, "null argument");
    Box &b = *B(box);
    if (b.batch) return fail(LCS_ERR_STATE, "box is stepped by a batch");
    if (b.g.h < 3) return fail(LCS_ERR_INVALID, "a flow step needs halo_size >= 3");
    if (cfg->localize_interval < 0) return fail(LCS_ERR_INVALID, "localize_interval must be >= 0");
    int rc;
    int np = (int)b.profiles.size();
    if (cfg->localize_interval > 0 && b.steps_host % cfg->localize_interval == 0)
        for (int q = 0; q < np; q++)
            if ((rc = launch_p_localize(b, b.profiles[q]))) return rc;
    b.steps_host++;
    // stage 1's FullHalo fill of U, V, W, P (evolve.jl:74-81) comes first: it
    // writes every ghost the CFL's FaceHalo fill would, with the same values,
    // and nothing writes the velocity in between, so the CFL reads it instead
    // HSMAC may swap field pointers with its ping-pong partners: read b.f at use
    void *UsP[4] = {b.f[LCS_FIELD_U], b.f[LCS_FIELD_V], b.f[LCS_FIELD_W], b.f[LCS_FIELD_P]};
    if ((rc = launch_ghost_fill(b, UsP, 4, b.g.h, 1, nullptr))) return rc;
    // timestep!: src/timestep.jl:1-27
    if ((rc = launch_cfl(b, cfg->cfl, cfg->dtmax, false))) return rc;
    for (int q = 0; q < np; q++)
        if ((rc = launch_p_timestep(b, b.profiles[q], cfg->dtmax, &b.S->dtp[q]))) return rc;
    if ((rc = launch_finalize_dt(b, np, cfg->dtmax))) return rc;

    // evolve!: src/evolve.jl:63-146 (FlowMode :15-61 when there are no particles)
    const DtRef dt1{&b.S->dt, 0.5}, dt2{&b.S->dt, 1.0};
    // stage 1 (ghosts filled above)
    for (int q = 0; q < np; q++)
        if ((rc = launch_p_stage(b, b.profiles[q], 1, dt1))) return rc;
    if ((rc = launch_rhs(b, 1, true, &b.f[LCS_FIELD_U2], &b.f[LCS_FIELD_U], dt1))) return rc;
    if ((rc = launch_hsmac(b, &b.f[LCS_FIELD_U2], dt1, cfg->hsmac, 0))) return rc;
    // stage 2
    void *Us2P[4] = {b.f[LCS_FIELD_U2], b.f[LCS_FIELD_V2], b.f[LCS_FIELD_W2], b.f[LCS_FIELD_P]};
    if ((rc = launch_ghost_fill(b, Us2P, 4, b.g.h, 1, nullptr))) return rc;
    for (int q = 0; q < np; q++)
        if ((rc = launch_p_stage(b, b.profiles[q], 2, dt2))) return rc;
    if ((rc = launch_rhs(b, 2, true, &b.f[LCS_FIELD_U], &b.f[LCS_FIELD_U], dt2))) return rc;
    const bool lf_defer = cfg->forcing.kind == LCS_FORCING_LINEAR_POWER && cfg->hsmac.itrmax >= 1;
    switch (cfg->forcing.kind) {
    case LCS_FORCING_NONE:
        break;
    case LCS_FORCING_LINEAR_POWER:
        // forcing! then couple!: the first HSMAC sweep applies the forcing as
        // it loads the velocity, saving a read and write of U, V, W
        if ((rc = lf_defer ? launch_force_linear_coef(b, cfg->forcing.power, hsmac_group_wraps(b.g, cfg->hsmac))
                           : launch_force_linear(b, cfg->forcing.power, dt2)))
            return rc;
        break;
    case LCS_FORCING_RC_ENERGY:
        if ((rc = launch_force_rcf(b, cfg->forcing.energy, cfg->forcing.kmin, cfg->forcing.kmax,
                                   cfg->forcing.coarse_size, dt2)))
            return rc;
        break;
    default:
        return fail(LCS_ERR_INVALID, "unknown forcing kind");
    }
    if ((rc = launch_hsmac(b, &b.f[LCS_FIELD_U], dt2, cfg->hsmac, 1, lf_defer))) return rc;
    hipLaunchKernelGGL(k_advance_state, dim3(1), dim3(1), 0, b.stream, b.S);
    LCS_HIP(hipGetLastError());
    if (log) {
        if ((rc = read_scalars(b))) return rc;
        fill_log(*b.S_host, log);
    }
    return LCS_OK;
}

int lcs_set_state(lcs_box *box, int64_t step, double t)
{
    if (check_box(box)) return fail(LCS_ERR_INVALID, "null argument");
    Box &b = *B(box);
    if (step < 0) return fail(LCS_ERR_INVALID, "step must be >= 0");
    LCS_HIP(hipSetDevice(b.ctx->device));
    hipLaunchKernelGGL(k_set_state, dim3(1), dim3(1), 0, b.stream, b.S, step, t);
    LCS_HIP(hipGetLastError());
    LCS_HIP(hipStreamSynchronize(b.stream));
    return LCS_OK;
}

int lcs_step_log_read(lcs_box *box, lcs_step_log *log)
{
    if (check_box(box) || !log) return fail(LCS_ERR_INVALID, "null argument");
    Box &b = *B(box);
    int rc = read_scalars(b);
    if (rc) return rc;
    fill_log(*b.S_host, log);
    return LCS_OK;
}

int lcs_profile(lcs_box *box, int enable)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    LCS_HIP(hipStreamSynchronize(b.stream));
    prof_collect(b);
    b.prof.entries.clear();
    b.prof.on = enable != 0;
    return LCS_OK;
}

int lcs_profile_read(lcs_box *box, int index, char *name, size_t name_len, double *total_ms, int64_t *launches,
                     double *bytes)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    Box &b = *B(box);
    LCS_HIP(hipStreamSynchronize(b.stream));
    prof_collect(b);
    if (index < 0 || index >= (int)b.prof.entries.size()) return fail(LCS_ERR_INVALID, "no such profile entry");
    const ProfEntry &e = b.prof.entries[index];
    if (name && name_len) {
        size_t n = std::min(name_len - 1, e.name.size());
        memcpy(name, e.name.data(), n);
        name[n] = 0;
    }
    if (total_ms) *total_ms = e.ms;
    if (launches) *launches = e.launches;
    if (bytes) *bytes = e.bytes;
    return LCS_OK;
}

int lcs_sync(lcs_box *box)
{
    if (check_box(box)) return LCS_ERR_INVALID;
    LCS_HIP(hipStreamSynchronize(B(box)->stream));
    return LCS_OK;
}

}  // extern "C"
