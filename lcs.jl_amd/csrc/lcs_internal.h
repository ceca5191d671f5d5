// Internal declarations of the MI355X backend: box state, device scalar
// block, kernel launchers.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lcs_mi355x.h"

namespace lcs {

// 2π and origin: src/LCS.jl:16-17
constexpr double kDomainLength = 6.283185307179586;
constexpr double kDomainOrigin = 0.0;
// src/particles/Particles.jl:18
constexpr int64_t kInvalidId = (int64_t(1) << 53) - 1;

// Index geometry of one halo array.  Domain index (i,j,k), 1-based, maps to
// linear offset (i-1+h) + sx*((j-1+h) + sy*(k-1+h)).
struct Geom {
    int nx, ny, nz, h;
    long sx, sy;  // sx = nx+2h, sy = ny+2h
    long len;     // elements per halo array
    __host__ __device__ long at(int i, int j, int k) const
    {
        return (long)(i + h - 1) + sx * ((long)(j + h - 1) + sy * (long)(k + h - 1));
    }
    __host__ __device__ long interior(int i, int j, int k) const
    {
        return (long)(i - 1) + (long)nx * ((long)(j - 1) + (long)ny * (long)(k - 1));
    }
    __host__ __device__ long cells() const { return (long)nx * ny * nz; }
};

inline Geom make_geom(int nx, int ny, int nz, int h)
{
    Geom g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.h = h;
    g.sx = nx + 2 * h;
    g.sy = ny + 2 * h;
    g.len = g.sx * g.sy * (long)(nz + 2 * h);
    return g;
}

// Per-box scalars that stay on the device so a whole step can be enqueued
// without host round trips.  Written and read only by kernels and by
// explicit host copies.
struct DevScalars {
    double dt;        // state.dt (src/state.jl:15-38)
    double dt_phase;  // explicit stage dt of a per-phase call
    double t;
    int64_t step;
    double umax[3];
    double dt_flow;
    double dt_part;
    // HSMAC call in flight
    int hs_done;
    int hs_itr;
    int hs_par;  // 1 when the current HSMAC state lives in the spare buffers
    int hs_pad;
    double hs_div;
    // log of the current step
    int itr[2];
    double div[2];
    // forcing
    double lf[5];  // <u2>, <U>, <V>, <W>, A
    double p;
    double e10;
    double rcf_en[8];
    double dtp[8];  // particle dt per profile
};
constexpr int kMaxProfiles = 8;

// Reduction partial slots (doubles) available to one box.
constexpr int kMaxPartBlocks = 16384;    // blocks per reducing pass
constexpr int kMaxSweepPasses = 14;      // core 2 + 6 faces x 2 colours
constexpr long kPartSlots = (long)kMaxPartBlocks * (kMaxSweepPasses + 4);
// the last partial slot: per-workgroup linear-forcing power of the sweep that
// applies the forcing on load (no other pass of a step uses it)
constexpr long kPowerSlot = (long)(kMaxSweepPasses + 3) * kMaxPartBlocks;

struct ParticleProfile {
    lcs_particle_desc desc;
    int nc[3];  // cell dims (n / grid_ratio)
    void *props[LCS_N_PROPERTIES];  // [0] int64 ids, others real
    void *hashes = nullptr;         // int64 ntotal
    void *perm = nullptr;           // int32 ntotal
    void *starts = nullptr;         // int64 (nc+2)^3
    void *stops = nullptr;
    void *scratch = nullptr;        // 8 bytes * ntotal: spare of the int64 ids, and transfer staging
    void *scratch_t = nullptr;      // esize * ntotal: spare of the float properties
    void *sort_tmp = nullptr;       // keys/values ping-pong + histograms
    size_t sort_tmp_bytes = 0;
    bool indexed = false;
    // Storage order.  Logical particle i (the order the C ABI presents, which
    // is the reference's array order) lives in slot loc[i] when loc_on; the
    // identity otherwise.  lcs_particles_localize reorders slots by position.
    void *loc = nullptr;      // uint32 ntotal
    void *loc_tmp = nullptr;  // uint32 ntotal
    bool loc_on = false;
};

// The spare buffer a property is permuted into before the two are swapped:
// the spare always has that property's element size, so swaps never mix sizes.
inline void *&prop_spare(ParticleProfile &pp, int prop) { return prop == LCS_PROP_ID ? pp.scratch : pp.scratch_t; }

struct Ctx {
    int device;
};

// Live kernel timing: HIP events recorded on the box stream around each
// launch of a named kernel, plus the launch's algorithmic HBM bytes.
struct ProfEntry {
    std::string name;
    double bytes = 0;
    int64_t launches = 0;
    double ms = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};
struct Profiler {
    bool on = false;
    std::vector<ProfEntry> entries;
    std::vector<hipEvent_t> pool;
};

struct Batch;

// Tile shapes of the fused RHS and HSMAC sweep kernels of one box.  Results
// never depend on them (every shape computes the same values); 0 means the
// size rule of tiles_default.  lcs_autotune picks them by timing.
struct TileShapes {
    int sweep_tx = 0, sweep_ty = 0, rhs_tx = 0, rhs_ty = 0;
    int sweep2_tx = 0, sweep2_ty = 0;  // two-sweep pass; 0: two single sweeps
};
TileShapes tiles_default(const Geom &g, int esize);
// candidate (tx, ty) pairs: kind 0 sweep, 1 RHS, 2 two-sweep pass; returns the count
int tile_candidates(int esize, int kind, int out[][2], int cap);

struct Box {
    Ctx *ctx = nullptr;
    lcs_box_desc desc;
    Geom g;
    int esize;  // 4 or 8
    int cus = 256;  // compute units of the device
    hipStream_t stream = nullptr;
    void *f[10] = {};  // LCS_FIELD_* device pointers (DU..DW lazily allocated)
    void *spare[4] = {};  // ping-pong partners of (U,V,W,P) for fused HSMAC sweeps
    DevScalars *S = nullptr;  // device scalar block
    DevScalars *S_host = nullptr;  // pinned mirror for readback
    double *part = nullptr;  // reduction partials (kPartSlots)
    void *xf = nullptr;      // compact buffer of the x-slab HSMAC passes (lazily allocated)
    size_t xf_bytes = 0;
    double *rcf = nullptr;   // coarse grids + cropped spectra
    size_t rcf_bytes = 0;
    void *rcf_modes = nullptr;       // forced-mode table on the device (batched forcing)
    int rcf_key[3] = {0, 0, 0};      // its kmin, kmax, coarse size
    std::vector<ParticleProfile> profiles;
    Profiler prof;
    size_t bytes = 0;
    int64_t steps_host = 0;  // steps taken by this box (lcs_step or a batch)
    Batch *batch = nullptr;  // the batch that steps this box, if any
    TileShapes tiles;        // fused-kernel tile shapes (0: size rule)
    double Re() const { return desc.U0_m_s * desc.L0_m / desc.NU_m2_s; }
};

// ---- error handling --------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define LCS_HIP(expr)                                                                       \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return ::lcs::fail(LCS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Records start/stop events around the launches in its scope when the box
// profiler is on (lcs_profile); a no-op otherwise.
struct ProfScope {
    Box &b;
    int idx = -1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ProfScope(Box &box, const char *name, double bytes);
    ~ProfScope();
};
// Takes back launches (and their algorithmic bytes) of a named kernel that
// returned without work, e.g. HSMAC sweeps enqueued after convergence; their
// time stays in the total.
void prof_uncount(Box &b, const char *name, int64_t launches, double bytes);

// ---- launchers (lcs_flow.hip) ----------------------------------------------
// Where a kernel needs the stage dt it reads factor * (*dtsrc) on the device.
struct DtRef {
    const double *src;
    double factor;
};

int launch_ghost_fill(Box &b, void *const *arrays, int narrays, int width, int full,
                      const int *skip_flag);
int launch_rhs(Box &b, int stage_in, bool fused, void *const out[3], void *const base[3], DtRef dt);
int launch_update(Box &b, void *const out[3], void *const base[3], DtRef dt);
// fill = false: the ghosts of U, V, W are already periodic copies (a FullHalo
// fill since the last write), so the FaceHalo fill is skipped
int launch_cfl(Box &b, double cfl, double dtmax, bool fill = true);  // writes S->umax, S->dt_flow
// Us points at the box's three velocity pointers of the stage; when the
// result ends up in the spare buffers these pointers (and P) are swapped.
// lf_deferred: the linear forcing whose coefficients launch_force_linear_coef
// left in S->lf is applied by the call's first sweep as it loads U, V, W
// (the same values k_lf_apply writes), which also sets S->p.
int launch_hsmac(Box &b, void **Us, DtRef dt, const lcs_hsmac_params &prm, int log_slot, bool lf_deferred = false);
int launch_force_linear(Box &b, double power, DtRef dt);  // writes S->lf, S->p
// writes S->lf only; face_only: the means read the x = 0, y = 0, z = 0
// ghosts alone, so a width-1 FaceHalo fill of U, V, W serves them where the
// next sweep loads every other ghost from its periodic image
int launch_force_linear_coef(Box &b, double power, bool face_only = false);
// a time-blocked HSMAC group's first sweep loads ghost U, V, W from their
// periodic images instead of a FullHalo fill (every ghost has one image)
bool hsmac_group_wraps(const Geom &g, const lcs_hsmac_params &p);
int launch_boxmean(Box &b, int c, double *coarse_dev);
// Flows.stat raw values into out_dev[15] (device): U, V, W mean/max/min, means of
// u2, s2, dudx^2..4, then the integral length when asked
int launch_flow_stat(Box &b, int integral_length, double *out_dev);
// Particles.stat / Particles.density of one profile (re-bin first)
int particle_stat(Box &b, ParticleProfile &pp, const lcs_particle_stat_params &p, double tau_eta,
                  lcs_particle_stats &o, double *edges, int64_t *npairs, double *gr, double *wavenumbers, double *Enp);
int particle_density(Box &b, ParticleProfile &pp, double *out);
// scratch: coarse grids, spectra and crop (rcf_scratch_doubles(c, kmax)
// doubles); nullptr uses the box's own buffer, allocated on first use
int launch_force_rcf(Box &b, double energy, int kmin, int kmax, int c, DtRef dt, double *scratch = nullptr);
size_t rcf_scratch_doubles(int c, int kmax);
int launch_init_spectrum(Box &b, uint64_t seed, double k0, double uprime);

// ---- launchers (lcs_particles.hip) -----------------------------------------
int launch_p_boundary(Box &b, ParticleProfile &pp, int stage);
int launch_p_motion(Box &b, ParticleProfile &pp, int stage);
int launch_p_update(Box &b, ParticleProfile &pp, int stage, DtRef dt);
int launch_p_stage(Box &b, ParticleProfile &pp, int stage, DtRef dt);
int launch_p_timestep(Box &b, ParticleProfile &pp, double dtmax, double *dtp_dev);
int launch_p_makeindex(Box &b, ParticleProfile &pp, int stage);
int launch_p_init(Box &b, ParticleProfile &pp, uint64_t seed, double st_min, double st_max,
                  double eps, int zero_slip);
int launch_finalize_dt(Box &b, int nprofiles, double dtmax);
int launch_p_localize(Box &b, ParticleProfile &pp);
// element moves through the storage map: dst[loc[i]] = src[i] (scatter) or
// dst[i] = src[loc[i]] (gather), i < ntotal, elements of esize bytes
int launch_p_loc_move(Box &b, ParticleProfile &pp, const void *src, void *dst, size_t esize, bool scatter);
// puts every property back into logical order and switches the map off
int launch_p_unlocalize(Box &b, ParticleProfile &pp);

// ---- batches of boxes (lcs_batch_*) ----------------------------------------
// A batch advances many independent boxes with one launch per phase.  Every
// batched kernel reads its box's pointers, sizes and coefficients from a
// device-resident BoxRec at run time, so the HSMAC ping-pong commit (a device
// kernel) can swap buffers without the host and a captured step graph stays
// valid.  Coefficients are stored as doubles holding the values rounded to
// the field type on the host, exactly as the single-box launchers form them.
struct DevProfile {
    long nvalid, ntotal;
    void *props[LCS_N_PROPERTIES];
    double mdx, mdy, mdz, mRe, mg, mkap;  // MotionCoef (field-type values)
    double kap;                           // taup/d^2 in double (particle dt)
    int drag;
    int blk0, nblk;    // particle blocks of this profile in the box's p_stage work (nvalid)
    int dblk0, dnblk;  // blocks of this profile in the box's max-diameter work (ntotal)
};
struct BoxRec {
    Geom g;
    void *f[7];      // U V W U2 V2 W2 P, where they currently live
    void *spare[4];  // ping-pong partners
    DevScalars *S;
    double *part;
    double Re, power, dx, dy, dz;
    double rhs_di[3], rhs_di2[3], rhs_Rei;
    double hs_dxi, hs_dyi, hs_dzi, hs_dl, hs_omega;
    int rhs_zc, hs_zc, hs_tx, hs_ty;
    int hs_tiles;        // workgroups of one fused sweep
    int face_blocks[3];  // blocks of one paired face pass, axes z, y, x
    void *xf;            // compact buffer of the x-slab pair (Box::xf), when the batch uses it
    int row_blocks;      // blocks of the row-walking reductions (CFL, forcing)
    int lf_force;        // linear-power forcing: applied by the first stage-2 sweep
    int nprof;
    // rc-energy forcing of the batch (rcf_c = 0: none): parameters, the
    // box's coarse-grid and spectrum buffer, its forced-mode table, and the
    // blocks of the apply pass (the single-box launch's grid)
    double *rcf_buf;
    const void *rcf_modes;
    double rcf_energy;
    int rcf_c, rcf_kmin, rcf_kmax, rcf_blocks;
    // the batch's stop flag (one int for all records): while it is nonzero
    // every batched kernel returns at once (b_halted), so the rest of a
    // captured step graph is skipped once a stage needs more HSMAC groups
    // than the graph holds
    int *halt;
    DevProfile prof[kMaxProfiles];
};
#if defined(__HIPCC__)
__device__ inline bool b_halted(const BoxRec &B) { return *(volatile const int *)B.halt != 0; }
#endif
struct WorkItem {
    int box, blk;
};
struct TileItem {
    int box, bx, by, bz;
};
// per-box HSMAC state read back by the host after each poll
struct HsPoll {
    int done, par, itr, pad;
    double div;
};
constexpr int kBatchRowBlocks = 64;
// fused RHS tile shapes of the batched kernels, chosen per size class
// (lcs_batch_autotune, lcs_batch_set_rhs_tiles); shape 0, the default, is
// narrower than the single-box tile so 32^3 boxes do not leave most of a
// workgroup idle
constexpr int kBatchRhsShapesMax = 4;
static const int kBatchRhsShapesF[kBatchRhsShapesMax][2] = {{32, 8}, {64, 8}, {64, 16}, {32, 16}};
static const int kBatchRhsShapesD[kBatchRhsShapesMax][2] = {{32, 4}, {32, 8}, {64, 8}, {0, 0}};
inline int batch_rhs_shapes(int esize) { return esize == 8 ? 3 : 4; }
inline const int *batch_rhs_shape(int esize, int q) { return esize == 8 ? kBatchRhsShapesD[q] : kBatchRhsShapesF[q]; }

// batched launchers (lcs_flow.hip / lcs_particles.hip); all enqueue on `s`
int launch_b_ghost_fill(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *work, int nwork, int set,
                        int narr, int r, int full, int hsmac);
int launch_b_cfl(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *rows, int nrows, int nbox,
                 double cfl, double dtmax);
// apply = false: the means and A only (the first stage-2 sweep applies the
// forcing, launch_b_hsmac_sweep with force, then launch_b_lf_power sets p)
int launch_b_lf(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *rows, int nrows, const int *boxes,
                int nbox, bool apply = true);
int launch_b_lf_power(hipStream_t s, const BoxRec *tab, const int *boxes, int nbox);
// rc-energy forcing of the listed boxes (src/flows/forcings/rcf.jl:21-191),
// the single-box kernels' arithmetic in one launch per phase: box means
// (work items: box, 3 c^3 blocks), forward DFT (box, crop slot), spectrum
// rescaling (one block per box), apply (box, rcf_blocks blocks), power
int launch_b_rcf(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *mean, int nmean, const WorkItem *dft,
                 int ndft, const int *boxes, int nbox, const WorkItem *apply, int napply);
// checks rc-energy parameters for box b and sets up its buffers and its
// forced-mode table on the device (synchronous; not in captured work)
int rcf_prepare(Box &b, int kmin, int kmax, int c);
int rcf_apply_blocks(const Box &b);
int rcf_crop_slots(int kmax);
// tiles of one shape (tx x ty, batch_rhs_shape); spare_out: write the result
// to the boxes' spare arrays instead of the stage's output (for timing)
int launch_b_rhs(hipStream_t s, int esize, int tx, int ty, const BoxRec *tab, const TileItem *tiles, int ntiles,
                 int stage, bool spare_out = false);
// tx: tile width of the fused sweep (32 or 64); the tiles listed must all be of that width
// force: boxes with lf_force apply the linear forcing of S->lf on load and
// leave their per-tile power in part[kPowerSlot]
// wrap: U, V, W of ghost cells are loaded from their periodic images (the
// sweep stands in for a FullHalo fill of U, V, W of width h before it)
// persist_cus > 0: a persistent grid of as many workgroups as that many CUs
// hold at once walks the tile list, drawing from per-XCD queues (queue: 9
// zeroed ints on the device, left zeroed) where the grid has 8 or more
// workgroups and queue is given, else by stride; 0: one workgroup per tile
int launch_b_hsmac_sweep(hipStream_t s, int esize, int coloring, int tx, int ty, const BoxRec *tab,
                         const TileItem *tiles, int ntiles, int set, int kind, int slot, double dtfac,
                         bool force = false, bool wrap = false, int persist_cus = 0, int *queue = nullptr);
// both colours over the x face slabs of every box (lcs_hsmac_xface.h): one
// work item per patch, partials of a box in part[slot][patch]; then one work
// item per 256 slab rows for the commit
int launch_b_hsmac_xface(hipStream_t s, int esize, int coloring, const BoxRec *tab, const WorkItem *work, int nwork,
                         int set, int domain, int slot, double dtfac);
int launch_b_hsmac_xface_commit(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *work, int nwork, int set,
                                int domain);
int launch_b_hsmac_faces(hipStream_t s, int esize, int coloring, const BoxRec *tab, const WorkItem *work,
                         int nwork, int set, int domain, int axis, int color, int slot, double dtfac);
// kinds of the passes since the last check: 0 = fused sweep, 1..3 = face pass of axis z, y, x
int launch_b_hsmac_check(hipStream_t s, const BoxRec *tab, int nbox, const int *kinds, int npass, double epsp,
                         int itrmin);
int launch_b_hsmac_ctl(hipStream_t s, BoxRec *tab, int nbox, int op, int arg, HsPoll *poll);
int launch_b_particles(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *work, int nwork, int stage);
int launch_b_particle_dt(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *work, int nwork, int nbox,
                         double dtmax);
int launch_b_state(hipStream_t s, const BoxRec *tab, int nbox, DevScalars *gather);
// HSMAC control ops for launch_b_hsmac_ctl
// Gate(arg): sets the stop flag to arg if any box has not converged;
// Unhalt clears it (the only op that runs while it is set).
enum { kHsBegin = 0, kHsFlip = 1, kHsEnd = 2, kHsCommit = 3, kHsPoll = 4, kHsGate = 5, kHsUnhalt = 6 };

// a box leaving its batch (lcs_box_destroy); a change of the box's buffers
// that the batch's device table must pick up
int batch_detach_box(Box &b);
void batch_mark_dirty(Box &b);

// ---- counter-based RNG shared with the oracle ------------------------------
__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline double rand_u01(uint64_t seed, uint64_t stream, uint64_t idx)
{
    uint64_t z = mix64(seed ^ mix64(stream ^ mix64(idx)));
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline double spacing(int n) { return kDomainLength / (double)n; }

}  // namespace lcs
