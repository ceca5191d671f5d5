/*
 * lcs_mi355x.h -- C ABI of the MI355X (gfx950) backend for the LCS.jl
 * per-time-step hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int return codes.
 * Every function returns LCS_OK (0) or a negative error class and sets a
 * thread-local message readable with lcs_last_error().
 *
 * Layout contract (the reference's, src/flows/buffer.jl:52-55 and
 * lib/offsets/src/Offsets.jl:23-43): every flow field is a column-major
 * (x fastest) array of (nx+2h)*(ny+2h)*(nz+2h) elements whose domain index
 * (i,j,k), 1-based, lives at 0-based offset (i-1+h) + sx*((j-1+h) + sy*(k-1+h))
 * with sx = nx+2h, sy = ny+2h.  dU/dt fields are interior-only nx*ny*nz.
 * Particle properties are SoA vectors of ntotal elements (ids int64, the rest
 * in the box's float type).
 *
 * Threading: one HIP stream per box.  Calls that return host scalars
 * synchronise that stream; all others are asynchronous.  Handles are
 * thread-compatible, not thread-safe.
 */
#ifndef LCS_MI355X_H
#define LCS_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error classes ---------------------------------------------------- */
#define LCS_OK 0
#define LCS_ERR_INVALID -1  /* bad argument / unsupported configuration */
#define LCS_ERR_HIP -2      /* HIP runtime failure */
#define LCS_ERR_CAPACITY -3 /* capacity or memory budget exceeded */
#define LCS_ERR_STATE -4    /* call out of order (e.g. no particles) */

/* ---- enums ------------------------------------------------------------ */
#define LCS_F64 0 /* the reference's float type (src/LCS.jl:20) */
#define LCS_F32 1

/* field ids (lcs_field_upload/download, lcs_fill_ghost masks) */
#define LCS_FIELD_U 0  /* Us[1]  */
#define LCS_FIELD_V 1  /* Us[2]  */
#define LCS_FIELD_W 2  /* Us[3]  */
#define LCS_FIELD_U2 3 /* Us2[1] */
#define LCS_FIELD_V2 4 /* Us2[2] */
#define LCS_FIELD_W2 5 /* Us2[3] */
#define LCS_FIELD_P 6
#define LCS_FIELD_DU 7 /* dUdts[1], interior-only */
#define LCS_FIELD_DV 8
#define LCS_FIELD_DW 9

/* particle property ids, in the order of Tuple(::Properties)
 * (src/particles/buffer.jl:32-35); N_PROPERTIES = 17 */
#define LCS_PROP_ID 0 /* int64 */
#define LCS_PROP_DIAM 1
#define LCS_PROP_X 2
#define LCS_PROP_Y 3
#define LCS_PROP_Z 4
#define LCS_PROP_X2 5
#define LCS_PROP_Y2 6
#define LCS_PROP_Z2 7
#define LCS_PROP_U 8
#define LCS_PROP_V 9
#define LCS_PROP_W 10
#define LCS_PROP_U2 11
#define LCS_PROP_V2 12
#define LCS_PROP_W2 13
#define LCS_PROP_DUDT 14
#define LCS_PROP_DVDT 15
#define LCS_PROP_DWDT 16
#define LCS_N_PROPERTIES 17

#define LCS_COLORING_RED_BLACK_FAST 0 /* lib/parallel/src/foraxes.jl:93-107 */
#define LCS_COLORING_RED_BLACK 1      /* lib/parallel/src/foraxes.jl:82-88 */

#define LCS_DRAG_LINEAR 0    /* src/particles/config.jl:124-128 */
#define LCS_DRAG_NONLINEAR 1 /* src/particles/config.jl:135-139 */

#define LCS_FORCING_NONE 0        /* NoForcing, src/flows/config.jl:113-115 */
#define LCS_FORCING_LINEAR_POWER 1 /* ConstantPowerLF, src/flows/config.jl:132-136 */
#define LCS_FORCING_RC_ENERGY 2    /* EnergyPreserveRCF, src/flows/config.jl:162-175 */

typedef struct lcs_ctx lcs_ctx;
typedef struct lcs_box lcs_box;

/* Grid + FlowParams + float type of one periodic box
 * (src/config.jl:56-60 Grid; src/flows/config.jl:11-18 FlowParams). */
typedef struct {
    int nx, ny, nz;
    int halo; /* grid.halo_size; flow phases need >= 3 */
    int dtype; /* LCS_F64 or LCS_F32 */
    double U0_m_s, L0_m, NU_m2_s;
} lcs_box_desc;

/* HSMAC couple parameters (src/flows/config.jl:88-96) */
typedef struct {
    int time_blocking;
    int overlap;
    int coloring; /* LCS_COLORING_* */
    double omega;
    double epsp;
    int itrmin;
    int itrmax;
} lcs_hsmac_params;

/* One particle profile (src/particles/config.jl:11-41,51-56,272-281) */
typedef struct {
    int64_t nvalid; /* population.valid on one process */
    int64_t ntotal; /* ceil(nvalid*(1+invalid_ratio)), src/particles/config.jl:22 */
    double gravity;
    double density_ratio;
    int drag;       /* LCS_DRAG_* */
    int grid_ratio; /* cell.grid_ratio */
} lcs_particle_desc;

/* Forcing selection for lcs_step (src/flows/config.jl:105-175) */
typedef struct {
    int kind; /* LCS_FORCING_* */
    double power;       /* linear-power */
    double energy;      /* rc-energy */
    int kmin, kmax;     /* rc-energy */
    int coarse_size;    /* rc-energy */
} lcs_forcing;

/* Everything timestep! + evolve! need beyond the box (src/config.jl:36-54 TimeStep) */
typedef struct {
    double cfl;
    double dtmax;
    lcs_hsmac_params hsmac;
    lcs_forcing forcing;
    /* lcs_particles_localize every profile before every localize_interval-th
     * lcs_step call on the box (0: never).  No counterpart in the reference;
     * it changes no result. */
    int localize_interval;
} lcs_step_config;

/* The log of one step: the reference returns it as a NamedTuple
 * (src/evolve.jl:127-145) plus the state update of src/evolve.jl:10-13. */
typedef struct {
    int64_t step;
    double t;
    double dt;
    int itr1, itr2;   /* couple1.itr, couple2.itr */
    double div1, div2; /* couple1.div, couple2.div */
    double p;          /* forcing.p */
    double e10;        /* forcing.e10 (rc-energy) */
} lcs_step_log;

/* ---- context / box ------------------------------------------------------ */
/* replaces Topologies.device! (src/topologies.jl:1-3) for one device */
int lcs_ctx_create(int device, lcs_ctx **out);
int lcs_ctx_destroy(lcs_ctx *ctx);

/* replaces LCS.Buffer / Flows.FlowBuffer (src/buffer.jl:27-34,
 * src/flows/buffer.jl:115-121): allocates Us, Us2, P in HBM */
int lcs_box_create(lcs_ctx *ctx, const lcs_box_desc *desc, lcs_box **out);
int lcs_box_destroy(lcs_box *box);
/* bytes of device memory the box holds now */
int lcs_box_bytes(const lcs_box *box, size_t *bytes);
/* the descriptor the box was created with and its number of particle
 * profiles (nprofiles may be NULL) */
int lcs_box_describe(const lcs_box *box, lcs_box_desc *desc, int *nprofiles);

/* host <-> device copies in the reference layout (blocking) */
int lcs_field_upload(lcs_box *box, int field, const void *host, size_t bytes);
int lcs_field_download(lcs_box *box, int field, void *host, size_t bytes);

/* ---- flow phases ---------------------------------------------------------- */
/* Topologies.synchalo! single-process (src/topologies.jl:36-45,
 * lib/topologies/src/halo.jl:138-159): field_mask bit f selects LCS_FIELD_f */
int lcs_fill_ghost(lcs_box *box, unsigned field_mask, int width, int full);

/* Flows.timestep (src/flows/timestep.jl:6-65): FaceHalo ghost fill of Us,
 * 4th-order cell-centred |u| maxima, dt.  umax3 may be NULL. */
int lcs_flow_timestep(lcs_box *box, double cfl, double dtmax, double *dt, double *umax3);

/* Flows.rhs! (src/flows/rhs.jl:9-81): stage 1 reads Us, stage 2 reads Us2;
 * writes dU/dV/dW */
int lcs_rhs(lcs_box *box, int stage);
/* Flows.update! (src/flows/update.jl:12-54): stage 1 Us2 = Us + dt*dUdt,
 * stage 2 Us = Us + dt*dUdt */
int lcs_update(lcs_box *box, int stage, double dt_stage);
/* rhs! followed by update! in one pass, without storing dU/dt */
int lcs_rhs_update(lcs_box *box, int stage, double dt_stage);

/* Flows.couple!(::HSMAC) (src/flows/couples/couple.jl:13-17,
 * src/flows/couples/hsmac.jl:52-201): stage 1 acts on Us2, stage 2 on Us */
int lcs_hsmac(lcs_box *box, int stage, double dt_stage, const lcs_hsmac_params *params, int *itr,
              double *div);

/* forcing!(::ConstantPowerLF) (src/flows/forcings/lf.jl:6-106) on Us */
int lcs_force_linear(lcs_box *box, double power, double dt, double *p);
/* forcing!(::EnergyPreserveRCF) (src/flows/forcings/rcf.jl:1-191) on Us;
 * box mean, spectrum rescaling and sparse inverse DFT all run on the device.
 * Needs 1 <= kmin <= kmax <= coarse_size/2, coarse_size dividing every grid
 * dimension, and at most 128 forced wavevectors (shells kmin..kmax counted
 * once per +-k pair: 9 for shell 1, 31 for shell 2, 105 for shell 4 alone,
 * so kmax <= 4); otherwise LCS_ERR_INVALID. */
int lcs_force_rcf(lcs_box *box, double energy, int kmin, int kmax, int coarse_size, double dt,
                  double *p);
/* lcs_force_rcf without any device allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  Same arguments, kernels and results as
 * lcs_force_rcf, but the coarse grids, spectra and crop live in one fixed
 * 1 MiB device array per GPU (131072 doubles: 3*c^3 + 12*(kmax+1)*(2kmax+1)^2
 * must fit, e.g. coarse_size <= 32 with kmax <= 4) instead of the box's own
 * buffer; larger requests fail with LCS_ERR_CAPACITY.  Every call on a GPU
 * shares that array, so calls that overlap on one GPU (boxes on different
 * streams) race; the array keeps the last call's contents, and every element
 * a call reads is written by that call first. */
int lcs_force_rcf_noalloc(lcs_box *box, double energy, int kmin, int kmax, int coarse_size, double dt,
                          double *p);
/* box-mean filter only (Flows.boxmean_l!, ext/AMDGPUExt.jl:18-41): coarse3
 * receives 3*c^3 values (U then V then W), column-major, as doubles */
int lcs_boxmean(lcs_box *box, int coarse_size, double *coarse3);

/* ---- particle statistics ------------------------------------------------------ */
/* Particles.Stat (src/particles/config.jl:66-108): RDF and spectrum settings. */
typedef struct {
    double contact_tolerance; /* rdf.contact_tolerance */
    int bin_size;             /* rdf.bin_size: number of log-spaced bins, 1..1024 */
    double min_radius;        /* rdf.min_radius */
    int search_cell_size;     /* rdf.search_cell_size: +-cells searched; rmax = it x min spacing */
    int max_wavenumber;       /* spectrum.max_wavenumber (0: no spectrum) */
    int nsamples;             /* spectrum.nsamples: directions per shell */
} lcs_particle_stat_params;

/* Particles.ParticleStat (src/particles/stat.jl:46-75) minus its arrays:
 * LCS.Summary (mean, max, min) of x, y, z, u, v, w, diam over 1:nvalid. */
typedef struct {
    double x_mean, x_max, x_min, y_mean, y_max, y_min, z_mean, z_max, z_min;
    double u_mean, u_max, u_min, v_mean, v_max, v_min, w_mean, w_max, w_min;
    double diam_mean, diam_max, diam_min;
    double taup;  /* of the mean diameter */
    double St;    /* taup / tau_eta */
    double gr_contact;
    int64_t npairs_contact;
    int nbins, nshells;
} lcs_particle_stats;

/* Particles.stat(fstat, buf, state, config, topo, iprofile)
 * (src/particles/stat.jl:80-155): re-bins the profile (makeindex!), then the
 * summaries, the radial distribution function over the cell index
 * (rdf_assume_ci, src/particles/rdf.jl:1-103) and the number-density spectrum
 * (src/particles/spectrum.jl:11-78).  tau_eta is FlowStat.tau_eta.  Arrays
 * (any may be NULL): edges[bin_size+1], npairs[bin_size], gr[bin_size],
 * wavenumbers[nshells], Enp[nshells] with nshells the wavenumbers of
 * (1, 2, 3, 4, 6, 8, ..., 768) up to max_wavenumber.  Synchronises. */
int lcs_particle_stat(lcs_box *box, int profile, const lcs_particle_stat_params *params, double tau_eta,
                      lcs_particle_stats *out, double *edges, int64_t *npairs, double *gr, double *wavenumbers,
                      double *Enp);
/* Particles.density (src/particles/stat.jl:171-217): re-bins the profile, then
 * the SPH density of each of the nvalid particles (cubic spline, h = smallest
 * grid spacing), in the order the re-binning leaves.  Synchronises. */
int lcs_particle_density(lcs_box *box, int profile, double *density);

/* ---- tile shapes ------------------------------------------------------------ */
/* No counterpart in the reference (its kernels are KernelAbstractions
 * workgroups of a fixed size).  The fused RHS and HSMAC sweep kernels march
 * tx x ty columns of cells up in z; the shape changes speed, never results.
 * Sweep shapes: (64|32) x (16|8), and in FP32 also 128x8 and 128x16 (two
 * cells per thread).  RHS shapes: FP32 (64|32) x (16|8), FP64
 * 64x8, 32x8, 32x4, 32x16.  A zero sweep shape means the size rule for all of them. */
typedef struct {
    int sweep_tx, sweep_ty;
    int rhs_tx, rhs_ty;
    int sweep2_tx, sweep2_ty; /* FP32: the d2 and d3 sweeps of a time-blocked group before itrmin as one
                                 pass, 32x24, 64x12 or 32x16; 0x0 = two single sweeps */
} lcs_tile_shapes;

/* Times every candidate shape on a scratch box with this box's x-y extent (at
 * most 256 planes deep) and keeps the fastest for this box; the choice is
 * cached per device, size and float type, so later boxes of the same size
 * take it at once.  Falls back to the size rule if the scratch box does not
 * fit in device memory.  chosen may be NULL. */
int lcs_autotune(lcs_box *box, lcs_tile_shapes *chosen);
/* Force shapes (NULL: back to the size rule); LCS_ERR_INVALID if a shape is
 * not one of the candidates. */
int lcs_set_tiles(lcs_box *box, const lcs_tile_shapes *shapes);
/* The shapes the next launches will use (the size rule resolved). */
int lcs_get_tiles(lcs_box *box, lcs_tile_shapes *shapes);

/* ---- flow statistics -------------------------------------------------------- */
/* Flows.FlowStat (src/flows/stat.jl:1-42): summaries (mean, max, min) of the
 * raw U, V, W interior values (LCS.Summary, src/stat.jl:4-22) and the derived
 * turbulence quantities of stat.jl:132-154. */
typedef struct {
    double u_mean, u_max, u_min;
    double v_mean, v_max, v_min;
    double w_mean, w_max, w_min;
    double k;         /* 0.5 (<u2> - |mean|^2) */
    double eps;       /* <2 S_ij S_ij> / Re */
    double tau_eta;   /* Kolmogorov time */
    double l_mu;      /* Taylor microscale */
    double l_eta;     /* Kolmogorov length */
    double Re_lambda;
    double urms;
    double k_eta;     /* (n/2) l_eta, n the smallest dimension */
    double skew;      /* -<dudx^3> / <dudx^2>^1.5 */
    double flat;      /* <dudx^4> / <dudx^2>^2 */
    double L;         /* integral length; 0 unless requested */
    double C_eps;     /* eps L / u'^3 */
} lcs_flow_stats;

/* Flows.stat(buf, state, config, topo) (src/flows/stat.jl:53-155): FullHalo
 * ghost fill on U, V, W, then the statistics of the current velocity.
 * integral_length != 0 also computes the integral length (stat.jl:157-182,
 * config flag flow.stat.integral_length).  Synchronises. */
int lcs_flow_stat(lcs_box *box, int integral_length, lcs_flow_stats *out);

/* ---- particles ------------------------------------------------------------ */
/* Particles.ParticleBuffer (src/particles/buffer.jl:124-135) for one profile;
 * profile index returned in *profile (0-based, in call order) */
int lcs_particles_add(lcs_box *box, const lcs_particle_desc *desc, int *profile);
/* the descriptor a profile was added with (nvalid = state.particles[i].nvalid) */
int lcs_particles_describe(const lcs_box *box, int profile, lcs_particle_desc *desc);
int lcs_particles_upload(lcs_box *box, int profile, int prop, const void *host, size_t bytes);
int lcs_particles_download(lcs_box *box, int profile, int prop, void *host, size_t bytes);
/* Particles.boundary! (src/particles/boundary.jl:33-120), single process */
int lcs_particles_boundary(lcs_box *box, int profile, int stage);
/* Particles.motion! (src/particles/motion.jl:15-84) */
int lcs_particles_motion(lcs_box *box, int profile, int stage);
/* Particles.update! (src/particles/update.jl:18-65) */
int lcs_particles_update(lcs_box *box, int profile, int stage, double dt_stage);
/* Particles.timestep (src/particles/timestep.jl:15-25) */
int lcs_particles_timestep(lcs_box *box, int profile, double dtmax, double *dt);
/* Particles.makeindex! (src/particles/makeindex.jl:9-67): stage 1 bins xss,
 * stage 2 bins xss2; permutes all 17 properties; stable sort */
int lcs_particles_makeindex(lcs_box *box, int profile, int stage);
/* No counterpart in the reference.  Reorders the profile's device storage by
 * the 8^3-cell tile of each particle's position, so the field gathers of
 * motion! read nearby memory.  Results do not change, and every other entry
 * point keeps presenting particles in their logical order (the reference's
 * array order): upload/download map through the storage order, and makeindex
 * restores the logical order before it sorts. */
int lcs_particles_localize(lcs_box *box, int profile);
/* the CellIndexBuffer after makeindex (src/particles/buffer.jl:45-75):
 * hashes (nvalid), starts/stops ((nc+2)^3, 1-based); any pointer may be NULL */
int lcs_particles_cellindex(lcs_box *box, int profile, int64_t *hashes, int64_t *starts,
                            int64_t *stops, int64_t *ncells);

/* ---- initial conditions (builder-defined; no counterpart in the reference) */
/* seeded solenoidal field with E(k) ~ k^4 exp(-2 (k/k0)^2), scaled to u' */
int lcs_init_spectrum(lcs_box *box, uint64_t seed, double k0, double uprime);
/* uniform positions, log-uniform Stokes numbers in [st_min, st_max] with
 * St = taup/taueta, taueta = sqrt(1/(eps*Re)); zero_slip != 0 sets the particle
 * velocity to the interpolated fluid velocity (needs filled ghosts), else rest */
int lcs_init_particles(lcs_box *box, int profile, uint64_t seed, double st_min, double st_max,
                       double eps, int zero_slip);

/* ---- whole step ------------------------------------------------------------ */
/* LCS.timestep! + LCS.evolve!(::FlowParticleMode / ::FlowMode)
 * (src/timestep.jl:1-27, src/evolve.jl:15-146).  All phases are enqueued
 * on the box stream; only the log is read back.  log may be NULL, in which
 * case the call does not synchronise. */
int lcs_step(lcs_box *box, const lcs_step_config *cfg, lcs_step_log *log);
/* the most recent step's log (synchronises) */
int lcs_step_log_read(lcs_box *box, lcs_step_log *log);
/* state.step = step; state.t = t, as LCSIO.load_resume_state! sets them on
 * resume (src/LCSIO/load.jl:1-21).  state.dt is recomputed by the next step. */
int lcs_set_state(lcs_box *box, int64_t step, double t);

/* ---- batches of boxes ---------------------------------------------------- */
/* No counterpart in the reference: an ensemble of independent boxes advanced
 * together, one kernel launch per phase for all of them.  A box added to a
 * batch is stepped only by lcs_batch_step (lcs_step and lcs_hsmac on it fail
 * with LCS_ERR_STATE); its own entry points for data (upload, download,
 * particles) keep working.  Each box keeps its own state, dt, step count and
 * HSMAC convergence, and advances exactly as lcs_step would advance it alone.
 * Boxes of one batch share the float type, halo size, cfl, dtmax and HSMAC
 * parameters; sizes, particle profiles and forcing (kind and parameters) may
 * differ.  Both forcings run batched: one launch per phase over every box with
 * that forcing. */
typedef struct lcs_batch lcs_batch;
int lcs_batch_create(lcs_ctx *ctx, lcs_batch **out);
int lcs_batch_destroy(lcs_batch *batch); /* detaches every box first */
int lcs_batch_add(lcs_batch *batch, lcs_box *box, const lcs_step_config *cfg);
int lcs_batch_remove(lcs_batch *batch, lcs_box *box);
int lcs_batch_size(lcs_batch *batch, int *nboxes);
/* one timestep! + evolve! of every box; logs (may be NULL) receives one entry
 * per box in the order of addition (at most nlogs) */
int lcs_batch_step(lcs_batch *batch, lcs_step_log *logs, int nlogs);
int lcs_batch_sync(lcs_batch *batch);
/* hipGraphs of the batched step; results are the same in every mode.
 * on = 1 (default): the first two steps after a membership change run
 * eagerly (a server's membership often changes again within a step or two);
 * from the third step on, each step is ONE graph: both stages with as many
 * HSMAC groups per stage as any of the last 8 steps needed, a device-side gate
 * after them that
 * stops the rest of the graph (every kernel returns at once) if a box still
 * needs groups, and the copies of both stages' convergence state to the host.
 * A stopped step goes on from the gate with polled groups, whose runs of
 * launches between convergence polls (one HSMAC group; stage 2 up to the
 * forcing; ...) are captured and replayed.  on = 2: those runs only, from
 * the second step on.  on = 3: one graph per step holding no groups beyond
 * those before itrmin (every step that needs more goes on from the gate; for
 * tests).  on = 0: eager launches. */
int lcs_batch_set_graphs(lcs_batch *batch, int on);
/* Steps taken since the batch was created: whole = one graph that ran to its
 * end, gated = one graph stopped at a gate and finished with polled groups,
 * segmented = eager or segment launches.  Any pointer may be NULL. */
int lcs_batch_graph_stats(lcs_batch *batch, int64_t *whole, int64_t *gated, int64_t *segmented);
/* No counterpart in the reference (its kernels launch one workgroup per
 * block).  cus < 0: each batched HSMAC sweep runs as a persistent grid, as
 * many workgroups as the device holds at once, each walking the tile table
 * (tiles b, b + G, b + 2G, ...); cus > 0: the grid that many CUs hold (at most
 * the device's); cus = 0 (default): one workgroup per tile.  Results are the
 * same in every case (a tile's arithmetic does not depend on the workgroup
 * that runs it).  Drops the batch's captured graphs. */
int lcs_batch_set_persistent(lcs_batch *batch, int cus);
/* No counterpart in the reference (its kernels have fixed workgroups).  The
 * batched HSMAC sweep tiles each size class of boxes (equal nx, ny, nz) with
 * one shape, 64x8, 32x8, 64x16 or 32x16, in z-chunks of a target length; by
 * default 32x8 and about 28 planes for every class.  lcs_batch_autotune times
 * the sweeps of one HSMAC group over the whole batch for every shape and
 * chunk target of 16, 28 and 48 planes, one class at a time with the others
 * at their current choice, in two rounds, and keeps the fastest per class.
 * The timed sweeps read the boxes' current fields and write only their
 * ping-pong partners, so no box changes.  Fields and sweep counts do not
 * depend on the tiling; the logged divergence and forcing power are sums of
 * per-tile partials, so their last bits do.  It then picks the fused RHS's
 * tile shape per class the same way (lcs_batch_set_rhs_tiles), timing
 * stage 1's RHS written into the spare arrays.  Synchronises; the next step
 * rebuilds the batch's tables and graphs. */
int lcs_batch_autotune(lcs_batch *batch);
/* Sets the tiling of the size class of box `index` (order of addition);
 * LCS_ERR_INVALID for another shape or zchunk < 1. */
int lcs_batch_set_tiles(lcs_batch *batch, int index, int tx, int ty, int zchunk);
/* The sweep tile of box `index` and its z-chunk length in planes. */
int lcs_batch_get_tiles(lcs_batch *batch, int index, int *tx, int *ty, int *zchunk);
/* No counterpart in the reference.  The fused RHS tile of the size class of
 * box `index`: FP32 32x8 (default), 64x8, 64x16 or 32x16; FP64 32x4
 * (default), 32x8 or 64x8.  lcs_batch_autotune also picks it per class, by
 * timing stage 1's RHS into the boxes' spare arrays (no box changes).  Fields
 * do not depend on it. */
int lcs_batch_set_rhs_tiles(lcs_batch *batch, int index, int tx, int ty);
int lcs_batch_get_rhs_tiles(lcs_batch *batch, int index, int *tx, int *ty);

/* ---- serving ---------------------------------------------------------------- */
/* No counterpart in the reference.  A request loop over one batch: runs are
 * submitted as requests and admitted while a slot (max_live) is free and
 * their device-memory footprint fits what the budget (memory_budget) has
 * left.  Admission follows arrival order, except that a request waiting for
 * memory lets later requests that fit go first, until it has been passed
 * over max_live times.  A request larger than the whole budget is refused at
 * submit with LCS_ERR_CAPACITY.  An admitted run gets its own box (seeded
 * spectrum, one particle profile with log-uniform Stokes numbers, zero
 * slip), and lcs_server_step advances all live runs together, one batched
 * step per call.  A run that reaches its step count leaves the batch at once
 * and its slot goes to the next request; its box stays readable
 * (lcs_server_box) until lcs_server_release.  Times are seconds on the
 * server's monotonic clock since lcs_server_create. */
typedef struct lcs_server lcs_server;
#define LCS_RUN_QUEUED 0
#define LCS_RUN_LIVE 1
#define LCS_RUN_DONE 2
#define LCS_RUN_FAILED 3
typedef struct {
    int n;                /* cube edge: n^3 cells, halo 3 */
    int dtype;            /* LCS_F32 / LCS_F64 */
    int64_t particles;    /* particles of the run's one profile (0: flow only) */
    double st_min, st_max; /* Stokes numbers, log-uniform */
    int forcing;          /* LCS_FORCING_*: rc-energy uses kmin 1, kmax 2, coarse 8 */
    double forcing_param; /* linear-power P0, or rc-energy energy */
    int steps;            /* steps to run */
    uint64_t seed;        /* spectrum and particle seed */
    double k0;            /* spectrum peak wavenumber (u' = 1) */
    double Re;            /* U0 L0 / nu with U0 = L0 = 1 */
} lcs_request;
typedef struct {
    size_t memory_budget; /* bytes of device memory the live runs may hold; 0: nine tenths of the
                             device's free memory at lcs_server_create, at most 288e9 */
    int max_live;         /* runs in the batch at once */
    double cfl, dtmax;
    lcs_hsmac_params hsmac;
    int localize_interval;
} lcs_server_config;
typedef struct {
    int64_t id;
    int status;           /* LCS_RUN_* */
    int64_t steps_done;
    double submit_s, start_s, finish_s;
    size_t bytes;         /* device bytes held by the run */
} lcs_run_status;
int lcs_server_create(lcs_ctx *ctx, const lcs_server_config *cfg, lcs_server **out);
int lcs_server_destroy(lcs_server *server);
int lcs_server_submit(lcs_server *server, const lcs_request *req, int64_t *id);
/* admit what fits, advance every live run one step, retire finished runs */
int lcs_server_step(lcs_server *server, int *n_live, int *n_finished);
int lcs_server_status(lcs_server *server, int64_t id, lcs_run_status *status);
int lcs_server_counts(lcs_server *server, int *queued, int *live, size_t *memory_used);
int lcs_server_box(lcs_server *server, int64_t id, lcs_box **box);
int lcs_server_release(lcs_server *server, int64_t id);

/* ---- measurement hook (no counterpart in the reference) ---------------------
 * enable != 0 starts recording HIP events on the box stream around every
 * kernel launch, keyed by kernel name, together with each launch's
 * algorithmic HBM bytes; enable == 0 stops and clears.  lcs_profile_read
 * returns entry `index` (0-based; LCS_ERR_INVALID past the end): name, total
 * milliseconds, launch count and total algorithmic bytes (synchronises). */
int lcs_profile(lcs_box *box, int enable);
int lcs_profile_read(lcs_box *box, int index, char *name, size_t name_len, double *total_ms,
                     int64_t *launches, double *bytes);

int lcs_sync(lcs_box *box);
const char *lcs_last_error(void);
/* library version string */
const char *lcs_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LCS_MI355X_H */
