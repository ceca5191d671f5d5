// Batches of boxes: an ensemble of independent periodic boxes of uneven size
// advanced with one launch per phase (lcs_batch_*).  Each box keeps its own
// state, step count and HSMAC convergence; the batch owns the stream, a
// device table of per-box records (BoxRec) and the work tables that map
// workgroups to (box, block) or (box, tile).  The per-box arithmetic is the
// single-box kernels' own device code, so a box advanced in a batch gives the
// same result as the same box advanced alone with lcs_step.
#include <algorithm>
#include <array>
#include <cmath>
#include <deque>
#include <map>
#include <vector>

#include "lcs_internal.h"
#include "lcs_hsmac_tile.h"
#include "lcs_hsmac_xface.h"
#include "lcs_rhs_tile.h"

namespace lcs {

template <typename E>
struct DevVec {
    E *d = nullptr;
    int n = 0, cap = 0;
    int assign(const std::vector<E> &h, hipStream_t s)
    {
        n = (int)h.size();
        if (n > cap) {
            if (d) LCS_HIP(hipFree(d));
            d = nullptr;
            cap = std::max(n, 16);
            LCS_HIP(hipMalloc(&d, sizeof(E) * (size_t)cap));
        }
        if (n) LCS_HIP(hipMemcpyAsync(d, h.data(), sizeof(E) * (size_t)n, hipMemcpyHostToDevice, s));
        return LCS_OK;
    }
    void release()
    {
        if (d) (void)hipFree(d);
        d = nullptr;
        n = cap = 0;
    }
};

struct Batch {
    Ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    int esize = 0;
    lcs_hsmac_params hs{};
    std::vector<Box *> boxes;
    std::vector<lcs_step_config> cfgs;
    std::vector<hipStream_t> own_streams;
    bool dirty = true;
    // device state
    std::vector<BoxRec> tab_h;
    DevVec<BoxRec> tab;
    DevVec<WorkItem> ghost_full, ghost_one, rows, lf_rows, parts, diams, faces[3];
    DevVec<WorkItem> xf_work, xf_commit;  // x-slab pair: patches, then 256-row blocks
    bool xface = false;                   // every box runs its x slabs as the pair
    DevVec<int> lf_boxes;            // boxes with linear-power forcing
    std::vector<int> rcf_boxes;      // boxes with rc-energy forcing
    DevVec<int> rcf_boxes_d;
    DevVec<WorkItem> rcf_mean, rcf_dft, rcf_apply;  // work of the batched rc-energy forcing
    DevVec<TileItem> rhs_tiles[kBatchRhsShapesMax];  // RHS tiles by shape: batch_rhs_shape
    DevVec<TileItem> hs_tiles[4];                    // sweep tiles by shape: kBatchSweepShapes
    HsPoll *poll_d = nullptr, *poll_h = nullptr;
    DevScalars *sc_d = nullptr, *sc_h = nullptr;
    int cap_boxes = 0;
    // hipGraphs.  The first two steps after a membership change run eagerly
    // (the first sets the kernels' attributes; a server's membership often
    // changes again within a step or two, and a capture that is not replayed
    // costs more than it saves).  From the third step on each step is one
    // graph (step_graphs, keyed by the post groups it holds per stage), and
    // the runs of launches between host polls (gexec, keyed by segment id)
    // are captured as a stopped step graph needs them; with graph_mode 2
    // the segments alone, from the second step on.
    bool graphs = true;
    int stable = 0;  // steps since the tables were last rebuilt
    int graph_mode = 1;  // lcs_batch_set_graphs: 1 whole steps, 2 segments only, 3 whole steps, fewest groups
    std::map<std::array<int, 2>, hipGraphExec_t> step_graphs;
    std::array<int, 2> post_seen = {0, 0};        // post groups the last call of each stage needed
    std::deque<std::array<int, 2>> post_hist;     // ... over the last kPostWindow steps
    int *halt_d = nullptr, *halt_h = nullptr;     // stop flag of the step graphs (BoxRec::halt)
    int64_t n_whole = 0, n_gated = 0, n_segmented = 0;  // steps by path (lcs_batch_graph_stats)
    // the next fused sweep applies the deferred linear forcing of the
    // linear-power boxes (set around stage 2's HSMAC only)
    bool force_next = false;
    bool wrap_next = false;  // the next sweep loads ghost U, V, W from their images
    bool wrap_ok = false;    // every box has n >= h on each axis
    std::map<int, hipGraphExec_t> gexec;
    // sweep tiling per size class (nx, ny, nz): index into kBatchSweepShapes
    // and z-chunk target in planes (lcs_batch_autotune, lcs_batch_set_tiles)
    std::map<std::array<int, 3>, std::array<int, 2>> sweep_tune;
    // RHS tile shape per size class: index into batch_rhs_shape (default 0)
    std::map<std::array<int, 3>, int> rhs_tune;
    // lcs_batch_set_persistent: persist_cus > 0 runs the sweeps as persistent
    // grids over the tile table, as many workgroups as that many CUs hold at
    // once; 0: one workgroup per tile.  cus: the device's CU count.
    int persist_cus = 0;
    int cus = 256;
    int *queue_d = nullptr;  // the persistent sweep's per-XCD tile queues (9 ints, zero between launches)
};

constexpr int kPostWindow = 8;

static void drop_graphs(Batch &B)
{
    for (auto &kv : B.gexec) (void)hipGraphExecDestroy(kv.second);
    B.gexec.clear();
    for (auto &kv : B.step_graphs) (void)hipGraphExecDestroy(kv.second);
    B.step_graphs.clear();
    B.post_hist.clear();
    B.stable = 0;
}

// Enqueues a run of launches with no host synchronisation inside, through its
// captured graph once the batch is warm.  The launches take every per-box
// value from the device table or the device scalars, so a replay is the same
// work as the eager run.
template <typename F>
static int segment(Batch &B, int id, F &&enqueue)
{
    if (!B.graphs || B.stable < (B.graph_mode == 2 ? 1 : 2)) return enqueue();
    auto it = B.gexec.find(id);
    if (it == B.gexec.end()) {
        LCS_HIP(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(B.stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(LCS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return fail(LCS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        it = B.gexec.emplace(id, x).first;
    }
    LCS_HIP(hipGraphLaunch(it->second, B.stream));
    return LCS_OK;
}

// Work items are listed box by box, then permuted so that workgroup L, which
// the hardware runs on XCD L % kXcds, takes item t of one contiguous range per
// XCD (the inverse of the order used by the single-box tile kernels).
template <typename E>
static std::vector<E> xcd_order(const std::vector<E> &v)
{
    const int n = (int)v.size(), X = 8;
    std::vector<E> out(v.size());
    const int per = n / X, rem = n % X;
    for (int L = 0; L < n; L++) {
        const int x = L % X, i = L / X;
        const int t = x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
        out[L] = v[t];
    }
    return out;
}

static int blocks_of(long n, int per = 256) { return (int)((n + per - 1) / per); }

// Sweep tile shapes of a batch, one launch per shape in use (fields do not
// depend on them; the per-tile partial sums of the divergence do).  Until
// lcs_batch_autotune or lcs_batch_set_tiles chooses per size class, every box
// takes 32 x 8 tiles in z-chunks of about 28 planes, so a sweep is one
// launch: in a mixed 96-box ensemble this beat one
// launch of 64 x 8 tiles plus one of 32 x 8 tiles chosen by lane waste per
// box (23.5 vs 24.4 ms a step), and 16-row tiles were slower (25.2 ms).
static const int kBatchSweepShapes[4][2] = {{64, 8}, {32, 8}, {64, 16}, {32, 16}};
constexpr int kBatchSweepChunk = 28;
// z-chunk targets the autotuner tries
static const int kBatchSweepChunks[3] = {16, 28, 48};

static std::array<int, 2> sweep_tiling(const Batch &B, const Geom &g)
{
    const auto it = B.sweep_tune.find({g.nx, g.ny, g.nz});
    return it == B.sweep_tune.end() ? std::array<int, 2>{1, kBatchSweepChunk} : it->second;
}

static int rhs_shape(const Batch &B, const Geom &g)
{
    const auto it = B.rhs_tune.find({g.nx, g.ny, g.nz});
    // (a choice made for the other float type falls back to the default)
    return it == B.rhs_tune.end() || it->second >= batch_rhs_shapes(B.esize) ? 0 : it->second;
}

// the fused RHS of a stage over every box, one launch per tile shape in use
static int b_rhs(Batch &B, int stage, bool spare_out = false)
{
    for (int q = 0; q < batch_rhs_shapes(B.esize); q++) {
        const int *sh = batch_rhs_shape(B.esize, q);
        int rc = launch_b_rhs(B.stream, B.esize, sh[0], sh[1], B.tab.d, B.rhs_tiles[q].d, B.rhs_tiles[q].n, stage,
                              spare_out);
        if (rc) return rc;
    }
    return LCS_OK;
}

template <typename T>
static void fill_coefs(const Box &b, const lcs_hsmac_params &hs, BoxRec &r)
{
    const Geom &g = b.g;
    const double d[3] = {spacing(g.nx), spacing(g.ny), spacing(g.nz)};
    for (int a = 0; a < 3; a++) {
        const T di = T(1.0 / d[a]);
        r.rhs_di[a] = di;
        r.rhs_di2[a] = (double)(T)(di * di);
    }
    r.rhs_Rei = (double)T(1.0 / b.Re());
    const T dxi = T(1.0 / d[0]), dyi = T(1.0 / d[1]), dzi = T(1.0 / d[2]);
    r.hs_dxi = dxi;
    r.hs_dyi = dyi;
    r.hs_dzi = dzi;
    r.hs_dl = (double)(T)(T(1) / (T(2) * ((dxi * dxi + dyi * dyi) + dzi * dzi)));
    r.hs_omega = (double)T(hs.omega);
    r.dx = d[0];
    r.dy = d[1];
    r.dz = d[2];
}

template <typename T>
static void fill_profile(const Box &b, const ParticleProfile &pp, DevProfile &p)
{
    p.nvalid = pp.desc.nvalid;
    p.ntotal = pp.desc.ntotal;
    for (int q = 0; q < LCS_N_PROPERTIES; q++) p.props[q] = pp.props[q];
    p.mdx = (double)T(spacing(b.g.nx));
    p.mdy = (double)T(spacing(b.g.ny));
    p.mdz = (double)T(spacing(b.g.nz));
    p.mRe = (double)T(b.Re());
    p.mg = (double)T(pp.desc.gravity);
    p.mkap = (double)T(((1.0 / 18.0) * pp.desc.density_ratio) * b.Re());
    p.kap = ((1.0 / 18.0) * pp.desc.density_ratio) * b.Re();
    p.drag = pp.desc.drag;
}

// the record of box i from the host mirror of its pointers
static void refresh_record(Batch &B, int i)
{
    Box &b = *B.boxes[i];
    BoxRec &r = B.tab_h[i];
    for (int q = 0; q < 7; q++) r.f[q] = b.f[q];
    for (int q = 0; q < 4; q++) r.spare[q] = b.spare[q];
    for (int q = 0; q < (int)b.profiles.size(); q++)
        for (int k = 0; k < LCS_N_PROPERTIES; k++) r.prof[q].props[k] = b.profiles[q].props[k];
}

// The table is small (a few KB per box): it is uploaded once the stream has
// drained, so no kernel reads a record while it changes.
static int upload_table(Batch &B)
{
    LCS_HIP(hipStreamSynchronize(B.stream));
    int rc = B.tab.assign(B.tab_h, B.stream);
    if (rc) return rc;
    LCS_HIP(hipStreamSynchronize(B.stream));
    return LCS_OK;
}

static int rebuild(Batch &B)
{
    drop_graphs(B);
    const int n = (int)B.boxes.size();
    if (n > B.cap_boxes) {
        if (B.poll_d) LCS_HIP(hipFree(B.poll_d));
        if (B.poll_h) LCS_HIP(hipHostFree(B.poll_h));
        if (B.sc_d) LCS_HIP(hipFree(B.sc_d));
        if (B.sc_h) LCS_HIP(hipHostFree(B.sc_h));
        B.cap_boxes = std::max(n, 2 * B.cap_boxes);
        // two polls per box: stage 1 and stage 2 of a step graph
        LCS_HIP(hipMalloc(&B.poll_d, sizeof(HsPoll) * 2 * (size_t)B.cap_boxes));
        LCS_HIP(hipHostMalloc((void **)&B.poll_h, sizeof(HsPoll) * 2 * (size_t)B.cap_boxes, hipHostMallocDefault));
        LCS_HIP(hipMalloc(&B.sc_d, sizeof(DevScalars) * (size_t)B.cap_boxes));
        LCS_HIP(hipHostMalloc((void **)&B.sc_h, sizeof(DevScalars) * (size_t)B.cap_boxes, hipHostMallocDefault));
    }
    B.tab_h.assign(n, BoxRec{});
    B.wrap_ok = true;
    for (const Box *b : B.boxes)
        if (b->g.nx < b->g.h || b->g.ny < b->g.h || b->g.nz < b->g.h) B.wrap_ok = false;
    std::vector<WorkItem> gfull, gone, rows, lf_rows, parts, diams, faces[3], xfw, xfc, rmean, rdft, rapply;
    std::vector<int> lf_boxes;
    B.rcf_boxes.clear();
    std::vector<TileItem> rtiles[kBatchRhsShapesMax], htiles[4];
    const int domain = B.hs.time_blocking ? 1 : 3;
    B.xface = n > 0;
    for (int i = 0; i < n; i++) {
        const Geom &g = B.boxes[i]->g;
        Box3 xs[2];
        hs_faces(g, domain, 2, xs[0], xs[1]);
        const long pn = B.boxes[i]->esize == 8 ? XfPatch<double>::N : XfPatch<float>::N;
        const long patches = 2 * ((g.ny - 1 + pn - 1) / pn) * ((g.nz - 1 + pn - 1) / pn);
        if (!xface_fits(g, xs) || patches > kMaxPartBlocks) B.xface = false;  // one partial per patch
    }
    for (int i = 0; i < n; i++) {
        Box &b = *B.boxes[i];
        const Geom &g = b.g;
        for (int q = 0; q < 4; q++)
            if (!b.spare[q]) {
                LCS_HIP(hipMalloc(&b.spare[q], (size_t)g.len * b.esize));
                b.bytes += (size_t)g.len * b.esize;
            }
        BoxRec &r = B.tab_h[i];
        r.g = g;
        r.S = b.S;
        r.halt = B.halt_d;
        r.part = b.part;
        r.Re = b.Re();
        r.power = B.cfgs[i].forcing.kind == LCS_FORCING_LINEAR_POWER ? B.cfgs[i].forcing.power : 0.0;
        if (b.esize == 8)
            fill_coefs<double>(b, B.hs, r);
        else
            fill_coefs<float>(b, B.hs, r);
        // fused RHS tiles of the class's shape, z-chunks of about 24 planes
        // (each re-reads six halo planes; 64-plane chunks were 0.2 ms a step slower)
        const int rq = rhs_shape(B, g);
        const int rtx = batch_rhs_shape(b.esize, rq)[0], rty = batch_rhs_shape(b.esize, rq)[1];
        const int rgx = (g.nx + rtx - 1) / rtx, rgy = (g.ny + rty - 1) / rty;
        const int rgz = std::max(1, g.nz / 24);
        r.rhs_zc = (g.nz + rgz - 1) / rgz;
        for (int z = 0; z * r.rhs_zc < g.nz; z++)
            for (int y = 0; y < rgy; y++)
                for (int x = 0; x < rgx; x++) rtiles[rq].push_back({i, x, y, z});
        // fused HSMAC sweep tiles over the whole array, in z-chunks (by default
        // of about 28 planes: short chunks balance the launch across boxes of
        // different sizes despite their five re-read halo planes; 64-plane
        // chunks: 25.1 ms a step, 12-plane chunks: 24.1 ms)
        const int zplanes = g.nz + 2 * g.h;
        const std::array<int, 2> tiling = sweep_tiling(B, g);
        const int shape = tiling[0];
        r.hs_tx = kBatchSweepShapes[shape][0];
        r.hs_ty = kBatchSweepShapes[shape][1];
        const int hgx = (g.nx + 2 * g.h + r.hs_tx - 1) / r.hs_tx, hgy = (g.ny + 2 * g.h + r.hs_ty - 1) / r.hs_ty;
        const int hgz = std::max(1, zplanes / tiling[1]);
        r.hs_zc = (zplanes + hgz - 1) / hgz;
        int nz_used = (zplanes + r.hs_zc - 1) / r.hs_zc;
        r.hs_tiles = hgx * hgy * nz_used;
        for (int z = 0; z < nz_used; z++)
            for (int y = 0; y < hgy; y++)
                for (int x = 0; x < hgx; x++) htiles[shape].push_back({i, x, y, z});
        // face passes of the overlap order; the x slabs as the pair when every box can
        for (int a = 0; a < 3; a++) {
            Box3 f0, f1;
            hs_faces(g, domain, a, f0, f1);
            r.face_blocks[a] = std::max(1, blocks_of(pass_count(f0) + pass_count(f1)));
            if (a == 2 && B.xface) continue;
            for (int k = 0; k < r.face_blocks[a]; k++) faces[a].push_back({i, k});
        }
        r.xf = nullptr;
        if (B.xface) {
            Box3 xs[2];
            hs_faces(g, domain, 2, xs[0], xs[1]);
            const size_t need = (size_t)kXfVals * SlabAxes<0>::rows(xs[0]) * b.esize;
            if (b.xf_bytes < need) {
                if (b.xf) {
                    LCS_HIP(hipStreamSynchronize(b.stream));
                    LCS_HIP(hipFree(b.xf));
                    b.bytes -= b.xf_bytes;
                }
                b.xf = nullptr;
                b.xf_bytes = 0;
                LCS_HIP(hipMalloc(&b.xf, need));
                b.xf_bytes = need;
                b.bytes += need;
            }
            r.xf = b.xf;
            const int pn = b.esize == 8 ? XfPatch<double>::N : XfPatch<float>::N;
            r.face_blocks[2] = 2 * ((g.ny - 1 + pn - 1) / pn) * ((g.nz - 1 + pn - 1) / pn);
            for (int k = 0; k < r.face_blocks[2]; k++) xfw.push_back({i, k});
            const long cb = (SlabAxes<0>::rows(xs[0]) + 255) / 256;
            for (long k = 0; k < cb; k++) xfc.push_back({i, (int)k});
        }
        // ghost fills (their cell counts must fit the kernel's 32-bit indices)
        {
            const long X = g.nx + 2L * g.h, Y = g.ny + 2L * g.h;
            if (2L * g.h * (X * Y + g.nz * X + (long)g.ny * g.nz) >= (1L << 31))
                return fail(LCS_ERR_INVALID, "box too large for the batched ghost fill");
        }
        auto ghost = [&](int rr, int full, std::vector<WorkItem> &w) {
            const long X = full ? g.nx + 2 * rr : g.nx, Y = full ? g.ny + 2 * rr : g.ny;
            const long tot = 2L * rr * X * Y + 2L * rr * g.nz * X + 2L * rr * g.ny * g.nz;
            for (int k = 0; k < blocks_of(tot); k++) w.push_back({i, k});
        };
        ghost(g.h, 1, gfull);
        ghost(1, 1, gone);
        // row-walking reductions
        r.row_blocks = std::min(g.ny * g.nz, kBatchRowBlocks);
        for (int k = 0; k < r.row_blocks; k++) rows.push_back({i, k});
        r.lf_force = B.cfgs[i].forcing.kind == LCS_FORCING_LINEAR_POWER;
        if (B.cfgs[i].forcing.kind == LCS_FORCING_LINEAR_POWER) {
            lf_boxes.push_back(i);
            for (int k = 0; k < r.row_blocks; k++) lf_rows.push_back({i, k});
        } else if (B.cfgs[i].forcing.kind == LCS_FORCING_RC_ENERGY) {
            const lcs_forcing &f = B.cfgs[i].forcing;
            int rc = rcf_prepare(b, f.kmin, f.kmax, f.coarse_size);
            if (rc) return rc;
            B.rcf_boxes.push_back(i);
            r.rcf_buf = b.rcf;
            r.rcf_modes = b.rcf_modes;
            r.rcf_energy = f.energy;
            r.rcf_c = f.coarse_size;
            r.rcf_kmin = f.kmin;
            r.rcf_kmax = f.kmax;
            r.rcf_blocks = rcf_apply_blocks(b);
            const int c = f.coarse_size;
            for (int k = 0; k < 3 * c * c * c; k++) rmean.push_back({i, k});
            for (int k = 0; k < rcf_crop_slots(f.kmax); k++) rdft.push_back({i, k});
            for (int k = 0; k < r.rcf_blocks; k++) rapply.push_back({i, k});
        }
        // particles
        r.nprof = (int)b.profiles.size();
        int blk = 0, dblk = 0;
        for (int q = 0; q < r.nprof; q++) {
            const ParticleProfile &pp = b.profiles[q];
            DevProfile &p = r.prof[q];
            if (b.esize == 8)
                fill_profile<double>(b, pp, p);
            else
                fill_profile<float>(b, pp, p);
            p.blk0 = blk;
            p.nblk = blocks_of(pp.desc.nvalid);
            p.dblk0 = dblk;
            p.dnblk = std::max(1, blocks_of(pp.desc.ntotal));
            blk += p.nblk;
            dblk += p.dnblk;
        }
        for (int k = 0; k < blk; k++) parts.push_back({i, k});
        for (int k = 0; k < dblk; k++) diams.push_back({i, k});
        // every count of per-block partials must fit one partial slot
        int most = std::max(r.hs_tiles, r.row_blocks);
        for (int a = 0; a < 3; a++) most = std::max(most, r.face_blocks[a]);
        for (int q = 0; q < r.nprof; q++) most = std::max(most, r.prof[q].dnblk);
        if (most > kMaxPartBlocks) return fail(LCS_ERR_CAPACITY, "box too large for a batch");
        refresh_record(B, i);
    }
    int rc;
    if ((rc = upload_table(B))) return rc;
    if ((rc = B.ghost_full.assign(xcd_order(gfull), B.stream)) ||
        (rc = B.ghost_one.assign(xcd_order(gone), B.stream)) || (rc = B.rows.assign(xcd_order(rows), B.stream)) ||
        (rc = B.lf_rows.assign(xcd_order(lf_rows), B.stream)) || (rc = B.lf_boxes.assign(lf_boxes, B.stream)) ||
        (rc = B.parts.assign(xcd_order(parts), B.stream)) || (rc = B.diams.assign(xcd_order(diams), B.stream)))
        return rc;
    for (int q = 0; q < 4; q++)
        if ((rc = B.hs_tiles[q].assign(xcd_order(htiles[q]), B.stream))) return rc;
    for (int q = 0; q < kBatchRhsShapesMax; q++)
        if ((rc = B.rhs_tiles[q].assign(xcd_order(rtiles[q]), B.stream))) return rc;
    for (int a = 0; a < 3; a++)
        if ((rc = B.faces[a].assign(xcd_order(faces[a]), B.stream))) return rc;
    if ((rc = B.xf_work.assign(xcd_order(xfw), B.stream)) || (rc = B.xf_commit.assign(xcd_order(xfc), B.stream)))
        return rc;
    if ((rc = B.rcf_boxes_d.assign(B.rcf_boxes, B.stream)) || (rc = B.rcf_mean.assign(xcd_order(rmean), B.stream)) ||
        (rc = B.rcf_dft.assign(xcd_order(rdft), B.stream)) || (rc = B.rcf_apply.assign(xcd_order(rapply), B.stream)))
        return rc;
    LCS_HIP(hipStreamSynchronize(B.stream));  // host vectors above are temporaries
    B.dirty = false;
    return LCS_OK;
}

// ---- HSMAC over the whole batch (couple!: src/flows/couples/hsmac.jl:52-201) --
static int b_fill(Batch &B, int set, int narr, int width, int hsmac)
{
    DevVec<WorkItem> &w = width == 1 ? B.ghost_one : B.ghost_full;
    return launch_b_ghost_fill(B.stream, B.esize, B.tab.d, w.d, w.n, set, narr, width, 1, hsmac);
}

static int b_sweep(Batch &B, int set, int kind, int slot, double dtfac)
{
    int rc;
    const bool force = B.force_next, wrap = B.wrap_next;
    B.force_next = B.wrap_next = false;
    for (int q = 0; q < 4; q++)
        if ((rc = launch_b_hsmac_sweep(B.stream, B.esize, B.hs.coloring, kBatchSweepShapes[q][0],
                                       kBatchSweepShapes[q][1], B.tab.d, B.hs_tiles[q].d, B.hs_tiles[q].n, set, kind,
                                       slot, dtfac, force, wrap, B.persist_cus, B.queue_d)))
            return rc;
    if (force && (rc = launch_b_lf_power(B.stream, B.tab.d, B.lf_boxes.d, B.lf_boxes.n))) return rc;
    return launch_b_hsmac_ctl(B.stream, B.tab.d, (int)B.boxes.size(), kHsFlip, 0, nullptr);
}

// core sweep + paired face passes (overlap order); returns the pass kinds
static int b_core_faces(Batch &B, int set, int domain, double dtfac, std::vector<int> &kinds)
{
    int rc = b_sweep(B, set, 0, 0, dtfac);
    if (rc) return rc;
    kinds.assign(1, 0);
    for (int a = 0; a < 3; a++) {
        if (a == 2 && B.xface) {
            const int slot = (int)kinds.size();
            rc = launch_b_hsmac_xface(B.stream, B.esize, B.hs.coloring, B.tab.d, B.xf_work.d, B.xf_work.n, set, domain,
                                      slot, dtfac);
            if (rc || (rc = launch_b_hsmac_xface_commit(B.stream, B.esize, B.tab.d, B.xf_commit.d, B.xf_commit.n, set,
                                                        domain)))
                return rc;
            kinds.push_back(1 + a);
            continue;
        }
        for (int c = 0; c < 2; c++) {
            const int slot = (int)kinds.size();
            rc = launch_b_hsmac_faces(B.stream, B.esize, B.hs.coloring, B.tab.d, B.faces[a].d, B.faces[a].n, set,
                                      domain, a, c, slot, dtfac);
            if (rc) return rc;
            kinds.push_back(1 + a);
        }
    }
    return LCS_OK;
}

static int b_check(Batch &B, const std::vector<int> &kinds)
{
    return launch_b_hsmac_check(B.stream, B.tab.d, (int)B.boxes.size(), kinds.data(), (int)kinds.size(), B.hs.epsp,
                                B.hs.itrmin);
}

static int b_poll(Batch &B, bool &all_done)
{
    const int n = (int)B.boxes.size();
    int rc = launch_b_hsmac_ctl(B.stream, B.tab.d, n, kHsPoll, 0, B.poll_d);
    if (rc) return rc;
    LCS_HIP(hipMemcpyAsync(B.poll_h, B.poll_d, sizeof(HsPoll) * (size_t)n, hipMemcpyDeviceToHost, B.stream));
    LCS_HIP(hipStreamSynchronize(B.stream));
    all_done = true;
    for (int i = 0; i < n; i++) all_done = all_done && B.poll_h[i].done;
    return LCS_OK;
}

// One iteration group: _couple_naive! does one sweep per check (hsmac.jl:77-129),
// _couple_time_blocking! three sweeps over shrinking regions after one fill
// (hsmac.jl:131-201).
static int b_group(Batch &B, int set, double dtfac)
{
    const lcs_hsmac_params &p = B.hs;
    std::vector<int> kinds;
    int rc;
    if (!p.time_blocking) {
        if ((rc = b_fill(B, set, p.overlap ? 3 : 4, 1, 1))) return rc;
        if (p.overlap) {
            if ((rc = b_core_faces(B, set, 3, dtfac, kinds))) return rc;
        } else {
            if ((rc = b_sweep(B, set, 3, 0, dtfac))) return rc;
            kinds.assign(1, 0);
        }
        return b_check(B, kinds);
    }
    // the FullHalo fill of U, V, W folds into the loads of the group's first
    // sweep where every box's ghosts have one image each
    if (B.wrap_ok)
        B.wrap_next = true;
    else if ((rc = b_fill(B, set, 3, B.boxes[0]->g.h, 1)))
        return rc;
    if (p.overlap) {
        if ((rc = b_core_faces(B, set, 1, dtfac, kinds))) return rc;
    } else {
        if ((rc = b_sweep(B, set, 1, 0, dtfac))) return rc;
        kinds.assign(1, 0);
    }
    if ((rc = b_check(B, kinds))) return rc;
    kinds.assign(1, 0);
    if ((rc = b_sweep(B, set, 2, 0, dtfac)) || (rc = b_check(B, kinds))) return rc;
    if ((rc = b_sweep(B, set, 3, 0, dtfac)) || (rc = b_check(B, kinds))) return rc;
    return LCS_OK;
}

// HSMAC groups of a call: all of them, and those before the first one that
// may end the iteration (itrmin), which run without a poll after them
static int hs_groups(const Batch &B) { return B.hs.time_blocking ? (B.hs.itrmax + 2) / 3 : B.hs.itrmax; }
static int hs_pre_groups(const Batch &B)
{
    const int ngroups = hs_groups(B);
    int pre = 0;
    while (pre < ngroups && (B.hs.time_blocking ? 3 * pre + 2 : pre) < B.hs.itrmin) pre++;
    return pre;
}

// The groups after the pre groups that a finished HSMAC call needed, from the
// poll of every box: a box's final itr counts its checks that did not
// converge, so it converged in group itr / 3 (time-blocked: three checks a
// group) or itr (naive); a box that never converged ran every group.
static int hs_post_needed(const Batch &B, const HsPoll *poll)
{
    const int ngroups = hs_groups(B), n = (int)B.boxes.size();
    int need = 0;
    for (int i = 0; i < n; i++) {
        const int g = poll[i].done ? (B.hs.time_blocking ? poll[i].itr / 3 : poll[i].itr) + 1 : ngroups;
        need = std::max(need, std::min(g, ngroups));
    }
    return std::max(0, need - hs_pre_groups(B));
}

// begin + the pre groups of a call on `set`
static int enqueue_hsmac_head(Batch &B, int set, double dtfac)
{
    const int pre = hs_pre_groups(B);
    int r = launch_b_hsmac_ctl(B.stream, B.tab.d, (int)B.boxes.size(), kHsBegin, 0, nullptr);
    for (int g = 0; g < pre && r == LCS_OK; g++) r = b_group(B, set, dtfac);
    return r;
}

// swaps the host mirror of the arrays of `set` for every box whose result
// lives in the spares (the device table was swapped by the commit kernel)
static void mirror_commit(Batch &B, int set, const HsPoll *poll)
{
    for (int i = 0; i < (int)B.boxes.size(); i++)
        if (poll[i].par) {
            Box &b = *B.boxes[i];
            for (int q = 0; q < 4; q++) {
                const int f = q < 3 ? 3 * set + q : LCS_FIELD_P;
                std::swap(b.f[f], b.spare[q]);
            }
            refresh_record(B, i);
        }
}

// Groups g0.. of a call on `set`, each one segment followed by a poll of every
// box's convergence flag, then the end, commit and host mirror swap.
static int b_hsmac_from(Batch &B, int set, double dtfac, int slot, int seg_id, int g0)
{
    const int n = (int)B.boxes.size();
    const int ngroups = hs_groups(B);
    int rc;
    bool done = false, polled_last = false;
    for (int g = g0; g < ngroups && !done; g++) {
        const int id = B.force_next ? seg_id + 2 : seg_id + 1;
        rc = segment(B, id, [&]() { return b_group(B, set, dtfac); });
        B.force_next = false;
        if (rc || (rc = b_poll(B, done))) return rc;
        polled_last = true;
    }
    if ((rc = launch_b_hsmac_ctl(B.stream, B.tab.d, n, kHsEnd, slot, nullptr))) return rc;
    if (!polled_last && (rc = b_poll(B, done))) return rc;
    // the final parity decides where each box's result lives: swap on the
    // device (commit) and in the host mirror
    if ((rc = launch_b_hsmac_ctl(B.stream, B.tab.d, n, kHsCommit, set, nullptr))) return rc;
    mirror_commit(B, set, B.poll_h);
    B.post_seen[set == 1 ? 0 : 1] = hs_post_needed(B, B.poll_h);
    return LCS_OK;
}

// set 1 acts on (U2, V2, W2, P) (stage 1), set 0 on (U, V, W, P) (stage 2).
// Groups before the first one that may end the iteration (itrmin) run
// unconditionally with the begin kernel in one segment; after that each group
// is one segment followed by a poll of every box's convergence flag.
static int b_hsmac(Batch &B, int set, double dtfac, int slot, int seg_id)
{
    const int pre = hs_pre_groups(B);
    int rc = segment(B, seg_id, [&]() { return enqueue_hsmac_head(B, set, dtfac); });
    if (rc) return rc;
    // The deferred linear forcing belongs to the first sweep that runs.  A
    // replayed segment does not pass through b_sweep, so the host flag is
    // cleared here for captured and eager runs alike.  If no group precedes
    // the first poll (small itrmin), the forced group is a segment of its
    // own: the group segment that later polls replay must never force.
    if (pre > 0) B.force_next = false;
    return b_hsmac_from(B, set, dtfac, slot, seg_id, pre);
}

// timestep! (CFL, particle dt) and stage 1 up to couple!; stage 1's FullHalo
// fill comes first and serves the CFL too (it writes every ghost a FaceHalo
// fill would, with the same values)
static int enqueue_stage1_prefix(Batch &B)
{
    const int n = (int)B.boxes.size(), h = B.boxes[0]->g.h;
    const lcs_step_config &c0 = B.cfgs[0];
    int r;
    if ((r = launch_b_ghost_fill(B.stream, B.esize, B.tab.d, B.ghost_full.d, B.ghost_full.n, 0, 4, h, 1, 0)) ||
        (r = launch_b_cfl(B.stream, B.esize, B.tab.d, B.rows.d, B.rows.n, n, c0.cfl, c0.dtmax)) ||
        (r = launch_b_particle_dt(B.stream, B.esize, B.tab.d, B.diams.d, B.diams.n, n, c0.dtmax)) ||
        (r = launch_b_particles(B.stream, B.esize, B.tab.d, B.parts.d, B.parts.n, 1)) ||
        (r = b_rhs(B, 1)))
        return r;
    return LCS_OK;
}

// stage 2 up to forcing!: the linear-power boxes' means (their forcing is
// applied by the first sweep of couple!), then the rc-energy boxes' forcing
static bool lf_deferred(const Batch &B) { return B.lf_boxes.n > 0 && B.hs.itrmax >= 1; }
static int enqueue_stage2_prefix(Batch &B)
{
    const int h = B.boxes[0]->g.h;
    const bool lf_defer = lf_deferred(B);
    int r;
    if ((r = launch_b_ghost_fill(B.stream, B.esize, B.tab.d, B.ghost_full.d, B.ghost_full.n, 1, 4, h, 1, 0)) ||
        (r = launch_b_particles(B.stream, B.esize, B.tab.d, B.parts.d, B.parts.n, 2)) ||
        (r = b_rhs(B, 2)))
        return r;
    if (B.lf_boxes.n > 0) {
        // the means read the x = 0, y = 0, z = 0 ghosts alone; where the
        // first sweep loads every ghost from its image, width 1 serves them
        const bool one = lf_defer && B.hs.time_blocking && B.wrap_ok;
        const DevVec<WorkItem> &gw = one ? B.ghost_one : B.ghost_full;
        if ((r = launch_b_ghost_fill(B.stream, B.esize, B.tab.d, gw.d, gw.n, 0, 3, one ? 1 : h, 1, 0)) ||
            (r = launch_b_lf(B.stream, B.esize, B.tab.d, B.lf_rows.d, B.lf_rows.n, B.lf_boxes.d, B.lf_boxes.n,
                             !lf_defer)))
            return r;
    }
    return launch_b_rcf(B.stream, B.esize, B.tab.d, B.rcf_mean.d, B.rcf_mean.n, B.rcf_dft.d, B.rcf_dft.n,
                        B.rcf_boxes_d.d, B.rcf_boxes_d.n, B.rcf_apply.d, B.rcf_apply.n);
}

static void fill_logs(Batch &B, lcs_step_log *logs, int nlogs)
{
    for (int i = 0; i < (int)B.boxes.size() && i < nlogs; i++) {
        const DevScalars &s = B.sc_h[i];
        lcs_step_log &l = logs[i];
        l.step = s.step;
        l.t = s.t;
        l.dt = s.dt;
        l.itr1 = s.itr[0];
        l.itr2 = s.itr[1];
        l.div1 = s.div[0];
        l.div2 = s.div[1];
        l.p = s.p;
        l.e10 = s.e10;
    }
}

// Stage 2 from its prefix to the end of the step (the segmented path).
static int step_from_stage2(Batch &B, lcs_step_log *logs, int nlogs)
{
    const int n = (int)B.boxes.size();
    int rc = segment(B, 2, [&]() { return enqueue_stage2_prefix(B); });
    if (rc) return rc;
    // the first sweep of stage 2's couple! applies the linear forcing as it
    // loads U, V, W (its coefficients were set above)
    B.force_next = lf_deferred(B);
    rc = b_hsmac(B, 0, 1.0, 1, 20);
    B.force_next = false;
    if (rc) return rc;
    if ((rc = launch_b_state(B.stream, B.tab.d, n, logs ? B.sc_d : nullptr))) return rc;
    if (logs) {
        LCS_HIP(hipMemcpyAsync(B.sc_h, B.sc_d, sizeof(DevScalars) * (size_t)n, hipMemcpyDeviceToHost, B.stream));
        LCS_HIP(hipStreamSynchronize(B.stream));
        fill_logs(B, logs, nlogs);
    }
    return LCS_OK;
}

// Post groups a step graph holds per stage: the most any of the last
// kPostWindow calls of that stage needed, and at least one where no group
// precedes the first check (so the forced first sweep of stage 2 is always
// inside the graph).  Mode 3 holds the least.
static std::array<int, 2> step_graph_groups(const Batch &B)
{
    const int pre = hs_pre_groups(B), room = hs_groups(B) - pre;
    std::array<int, 2> G = {0, 0};
    for (int s = 0; s < 2; s++) {
        if (B.graph_mode != 3)
            for (const auto &h : B.post_hist) G[s] = std::max(G[s], h[s]);
        G[s] = std::min(std::max(G[s], pre == 0 ? 1 : 0), room);
    }
    return G;
}

// One step as one captured graph: stage 1 (prefix, begin, pre groups, G1
// more groups), a gate that stops the graph if a box still needs groups,
// stage 1's end, commit and poll, stage 2 likewise, the state update, and the
// copies of both polls, the scalars and the stop flag to the host.  Returns
// the stop flag: 0, or the stage whose gate stopped the graph.
static int step_graph(Batch &B, const std::array<int, 2> &G, int &halted)
{
    const int n = (int)B.boxes.size();
    const int pre = hs_pre_groups(B), ngroups = hs_groups(B);
    auto it = B.step_graphs.find(G);
    if (it == B.step_graphs.end()) {
        LCS_HIP(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        auto enqueue = [&]() {
            int r;
            const std::array<int, 2> sets = {1, 0};
            for (int s = 0; s < 2; s++) {
                const int set = sets[s];
                const double dtfac = s == 0 ? 0.5 : 1.0;
                if ((r = s == 0 ? enqueue_stage1_prefix(B) : enqueue_stage2_prefix(B))) return r;
                B.force_next = s == 1 && lf_deferred(B);
                if ((r = enqueue_hsmac_head(B, set, dtfac))) return r;
                for (int g = 0; g < G[s]; g++)
                    if ((r = b_group(B, set, dtfac))) return r;
                B.force_next = false;
                if (pre + G[s] < ngroups && (r = launch_b_hsmac_ctl(B.stream, B.tab.d, n, kHsGate, s + 1, nullptr)))
                    return r;
                if ((r = launch_b_hsmac_ctl(B.stream, B.tab.d, n, kHsEnd, s, nullptr)) ||
                    (r = launch_b_hsmac_ctl(B.stream, B.tab.d, n, kHsCommit, set, nullptr)) ||
                    (r = launch_b_hsmac_ctl(B.stream, B.tab.d, n, kHsPoll, 0, B.poll_d + (long)s * n)))
                    return r;
            }
            if ((r = launch_b_state(B.stream, B.tab.d, n, B.sc_d))) return r;
            LCS_HIP(hipMemcpyAsync(B.poll_h, B.poll_d, sizeof(HsPoll) * 2 * (size_t)n, hipMemcpyDeviceToHost,
                                   B.stream));
            LCS_HIP(hipMemcpyAsync(B.sc_h, B.sc_d, sizeof(DevScalars) * (size_t)n, hipMemcpyDeviceToHost, B.stream));
            LCS_HIP(hipMemcpyAsync(B.halt_h, B.halt_d, sizeof(int), hipMemcpyDeviceToHost, B.stream));
            return LCS_OK;
        };
        const int rc = enqueue();
        B.force_next = B.wrap_next = false;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(B.stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(LCS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return fail(LCS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        it = B.step_graphs.emplace(G, x).first;
    }
    LCS_HIP(hipGraphLaunch(it->second, B.stream));
    LCS_HIP(hipStreamSynchronize(B.stream));
    halted = *B.halt_h;
    return LCS_OK;
}

static void record_post(Batch &B)
{
    B.post_hist.push_back(B.post_seen);
    while ((int)B.post_hist.size() > kPostWindow) B.post_hist.pop_front();
}

static int batch_step(Batch &B, lcs_step_log *logs, int nlogs)
{
    const int n = (int)B.boxes.size();
    if (n == 0) return LCS_OK;
    int rc;
    if (B.dirty && (rc = rebuild(B))) return rc;
    // storage order of particles for locality, per box on its own schedule
    bool moved = false;
    for (int i = 0; i < n; i++) {
        Box &b = *B.boxes[i];
        const int iv = B.cfgs[i].localize_interval;
        if (iv > 0 && b.steps_host % iv == 0 && !b.profiles.empty()) {
            for (auto &pp : b.profiles)
                if ((rc = launch_p_localize(b, pp))) return rc;
            refresh_record(B, i);
            moved = true;
        }
        b.steps_host++;
    }
    if (moved && (rc = upload_table(B))) return rc;
    const int pre = hs_pre_groups(B);
    // the whole step as one graph from the third step after a membership
    // change, once the post-group counts of a step are known
    if (B.graphs && B.graph_mode != 2 && B.stable >= 2 && !B.post_hist.empty()) {
        const std::array<int, 2> G = step_graph_groups(B);
        int halted = 0;
        if ((rc = step_graph(B, G, halted))) return rc;
        if (!halted) {
            B.n_whole++;
            mirror_commit(B, 1, B.poll_h);
            mirror_commit(B, 0, B.poll_h + n);
            B.post_seen = {hs_post_needed(B, B.poll_h), hs_post_needed(B, B.poll_h + n)};
            record_post(B);
            if (logs) fill_logs(B, logs, nlogs);
            return LCS_OK;
        }
        // a stage needs more groups than the graph held: every kernel after
        // its gate returned at once; go on from there with polled groups
        B.n_gated++;
        if ((rc = launch_b_hsmac_ctl(B.stream, B.tab.d, n, kHsUnhalt, 0, nullptr))) return rc;
        if (halted == 1) {
            if ((rc = b_hsmac_from(B, 1, 0.5, 0, 10, pre + G[0])) || (rc = step_from_stage2(B, logs, nlogs)))
                return rc;
        } else {
            std::vector<HsPoll> p1(B.poll_h, B.poll_h + n);  // stage 1 finished inside the graph
            mirror_commit(B, 1, p1.data());
            B.post_seen[0] = hs_post_needed(B, p1.data());
            B.force_next = false;  // the graph's first stage-2 sweep applied it
            if ((rc = b_hsmac_from(B, 0, 1.0, 1, 20, pre + G[1]))) return rc;
            if ((rc = launch_b_state(B.stream, B.tab.d, n, logs ? B.sc_d : nullptr))) return rc;
            if (logs) {
                LCS_HIP(hipMemcpyAsync(B.sc_h, B.sc_d, sizeof(DevScalars) * (size_t)n, hipMemcpyDeviceToHost,
                                       B.stream));
                LCS_HIP(hipStreamSynchronize(B.stream));
                fill_logs(B, logs, nlogs);
            }
        }
        record_post(B);
        return LCS_OK;
    }
    B.n_segmented++;
    rc = segment(B, 1, [&]() { return enqueue_stage1_prefix(B); });
    if (rc || (rc = b_hsmac(B, 1, 0.5, 0, 10))) return rc;
    if ((rc = step_from_stage2(B, logs, nlogs))) return rc;
    record_post(B);
    B.stable++;
    return LCS_OK;
}

// ---- sweep tiling per size class ----------------------------------------------
// Fastest of `reps` runs of the sweeps of one HSMAC group (time-blocked: the
// group's three sweeps; naive: one) over the whole batch with the current
// tiling.  Begin sets every box's parity to 0 and nothing flips it, so each
// sweep reads the boxes' current arrays and writes only their ping-pong
// partners, which hold no state between HSMAC calls: no box changes.
static int time_group_sweeps(Batch &B, int reps, hipEvent_t e0, hipEvent_t e1, float &best)
{
    int rc;
    if ((rc = rebuild(B))) return rc;
    const int n = (int)B.boxes.size();
    std::vector<int> kinds;
    if (B.hs.time_blocking)
        kinds = {B.hs.overlap ? 0 : 1, 2, 3};
    else
        kinds = {B.hs.overlap ? 0 : 3};
    auto run = [&]() {
        int r = launch_b_hsmac_ctl(B.stream, B.tab.d, n, kHsBegin, 0, nullptr);
        for (int k : kinds)
            for (int q = 0; q < 4 && r == LCS_OK; q++)
                r = launch_b_hsmac_sweep(B.stream, B.esize, B.hs.coloring, kBatchSweepShapes[q][0],
                                         kBatchSweepShapes[q][1], B.tab.d, B.hs_tiles[q].d, B.hs_tiles[q].n, 0, k, 0,
                                         1.0, false, false, B.persist_cus, B.queue_d);
        return r;
    };
    if ((rc = run())) return rc;
    best = 1e30f;
    for (int r = 0; r < reps; r++) {
        LCS_HIP(hipEventRecord(e0, B.stream));
        if ((rc = run())) return rc;
        LCS_HIP(hipEventRecord(e1, B.stream));
        LCS_HIP(hipEventSynchronize(e1));
        float ms = 0;
        LCS_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    return LCS_OK;
}

// Fastest of `reps` runs of stage 1's fused RHS over the whole batch with the
// current RHS tiling, written to the boxes' spare arrays: no box changes.
static int time_rhs(Batch &B, int reps, hipEvent_t e0, hipEvent_t e1, float &best)
{
    int rc;
    if ((rc = rebuild(B))) return rc;
    if ((rc = b_rhs(B, 1, true))) return rc;
    best = 1e30f;
    for (int r = 0; r < reps; r++) {
        LCS_HIP(hipEventRecord(e0, B.stream));
        if ((rc = b_rhs(B, 1, true))) return rc;
        LCS_HIP(hipEventRecord(e1, B.stream));
        LCS_HIP(hipEventSynchronize(e1));
        float ms = 0;
        LCS_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    return LCS_OK;
}

// Coordinate descent over size classes: each class in turn tries every shape
// and chunk target with the other classes at their current choice, and keeps
// a candidate only if it beats the current one by more than 1 %.  Then the
// same for the fused RHS's tile shape.
static int batch_autotune(Batch &B)
{
    if (B.boxes.empty()) return LCS_OK;
    LCS_HIP(hipStreamSynchronize(B.stream));
    std::vector<std::array<int, 3>> classes;
    for (const Box *b : B.boxes) {
        const std::array<int, 3> c = {b->g.nx, b->g.ny, b->g.nz};
        if (std::find(classes.begin(), classes.end(), c) == classes.end()) classes.push_back(c);
        if (!B.sweep_tune.count(c)) B.sweep_tune[c] = {1, kBatchSweepChunk};
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(LCS_ERR_HIP, "hipEventCreate failed");
    }
    int rc = LCS_OK;
    const int reps = 5;
    for (int round = 0; round < 2 && rc == LCS_OK; round++)
        for (const auto &c : classes) {
            std::array<int, 2> best = B.sweep_tune[c];
            float best_ms = 0;
            B.dirty = true;
            if ((rc = time_group_sweeps(B, reps, e0, e1, best_ms))) break;
            for (int q = 0; q < 4 && rc == LCS_OK; q++)
                for (int z : kBatchSweepChunks) {
                    const std::array<int, 2> cand = {q, z};
                    if (cand == best) continue;
                    B.sweep_tune[c] = cand;
                    B.dirty = true;
                    float ms = 0;
                    // a tiling whose per-box tile count does not fit a partial
                    // slot is skipped, not an error
                    rc = time_group_sweeps(B, reps, e0, e1, ms);
                    if (rc == LCS_ERR_CAPACITY) {
                        rc = LCS_OK;
                        continue;
                    }
                    if (rc) break;
                    if (ms < 0.99f * best_ms) {
                        best = cand;
                        best_ms = ms;
                    }
                }
            B.sweep_tune[c] = best;
        }
    for (int round = 0; round < 2 && rc == LCS_OK; round++)
        for (const auto &c : classes) {
            if (!B.rhs_tune.count(c)) B.rhs_tune[c] = 0;
            int best = B.rhs_tune[c];
            float best_ms = 0;
            B.dirty = true;
            if ((rc = time_rhs(B, reps, e0, e1, best_ms))) break;
            for (int q = 0; q < batch_rhs_shapes(B.esize) && rc == LCS_OK; q++) {
                if (q == best) continue;
                B.rhs_tune[c] = q;
                B.dirty = true;
                float ms = 0;
                if ((rc = time_rhs(B, reps, e0, e1, ms))) break;
                if (ms < 0.99f * best_ms) {
                    best = q;
                    best_ms = ms;
                }
            }
            B.rhs_tune[c] = best;
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B.dirty = true;  // the next step rebuilds the tables (and drops the graphs)
    if (rc == LCS_OK && (rc = rebuild(B))) return rc;
    return rc;
}

static bool same_hsmac(const lcs_hsmac_params &a, const lcs_hsmac_params &b)
{
    return a.time_blocking == b.time_blocking && a.overlap == b.overlap && a.coloring == b.coloring &&
           a.omega == b.omega && a.epsp == b.epsp && a.itrmin == b.itrmin && a.itrmax == b.itrmax;
}

static int detach(Batch &B, int i)
{
    LCS_HIP(hipStreamSynchronize(B.stream));
    Box &b = *B.boxes[i];
    b.stream = B.own_streams[i];
    b.batch = nullptr;
    B.boxes.erase(B.boxes.begin() + i);
    B.cfgs.erase(B.cfgs.begin() + i);
    B.own_streams.erase(B.own_streams.begin() + i);
    B.dirty = true;
    return LCS_OK;
}

int batch_detach_box(Box &b)
{
    Batch *B = b.batch;
    if (!B) return LCS_OK;
    for (size_t i = 0; i < B->boxes.size(); i++)
        if (B->boxes[i] == &b) return detach(*B, (int)i);
    return LCS_OK;
}

void batch_mark_dirty(Box &b)
{
    if (b.batch) b.batch->dirty = true;
}

}  // namespace lcs

using namespace lcs;

static Batch *BT(lcs_batch *p) { return reinterpret_cast<Batch *>(p); }

extern "C" {

int lcs_batch_create(lcs_ctx *ctx, lcs_batch **out)
{
    if (!ctx || !out) return fail(LCS_ERR_INVALID, "null argument");
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    LCS_HIP(hipSetDevice(c->device));
    Batch *B = new Batch;
    B->ctx = c;
    if (hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || B->cus < 1)
        B->cus = 256;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        return fail(LCS_ERR_HIP, "hipStreamCreate failed");
    }
    if (hipMalloc(&B->halt_d, sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&B->halt_h, sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipMemset(B->halt_d, 0, sizeof(int)) != hipSuccess || hipMalloc(&B->queue_d, 9 * sizeof(int)) != hipSuccess ||
        hipMemset(B->queue_d, 0, 9 * sizeof(int)) != hipSuccess) {
        if (B->halt_d) (void)hipFree(B->halt_d);
        if (B->halt_h) (void)hipHostFree(B->halt_h);
        if (B->queue_d) (void)hipFree(B->queue_d);
        (void)hipStreamDestroy(B->stream);
        delete B;
        return fail(LCS_ERR_HIP, "stop flag allocation failed");
    }
    *B->halt_h = 0;
    *out = reinterpret_cast<lcs_batch *>(B);
    return LCS_OK;
}

int lcs_batch_destroy(lcs_batch *batch)
{
    if (!batch) return LCS_OK;
    Batch *B = BT(batch);
    (void)hipSetDevice(B->ctx->device);
    drop_graphs(*B);
    while (!B->boxes.empty()) (void)detach(*B, (int)B->boxes.size() - 1);
    DevVec<WorkItem> *w[] = {&B->ghost_full, &B->ghost_one, &B->rows,     &B->lf_rows,  &B->parts,
                             &B->diams,      &B->faces[0],  &B->faces[1], &B->faces[2], &B->xf_work,
                             &B->xf_commit};
    for (auto *v : w) v->release();
    B->lf_boxes.release();
    B->rcf_boxes_d.release();
    B->rcf_mean.release();
    B->rcf_dft.release();
    B->rcf_apply.release();
    for (auto &t : B->rhs_tiles) t.release();
    for (auto &t : B->hs_tiles) t.release();
    B->tab.release();
    if (B->poll_d) (void)hipFree(B->poll_d);
    if (B->poll_h) (void)hipHostFree(B->poll_h);
    if (B->sc_d) (void)hipFree(B->sc_d);
    if (B->sc_h) (void)hipHostFree(B->sc_h);
    if (B->halt_d) (void)hipFree(B->halt_d);
    if (B->halt_h) (void)hipHostFree(B->halt_h);
    if (B->queue_d) (void)hipFree(B->queue_d);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
    return LCS_OK;
}

int lcs_batch_add(lcs_batch *batch, lcs_box *box, const lcs_step_config *cfg)
{
    if (!batch || !box || !cfg) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    Box &b = *reinterpret_cast<Box *>(box);
    if (b.batch) return fail(LCS_ERR_STATE, "box is already in a batch");
    if (b.ctx != B.ctx) return fail(LCS_ERR_INVALID, "box and batch belong to different contexts");
    if (b.g.h < 3) return fail(LCS_ERR_INVALID, "a flow step needs halo_size >= 3");
    if (b.g.nx < 8 || b.g.ny < 8 || b.g.nz < 8) return fail(LCS_ERR_INVALID, "batched boxes need n >= 8 per axis");
    if (cfg->forcing.kind != LCS_FORCING_NONE && cfg->forcing.kind != LCS_FORCING_LINEAR_POWER &&
        cfg->forcing.kind != LCS_FORCING_RC_ENERGY)
        return fail(LCS_ERR_INVALID, "unknown forcing kind");
    if (cfg->hsmac.coloring != LCS_COLORING_RED_BLACK_FAST && cfg->hsmac.coloring != LCS_COLORING_RED_BLACK)
        return fail(LCS_ERR_INVALID, "unsupported coloring");
    if (cfg->localize_interval < 0) return fail(LCS_ERR_INVALID, "localize_interval must be >= 0");
    if (!B.boxes.empty()) {
        if (b.esize != B.esize) return fail(LCS_ERR_INVALID, "all boxes of a batch share one float type");
        const lcs_step_config &c0 = B.cfgs[0];
        if (!same_hsmac(cfg->hsmac, B.hs) || cfg->cfl != c0.cfl || cfg->dtmax != c0.dtmax || b.g.h != B.boxes[0]->g.h)
            return fail(LCS_ERR_INVALID, "boxes of a batch share HSMAC parameters, cfl, dtmax and halo size");
    } else {
        B.esize = b.esize;
        B.hs = cfg->hsmac;
    }
    LCS_HIP(hipStreamSynchronize(b.stream));
    B.boxes.push_back(&b);
    B.cfgs.push_back(*cfg);
    B.own_streams.push_back(b.stream);
    b.stream = B.stream;  // every launcher of the box now enqueues on the batch stream
    b.batch = &B;
    B.dirty = true;
    return LCS_OK;
}

int lcs_batch_remove(lcs_batch *batch, lcs_box *box)
{
    if (!batch || !box) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    Box *b = reinterpret_cast<Box *>(box);
    for (size_t i = 0; i < B.boxes.size(); i++)
        if (B.boxes[i] == b) return detach(B, (int)i);
    return fail(LCS_ERR_INVALID, "box is not in this batch");
}

int lcs_batch_size(lcs_batch *batch, int *nboxes)
{
    if (!batch || !nboxes) return fail(LCS_ERR_INVALID, "null argument");
    *nboxes = (int)BT(batch)->boxes.size();
    return LCS_OK;
}

int lcs_batch_step(lcs_batch *batch, lcs_step_log *logs, int nlogs)
{
    if (!batch) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    LCS_HIP(hipSetDevice(B.ctx->device));
    return batch_step(B, logs, nlogs);
}

int lcs_batch_set_graphs(lcs_batch *batch, int on)
{
    if (!batch) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    if (on < 0 || on > 3) return fail(LCS_ERR_INVALID, "graph mode must be 0, 1, 2 or 3");
    LCS_HIP(hipStreamSynchronize(B.stream));
    drop_graphs(B);
    B.graphs = on != 0;
    if (on) B.graph_mode = on;
    return LCS_OK;
}

int lcs_batch_set_persistent(lcs_batch *batch, int cus)
{
    if (!batch) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    LCS_HIP(hipStreamSynchronize(B.stream));
    drop_graphs(B);  // captured launches keep the grid they were captured with
    B.persist_cus = cus < 0 ? B.cus : std::min(cus, B.cus);
    return LCS_OK;
}

int lcs_batch_graph_stats(lcs_batch *batch, int64_t *whole, int64_t *gated, int64_t *segmented)
{
    if (!batch) return fail(LCS_ERR_INVALID, "null argument");
    const Batch &B = *BT(batch);
    if (whole) *whole = B.n_whole;
    if (gated) *gated = B.n_gated;
    if (segmented) *segmented = B.n_segmented;
    return LCS_OK;
}

int lcs_batch_autotune(lcs_batch *batch)
{
    if (!batch) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    LCS_HIP(hipSetDevice(B.ctx->device));
    return batch_autotune(B);
}

int lcs_batch_set_tiles(lcs_batch *batch, int index, int tx, int ty, int zchunk)
{
    if (!batch) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    if (index < 0 || index >= (int)B.boxes.size()) return fail(LCS_ERR_INVALID, "box index out of range");
    int q = -1;
    for (int k = 0; k < 4; k++)
        if (kBatchSweepShapes[k][0] == tx && kBatchSweepShapes[k][1] == ty) q = k;
    if (q < 0) return fail(LCS_ERR_INVALID, "sweep tile shape must be 64x8, 32x8, 64x16 or 32x16");
    if (zchunk < 1) return fail(LCS_ERR_INVALID, "zchunk must be >= 1");
    const Geom &g = B.boxes[index]->g;
    LCS_HIP(hipStreamSynchronize(B.stream));
    B.sweep_tune[{g.nx, g.ny, g.nz}] = {q, zchunk};
    B.dirty = true;
    return LCS_OK;
}

int lcs_batch_get_tiles(lcs_batch *batch, int index, int *tx, int *ty, int *zchunk)
{
    if (!batch || !tx || !ty || !zchunk) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    if (index < 0 || index >= (int)B.boxes.size()) return fail(LCS_ERR_INVALID, "box index out of range");
    LCS_HIP(hipSetDevice(B.ctx->device));
    int rc;
    if (B.dirty && (rc = rebuild(B))) return rc;
    const BoxRec &r = B.tab_h[index];
    *tx = r.hs_tx;
    *ty = r.hs_ty;
    *zchunk = r.hs_zc;
    return LCS_OK;
}

int lcs_batch_set_rhs_tiles(lcs_batch *batch, int index, int tx, int ty)
{
    if (!batch) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    if (index < 0 || index >= (int)B.boxes.size()) return fail(LCS_ERR_INVALID, "box index out of range");
    int q = -1;
    for (int k = 0; k < batch_rhs_shapes(B.esize); k++)
        if (batch_rhs_shape(B.esize, k)[0] == tx && batch_rhs_shape(B.esize, k)[1] == ty) q = k;
    if (q < 0)
        return fail(LCS_ERR_INVALID, B.esize == 8 ? "FP64 batched RHS tile shape must be 32x4, 32x8 or 64x8"
                                                  : "FP32 batched RHS tile shape must be 32x8, 64x8, 64x16 or 32x16");
    const Geom &g = B.boxes[index]->g;
    LCS_HIP(hipStreamSynchronize(B.stream));
    B.rhs_tune[{g.nx, g.ny, g.nz}] = q;
    B.dirty = true;
    return LCS_OK;
}

int lcs_batch_get_rhs_tiles(lcs_batch *batch, int index, int *tx, int *ty)
{
    if (!batch || !tx || !ty) return fail(LCS_ERR_INVALID, "null argument");
    Batch &B = *BT(batch);
    if (index < 0 || index >= (int)B.boxes.size()) return fail(LCS_ERR_INVALID, "box index out of range");
    const int *sh = batch_rhs_shape(B.esize, rhs_shape(B, B.boxes[index]->g));
    *tx = sh[0];
    *ty = sh[1];
    return LCS_OK;
}

int lcs_batch_sync(lcs_batch *batch)
{
    if (!batch) return fail(LCS_ERR_INVALID, "null argument");
    LCS_HIP(hipStreamSynchronize(BT(batch)->stream));
    return LCS_OK;
}

}  // extern "C"
