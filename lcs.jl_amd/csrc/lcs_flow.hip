// Flow kernels of the MI355X backend (gfx950, wave64): periodic ghost fill,
// staggered 4th-order RHS (+ fused RK update), HSMAC colour passes with a
// device-side convergence test, CFL reduction, linear-power and rc-energy
// forcing, and the builder-defined spectrum initialiser.
//
// All kernels are templated on the field type T (float or double).  Scalars
// that later phases need (dt, HSMAC state, forcing means) live in the box's
// DevScalars block so a step never has to return to the host.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <map>
#include <mutex>

#include "lcs_internal.h"
#include "lcs_hsmac_tile.h"
#include "lcs_hsmac_tile2.h"
#include "lcs_hsmac_xface.h"
#include "lcs_rhs_tile.h"

namespace lcs {

// Launch set-up of a kernel with more dynamic LDS than the default: its LDS
// limit is raised and its resident workgroups per CU measured once per
// (kernel, device).
int kernel_setup(const void *func, int threads, size_t lds, int *per_cu)
{
    static std::mutex mu;
    static std::map<std::pair<const void *, int>, int> done;
    int dev = 0;
    LCS_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_pair(func, dev);
    auto it = done.find(key);
    if (it == done.end()) {
        LCS_HIP(hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int pc = 0;
        LCS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, func, threads, lds));
        it = done.emplace(key, std::max(pc, 1)).first;
    }
    if (per_cu) *per_cu = it->second;
    return LCS_OK;
}

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ inline int wrap1(int i, int n)
{
    int r = (i - 1) % n;
    if (r < 0) r += n;
    return r + 1;
}

// Sum / max of a double over a whole block (any shape, wave64), deterministic
// for a fixed block shape.  The result is valid in linear thread 0.
__device__ inline int lin_tid() { return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z); }
__device__ inline int block_threads() { return blockDim.x * blockDim.y * blockDim.z; }

__device__ inline double block_sum(double v, double *lds)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int t = lin_tid(), w = t >> 6, l = t & 63;
    if (l == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (t == 0) {
        int nw = (block_threads() + 63) >> 6;
        for (int q = 0; q < nw; q++) s += lds[q];
    }
    __syncthreads();
    return s;
}

__device__ inline double block_max(double v, double *lds)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    const int t = lin_tid(), w = t >> 6, l = t & 63;
    if (l == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (t == 0) {
        int nw = (block_threads() + 63) >> 6;
        s = lds[0];
        for (int q = 1; q < nw; q++) s = fmax(s, lds[q]);
    }
    __syncthreads();
    return s;
}

// Workgroups are dispatched round-robin over the XCDs (linear id L runs on
// XCD L % kXcds), each with its own L2.  Tile kernels take their tile from
// this map instead of blockIdx, so each XCD works on one contiguous range of
// tiles (x fastest) and neighbouring tiles share halo rows through one L2.
// The map is a bijection for any kXcds, so results never depend on it.
constexpr int kXcds = 8;
__device__ inline void xcd_tile(int &bx, int &by, int &bz)
{
    const int gx = gridDim.x, gy = gridDim.y;
    const int n = gx * gy * gridDim.z, L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int per = n / kXcds, rem = n % kXcds, x = L % kXcds, i = L / kXcds;
    const int t = x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
    bx = t % gx;
    by = (t / gx) % gy;
    bz = t / (gx * gy);
}

// Visits every interior cell once.  Each block walks one contiguous run of
// rows (j, k) with x over its threads, so the y/z neighbours a stencil reads
// were just read by the same block or by a block on the same XCD (runs are
// handed out XCD by XCD).  Rows shorter than half the block go several at a
// time (thread t takes row t / nx of each group, cell t % nx), so a 32-wide
// row does not leave seven eighths of a 256-thread block idle.
template <typename F>
__device__ inline void for_rows(const Geom &g, int b, int nb, F &&f)
{
    const int rows = g.ny * g.nz;
    const int per = (rows + nb - 1) / nb, r0 = b * per, r1 = min(rows, r0 + per);
    const int nt = (int)blockDim.x, R = g.nx <= nt / 2 ? nt / g.nx : 1;
    if (R > 1) {
        const int dr = (int)threadIdx.x / g.nx, i = 1 + (int)threadIdx.x % g.nx;
        if (dr >= R) return;
        for (int r = r0 + dr; r < r1; r += R) {
            const int j = r % g.ny + 1, k = r / g.ny + 1;
            f(i, j, k, g.at(i, j, k));
        }
        return;
    }
    int j = r0 % g.ny + 1, k = r0 / g.ny + 1;
    for (int r = r0; r < r1; r++) {
        const long base = g.at(0, j, k);
        for (int i = 1 + (int)threadIdx.x; i <= g.nx; i += nt) f(i, j, k, base + i);
        if (++j > g.ny) {
            j = 1;
            k++;
        }
    }
}

template <typename F>
__device__ inline void for_interior(const Geom &g, F &&f)
{
    const int nb = gridDim.x;
    const int per_x = nb / kXcds, rem = nb % kXcds, x = blockIdx.x % kXcds, i0 = blockIdx.x / kXcds;
    const int b = x < rem ? x * (per_x + 1) + i0 : rem * (per_x + 1) + (x - rem) * per_x + i0;
    for_rows(g, b, nb, f);
}

// blocks for a for_interior launch: every row gets a block up to the cap
static int grid_rows(const Geom &g, int cap = 4096) { return (int)std::max(1L, std::min<long>((long)g.ny * g.nz, cap)); }

static int grid_for(long work, int block, int cap = 4096)
{
    long g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// ---------------------------------------------------------------------------
// Stencils of src/diff.jl:1-31 (same operand order; the compiler may contract
// multiply-adds, which changes results only at rounding level).
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T interp_o2(T a, T b)
{
    return T(0.5) * (a + b);
}
template <typename T>
__device__ inline T interp_o4(T a, T b, T c, T d)
{
    return T(1.0 / 16.0) * (((-a + T(9) * b) + T(9) * c) - d);
}
template <typename T>
__device__ inline T diff2_o4(T a, T b, T c, T d, T e)
{
    return T(1.0 / 12.0) * ((((-a + T(16) * b) - T(30) * c) + T(16) * d) - e);
}
template <typename T>
__device__ inline T sum_flux_o4(T j1a, T j1b, T j3a, T j3b)
{
    return T(9.0 / 8.0) * (-j1a + j1b) - T((1.0 / 8.0) * (1.0 / 3.0)) * (-j3a + j3b);
}

// ---------------------------------------------------------------------------
// Periodic ghost fill (lib/topologies/src/halo.jl:138-159 single process):
// A[i] = A[mod1(i, n)] for every ghost cell of width r; FaceHalo (full=0) only
// cells with exactly one axis outside 1..n.  One thread per ghost cell,
// enumerated as z-slabs, then y-slabs, then x-slabs; up to four arrays.
// ---------------------------------------------------------------------------
template <typename T>
struct Ptr4 {
    T *p[4];
};

// Inside an HSMAC call (S != nullptr) the fill is skipped once the call has
// converged and acts on the spare buffer set B when the ping-pong parity says
// the current state lives there.
// Ghost cell t of the fill (z-slabs, then y-slabs, then x-slabs): its
// offset d and the offset s of its periodic source.  The counts of a fill fit
// 32 bits (the launchers check), so the index arithmetic is 32-bit.
__device__ inline void ghost_cell(const Geom &g, int r, int full, int cz, int cy, int t, long &d, long &s)
{
    const int X = full ? g.nx + 2 * r : g.nx;
    const int Y = full ? g.ny + 2 * r : g.ny;
    const int x0 = full ? 1 - r : 1, y0 = full ? 1 - r : 1;
    const int r2 = 2 * r;
    int i, j, k;
    if (t < cz) {
        int ii = t % X;
        int q = t / X;
        int jj = q % Y;
        int kk = q / Y;
        i = x0 + ii;
        j = y0 + jj;
        k = kk < r ? 1 - r + kk : g.nz + 1 + kk - r;
    } else if (t < cz + cy) {
        t -= cz;
        int ii = t % X;
        int q = t / X;
        int jj = q % r2;
        int kk = q / r2;
        i = x0 + ii;
        j = jj < r ? 1 - r + jj : g.ny + 1 + jj - r;
        k = 1 + kk;
    } else {
        t -= cz + cy;
        int ii = t % r2;
        int q = t / r2;
        int jj = q % g.ny;
        int kk = q / g.ny;
        i = ii < r ? 1 - r + ii : g.nx + 1 + ii - r;
        j = 1 + jj;
        k = 1 + kk;
    }
    d = g.at(i, j, k);
    s = g.at(wrap1(i, g.nx), wrap1(j, g.ny), wrap1(k, g.nz));
}

// ghost-cell counts of the z-, y- and x-slabs of a fill of width r
__host__ __device__ inline void ghost_counts(const Geom &g, int r, int full, long &cz, long &cy, long &cx)
{
    const long X = full ? g.nx + 2 * r : g.nx, Y = full ? g.ny + 2 * r : g.ny;
    cz = 2L * r * X * Y;
    cy = 2L * r * g.nz * X;
    cx = 2L * r * g.ny * g.nz;
}

template <typename T>
__global__ void k_ghost_fill(Geom g, Ptr4<T> A, Ptr4<T> Bs, int narr, int r, int full, long cz, long cy,
                             long total, const DevScalars *S)
{
    if (S && S->hs_done) return;
    const bool alt = S && S->hs_par;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    long d, s;
    ghost_cell(g, r, full, (int)cz, (int)cy, (int)t, d, s);
    for (int a = 0; a < narr; a++) {
        T *q = alt ? Bs.p[a] : A.p[a];
        q[d] = q[s];
    }
}

template <typename T>
static int ghost_fill_t(Box &b, void *const *arrays, void *const *alt, int narrays, int r, int full,
                        const DevScalars *S)
{
    const Geom &g = b.g;
    if (r < 1 || r > g.h) return fail(LCS_ERR_INVALID, "ghost width must be in 1..halo_size");
    if (narrays < 1 || narrays > 4) return fail(LCS_ERR_INVALID, "1..4 arrays per ghost fill");
    long cz, cy, cx;
    ghost_counts(g, r, full, cz, cy, cx);
    long total = cz + cy + cx;
    if (total >= (1L << 31)) return fail(LCS_ERR_INVALID, "grid too large for the ghost fill");
    Ptr4<T> A = {{nullptr, nullptr, nullptr, nullptr}}, Bs = A;
    for (int q = 0; q < narrays; q++) {
        A.p[q] = (T *)arrays[q];
        Bs.p[q] = alt ? (T *)alt[q] : (T *)arrays[q];
    }
    int blocks = (int)((total + 255) / 256);
    ProfScope ps(b, "ghost_fill", 2.0 * total * narrays * sizeof(T));
    hipLaunchKernelGGL(k_ghost_fill<T>, dim3(blocks), dim3(256), 0, b.stream, g, A, Bs, narrays, r, full, cz, cy,
                       total, S);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

int launch_ghost_fill(Box &b, void *const *arrays, int narrays, int width, int full, const int *)
{
    return b.esize == 8 ? ghost_fill_t<double>(b, arrays, nullptr, narrays, width, full, nullptr)
                        : ghost_fill_t<float>(b, arrays, nullptr, narrays, width, full, nullptr);
}

// ---------------------------------------------------------------------------
// RHS (src/flows/rhs.jl:44-252) fused over the three components and tiled
// through LDS (lcs_rhs_tile.h): each flux J1/J3 is formed once instead of being
// stored in six scratch arrays.  FUSED writes out = base + dUdt*dt (update!,
// src/flows/update.jl:37-51), otherwise out = dUdt (interior layout).
// ---------------------------------------------------------------------------
template <typename T>
struct RhsCoef {
    T di[3];
    T di2[3];
    T Rei;
};

template <typename T>
static RhsCoef<T> rhs_coef(const Box &b)
{
    RhsCoef<T> c;
    double d[3] = {spacing(b.g.nx), spacing(b.g.ny), spacing(b.g.nz)};
    for (int a = 0; a < 3; a++) {
        c.di[a] = T(1.0 / d[a]);
        c.di2[a] = c.di[a] * c.di[a];
    }
    c.Rei = T(1.0 / b.Re());
    return c;
}

// Each workgroup marches a TX x TY column up a chunk of zc planes.
template <typename T, int TX, int TY, bool FUSED>
__device__ inline void rhs_tile_block(const RhsArgs<T> &a, int bx, int by, int bz, T rdt, T *lds)
{
    using Tile = RhsTile<T, TX, TY>;
    typename Tile::Thread th;
    Tile::init(a, bx, by, bz, threadIdx.x, th);
    Tile::prologue_fetch(a, th);
    Tile::prologue_store(a, lds, th);
    Tile::prologue_queues(a, th);
    __syncthreads();
    const int z0 = th.z0, z1 = th.z1;
    for (int z = z0; z <= z1; z++) {
        if (z < z1) Tile::fetch(a, z, th);
        Tile::phase_fluxes(lds, z, threadIdx.x);
        __syncthreads();
        Tile::template phase_output<FUSED>(a, lds, z, rdt, th);
        if (z < z1) Tile::store(lds, z, th);
        __syncthreads();
    }
}

template <typename T, int TX, int TY, bool FUSED>
__global__ void __launch_bounds__(TX *TY) k_rhs_tile(RhsArgs<T> a, DtRef dt)
{
    extern __shared__ __align__(16) unsigned char rhs_lds_raw[];
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const T rdt = FUSED ? T(dt.factor * (*dt.src)) : T(0);
    rhs_tile_block<T, TX, TY, FUSED>(a, bx, by, bz, rdt, reinterpret_cast<T *>(rhs_lds_raw));
}

// Tile shapes: FP32 32x8 (39 KB of LDS, four workgroups per CU), FP64 32x4 (53 KB).
template <typename T, bool FUSED, int TX, int TY>
static int rhs_tile_launch(Box &b, const RhsArgs<T> &a0, DtRef dt)
{
    using Tile = RhsTile<T, TX, TY>;
    static_assert(Tile::LDS_ELEMS * sizeof(T) <= 160 * 1024, "tile exceeds the LDS of a CU");
    const Geom &g = a0.g;
    if (g.h < 3) return fail(LCS_ERR_INVALID, "the fourth-order RHS needs halo_size >= 3");
    RhsArgs<T> a = a0;
    const int gx = (g.nx + TX - 1) / TX, gy = (g.ny + TY - 1) / TY;
    // z-chunks: enough workgroups to cover the chip several times over, but
    // chunks of at least 8 planes (each chunk re-reads 6 halo planes)
    const long cols = (long)gx * gy;
    int gz = (int)std::max(1L, std::min((long)(g.nz + 7) / 8, (2048 + cols - 1) / cols));
    a.zc = (g.nz + gz - 1) / gz;
    gz = (g.nz + a.zc - 1) / a.zc;
    const size_t lds = Tile::LDS_ELEMS * sizeof(T);
    int rc = kernel_setup((const void *)k_rhs_tile<T, TX, TY, FUSED>, TX * TY, lds, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL((k_rhs_tile<T, TX, TY, FUSED>), dim3(gx, gy, gz), dim3(TX * TY), lds, b.stream, a, dt);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// Tile shapes (tx, ty) the fused kernels are built for (the LDS of one CU
// bounds them).  Sweep: 64 or 32 wide and 16 or 8 tall, and in FP32 also
// 128 x 8 and 128 x 16 (two cells per thread).  RHS: FP32 up to 64 x 16;
// FP64 64 x 8, 32 x 16 and smaller.
static const int kSweepShapesF[][2] = {{64, 16}, {64, 8}, {32, 16}, {32, 8}, {128, 8}, {128, 16}};
static const int kSweepShapesD[][2] = {{64, 16}, {64, 8}, {32, 16}, {32, 8}};
static const int kRhsShapesF[][2] = {{64, 16}, {64, 8}, {32, 16}, {32, 8}};
static const int kRhsShapesD[][2] = {{64, 8}, {32, 8}, {32, 4}, {32, 16}};
// two-sweep pass (FP32 only; (0, 0) = two single sweeps)
static const int kSweep2Shapes[][2] = {{32, 24}, {64, 12}, {32, 16}, {0, 0}};

int tile_candidates(int esize, int kind, int out[][2], int cap)
{
    const int(*src)[2] = kind == 0 ? (esize == 8 ? kSweepShapesD : kSweepShapesF)
                                   : kind == 2 ? kSweep2Shapes : (esize == 8 ? kRhsShapesD : kRhsShapesF);
    const int n = kind == 0 ? (esize == 8 ? 4 : 6) : kind == 2 ? (esize == 8 ? 0 : 4) : 4;
    for (int i = 0; i < n && i < cap; i++) {
        out[i][0] = src[i][0];
        out[i][1] = src[i][1];
    }
    return n;
}

// Size rule used until a box is tuned: tall tiles only where the planes are
// large enough to fill the chip with one 1024-thread workgroup per CU; sweep
// width by lane efficiency over the whole array row.
TileShapes tiles_default(const Geom &g, int esize)
{
    TileShapes t;
    const bool big = (long)g.nx * g.ny >= 256L * 256L;
    const int X = g.nx + 2 * g.h;
    const double e64 = (double)X / (64.0 * ((X + 63) / 64)), e32 = (double)X / (32.0 * ((X + 31) / 32));
    t.sweep_tx = e32 > e64 + 0.05 ? 32 : 64;
    t.sweep_ty = big ? 16 : 8;
    t.sweep2_tx = t.sweep2_ty = 0;  // the two-sweep pass only where the tuner measured it faster
    if (esize == 4) {
        t.rhs_tx = big ? 64 : 32;
        t.rhs_ty = big ? 16 : 8;
    } else {
        t.rhs_tx = 32;
        t.rhs_ty = 4;
    }
    return t;
}

template <typename T, bool FUSED>
static int rhs_shape(Box &b, const RhsArgs<T> &a, DtRef dt)
{
    const TileShapes t = b.tiles.rhs_tx ? b.tiles : tiles_default(b.g, b.esize);
    if constexpr (sizeof(T) == 4) {
        if (t.rhs_tx == 64 && t.rhs_ty == 16) return rhs_tile_launch<T, FUSED, 64, 16>(b, a, dt);
        if (t.rhs_tx == 64 && t.rhs_ty == 8) return rhs_tile_launch<T, FUSED, 64, 8>(b, a, dt);
        if (t.rhs_tx == 32 && t.rhs_ty == 16) return rhs_tile_launch<T, FUSED, 32, 16>(b, a, dt);
        if (t.rhs_tx == 32 && t.rhs_ty == 8) return rhs_tile_launch<T, FUSED, 32, 8>(b, a, dt);
    } else {
        if (t.rhs_tx == 64 && t.rhs_ty == 8) return rhs_tile_launch<T, FUSED, 64, 8>(b, a, dt);
        if (t.rhs_tx == 32 && t.rhs_ty == 8) return rhs_tile_launch<T, FUSED, 32, 8>(b, a, dt);
        if (t.rhs_tx == 32 && t.rhs_ty == 4) return rhs_tile_launch<T, FUSED, 32, 4>(b, a, dt);
        if (t.rhs_tx == 32 && t.rhs_ty == 16) return rhs_tile_launch<T, FUSED, 32, 16>(b, a, dt);
    }
    return fail(LCS_ERR_INVALID, "unsupported RHS tile shape");
}

template <typename T>
static int rhs_t(Box &b, int stage_in, bool fused, void *const out[3], void *const base[3], DtRef dt)
{
    const Geom &g = b.g;
    void *const *in = stage_in == 1 ? &b.f[LCS_FIELD_U] : &b.f[LCS_FIELD_U2];
    RhsCoef<T> c = rhs_coef<T>(b);
    RhsArgs<T> a;
    a.g = g;
    a.U = (const T *)in[0];
    a.V = (const T *)in[1];
    a.W = (const T *)in[2];
    a.P = (const T *)b.f[LCS_FIELD_P];
    for (int q = 0; q < 3; q++) {
        a.out[q] = (T *)out[q];
        // base == input (stage 1): read it from the tile's registers, not HBM
        a.base[q] = (fused && base[q] != in[q]) ? (const T *)base[q] : nullptr;
        a.di[q] = c.di[q];
        a.di2[q] = c.di2[q];
    }
    a.Rei = c.Rei;
    a.zc = g.nz;
    // algorithmic bytes: read U,V,W,P (+ the three base fields in stage 2), write 3
    ProfScope ps(b, fused ? "rhs_update" : "rhs",
                 (double)g.cells() * sizeof(T) * (fused && stage_in == 2 ? 10 : 7));
    return fused ? rhs_shape<T, true>(b, a, dt) : rhs_shape<T, false>(b, a, dt);
}

int launch_rhs(Box &b, int stage_in, bool fused, void *const out[3], void *const base[3], DtRef dt)
{
    return b.esize == 8 ? rhs_t<double>(b, stage_in, fused, out, base, dt)
                        : rhs_t<float>(b, stage_in, fused, out, base, dt);
}

// update! alone: out = base + dUdt*dt (src/flows/update.jl:37-51)
template <typename T>
__global__ void k_update(Geom g, T *o0, T *o1, T *o2, const T *b0, const T *b1, const T *b2,
                         const T *__restrict__ d0, const T *__restrict__ d1, const T *__restrict__ d2,
                         DtRef dt)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.cells()) return;
    int i = (int)(t % g.nx) + 1;
    long q = t / g.nx;
    int j = (int)(q % g.ny) + 1;
    int k = (int)(q / g.ny) + 1;
    long p = g.at(i, j, k);
    T rdt = T(dt.factor * (*dt.src));
    T du = d0[t] * rdt, dv = d1[t] * rdt, dw = d2[t] * rdt;
    T u = b0[p], v = b1[p], w = b2[p];
    o0[p] = u + du;
    o1[p] = v + dv;
    o2[p] = w + dw;
}

int launch_update(Box &b, void *const out[3], void *const base[3], DtRef dt)
{
    long cells = b.g.cells();
    int blocks = (int)((cells + 255) / 256);
    void *const *d = &b.f[LCS_FIELD_DU];
    ProfScope ps(b, "update", (double)cells * b.esize * 9);
    if (b.esize == 8)
        hipLaunchKernelGGL(k_update<double>, dim3(blocks), dim3(256), 0, b.stream, b.g, (double *)out[0],
                           (double *)out[1], (double *)out[2], (const double *)base[0],
                           (const double *)base[1], (const double *)base[2], (const double *)d[0],
                           (const double *)d[1], (const double *)d[2], dt);
    else
        hipLaunchKernelGGL(k_update<float>, dim3(blocks), dim3(256), 0, b.stream, b.g, (float *)out[0],
                           (float *)out[1], (float *)out[2], (const float *)base[0], (const float *)base[1],
                           (const float *)base[2], (const float *)d[0], (const float *)d[1],
                           (const float *)d[2], dt);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// ---------------------------------------------------------------------------
// HSMAC (src/flows/couples/hsmac.jl).
//
// A sweep over a box R is the reference's colour-0 pass followed by its
// colour-1 pass (foraxes.jl:73-77).  Full boxes run as ONE fused kernel that
// reads the current (U,V,W,P) and writes the result to the other buffer set
// (ping-pong): red corrections depend only on the sweep's input, black ones on
// the input plus the red corrections of their six neighbours, so each face's
// final value -- input + red correction + black correction, in that order --
// is computed by exactly one thread.  The thin face slabs of the overlap
// ordering run as in-place single-colour passes (cells of one colour share no
// face).  Every kernel also reduces div^2 over its cells inside (1:n+1)^3 into
// per-block partials; a one-block check kernel forms the RMS divergence and
// decides convergence on the device (hs_done), after which every later HSMAC
// kernel of the call returns immediately.  The device flag hs_par says which
// buffer set holds the current state; it flips only when a fused sweep ran.
// ---------------------------------------------------------------------------
template <typename T>
struct BufSets {
    T *a[4];  // (U, V, W, P) as the call started
    T *b[4];  // spare partners
};

// One colour pass over one or two (thin) boxes, in place on the current
// buffer set.  Two boxes share a launch only when no cell of one touches a
// face of the other (the opposite face slabs of one axis), so the pair gives
// the same result as the two passes in either order.

template <typename T, int COLORING>
__global__ void __launch_bounds__(256) k_hsmac_pass(Geom g, BufSets<T> bs, Box3 r0, Box3 r1, int n0, int n1,
                                                   int color, HsCoef<T> c, DtRef dt, const DevScalars *S,
                                                   double *part)
{
    __shared__ double lds[16];
    if (S->hs_done) return;
    const bool alt = S->hs_par != 0;
    T *U = alt ? bs.b[0] : bs.a[0], *V = alt ? bs.b[1] : bs.a[1];
    T *W = alt ? bs.b[2] : bs.a[2], *P = alt ? bs.b[3] : bs.a[3];
    const T dti = T(1.0 / (dt.factor * (*dt.src)));
    const long sy = g.sx, sz = g.sx * g.sy;
    double acc = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n0 + n1; t += gridDim.x * blockDim.x) {
        const bool second = t >= n0;
        const Box3 r = second ? r1 : r0;
        const int u = second ? t - n0 : t;
        const int a = r.lo[0];
        const int hw = (r.hi[0] - r.lo[0] + 2) / 2;  // cld(n, 2)
        const int ny = r.hi[1] - r.lo[1] + 1;
        const int ih = u % hw, q = u / hw;
        const int j = r.lo[1] + q % ny, k = r.lo[2] + q / ny;
        // RedBlackFast index map: lib/parallel/src/foraxes.jl:100-105
        const int i = a + 2 * ih + ((j + k - a + color) & 1);
        if (i > r.hi[0]) continue;
        if (!visited<COLORING>(r, i, j, k)) continue;
        const long p = g.at(i, j, k);
        T dudx = c.dxi * (-U[p - 1] + U[p]);
        T dvdy = c.dyi * (-V[p - sy] + V[p]);
        T dwdz = c.dzi * (-W[p - sz] + W[p]);
        T div = (dudx + dvdy) + dwdz;
        T dp = (-c.omega * div) * c.dl;
        P[p] += dp * dti;
        U[p - 1] -= dp * c.dxi;
        U[p] += dp * c.dxi;
        V[p - sy] -= dp * c.dyi;
        V[p] += dp * c.dyi;
        W[p - sz] -= dp * c.dzi;
        W[p] += dp * c.dzi;
        if (in_d3(g, i, j, k)) acc += (double)(div * div);
    }
    double s = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Both colours over the two x face slabs (lcs_hsmac_xface.h), launch 1:
// patches of slab rows, final values into the compact buffer.  The grid
// strides over the patches; each workgroup leaves its div^2 sum in part.
template <typename T, int COLORING>
__global__ void __launch_bounds__(XfPatch<T>::N *XfPatch<T>::N)
    k_hsmac_xface(Geom g, BufSets<T> bs, Box3 r0, Box3 r1, HsCoef<T> c, DtRef dt, const DevScalars *S, T *buf,
                  double *part)
{
    using X = XFaceOf<T, COLORING>;
    extern __shared__ __align__(16) unsigned char xf_lds_raw[];
    T *lds = reinterpret_cast<T *>(xf_lds_raw);
    __shared__ double red_lds[16];
    if (S->hs_done) return;
    const bool alt = S->hs_par != 0;
    XfArgs<T> a;
    a.g = g;
    a.U = alt ? bs.b[0] : bs.a[0];
    a.V = alt ? bs.b[1] : bs.a[1];
    a.W = alt ? bs.b[2] : bs.a[2];
    a.P = alt ? bs.b[3] : bs.a[3];
    a.R[0] = r0;
    a.R[1] = r1;
    a.c = c;
    a.dti = T(1.0 / (dt.factor * (*dt.src)));
    a.buf = buf;
    double acc = 0.0;
    const int nblk = X::blocks(r0);
    for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        typename X::Patch p;
        X::patch(a, blk, p);
        X::load(a, p, lds, threadIdx.x);
        __syncthreads();
        X::red(a, p, lds, threadIdx.x, acc);
        __syncthreads();
        X::black(a, p, lds, threadIdx.x, acc);
        __syncthreads();
        X::finals(a, p, lds, threadIdx.x);
        __syncthreads();
    }
    const double s = block_sum(acc, red_lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// launch 2: the compact buffer into U, V, W, P, one thread per slab row
template <typename T>
__global__ void __launch_bounds__(256) k_hsmac_xface_commit(Geom g, BufSets<T> bs, Box3 r0, Box3 r1,
                                                           const DevScalars *S, const T *buf)
{
    if (S->hs_done) return;
    const bool alt = S->hs_par != 0;
    T *U = alt ? bs.b[0] : bs.a[0], *V = alt ? bs.b[1] : bs.a[1];
    T *W = alt ? bs.b[2] : bs.a[2], *P = alt ? bs.b[3] : bs.a[3];
    const long rows = SlabAxes<0>::rows(r0);
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (long)gridDim.x * blockDim.x)
        XFaceOf<T, LCS_COLORING_RED_BLACK>::commit_row(g, r0, r1, U, V, W, P, buf, r);
}

// Fused red+black sweep over R, current -> other buffer set (the phase
// functions and their schedule are in lcs_hsmac_tile.h).  Returns the
// workgroup's div^2 sum (valid in thread 0).
template <typename T, int COLORING, int TY, int TX, bool FORCE = false>
__device__ inline double hsmac_sweep_block(const HsArgs<T> &a, int bx, int by, int bz, T *lds, double *red_lds,
                                           double *power = nullptr)
{
    using Tile = HsTile<T, COLORING, TY, TX, FORCE>;
    typename Tile::Thread th;
    Tile::init(a, bx, by, bz, threadIdx.x, th);
    Tile::prologue(a, lds, th);
    __syncthreads();
    const int q1 = Tile::last_step(th);
    if constexpr (Tile::NSLOT == 3) {
        for (int q = Tile::first_step(th); q <= q1; q += 3) {  // unrolled by the three register slots
            Tile::step_a(a, lds, q, th, 2);
            __syncthreads();
            Tile::step_b(a, lds, q, th, 0);
            __syncthreads();
            if (q + 1 > q1) break;
            Tile::step_a(a, lds, q + 1, th, 0);
            __syncthreads();
            Tile::step_b(a, lds, q + 1, th, 1);
            __syncthreads();
            if (q + 2 > q1) break;
            Tile::step_a(a, lds, q + 2, th, 1);
            __syncthreads();
            Tile::step_b(a, lds, q + 2, th, 2);
            __syncthreads();
        }
    } else {
        for (int q = Tile::first_step(th); q <= q1; q += 2) {  // unrolled by the two register slots
            Tile::step_a(a, lds, q, th, 1);
            __syncthreads();
            Tile::step_b(a, lds, q, th, 0);
            __syncthreads();
            if (q + 1 > q1) break;
            Tile::step_a(a, lds, q + 1, th, 0);
            __syncthreads();
            Tile::step_b(a, lds, q + 1, th, 1);
            __syncthreads();
        }
    }
    if (FORCE) *power = block_sum(th.pacc, red_lds);
    return block_sum(th.acc, red_lds);
}

// the two buffer sets of a sweep: the parity says which one is current
template <typename T>
__device__ inline void hs_bind(HsArgs<T> &a, const BufSets<T> &bs, bool par)
{
    a.Us = par ? bs.b[0] : bs.a[0];
    a.Vs = par ? bs.b[1] : bs.a[1];
    a.Ws = par ? bs.b[2] : bs.a[2];
    a.Ps = par ? bs.b[3] : bs.a[3];
    a.Ud = par ? bs.a[0] : bs.b[0];
    a.Vd = par ? bs.a[1] : bs.b[1];
    a.Wd = par ? bs.a[2] : bs.b[2];
    a.Pd = par ? bs.a[3] : bs.b[3];
}

// FORCE: the linear forcing of S->lf is applied to U, V, W as they are
// loaded, and each workgroup's forcing power goes to ppart.
template <typename T, int COLORING, int TY, int TX, bool FORCE = false>
__global__ void __launch_bounds__((HsTile<T, COLORING, TY, TX>::NT))
    k_hsmac_sweep(Geom g, BufSets<T> bs, Box3 R, HsCoef<T> c, DtRef dt, DevScalars *S, double *part, int zc,
                  double *ppart, int wrap)
{
    extern __shared__ __align__(16) unsigned char hs_lds_raw[];
    __shared__ double red_lds[16];
    if (S->hs_done) return;
    HsArgs<T> a;
    a.g = g;
    hs_bind(a, bs, S->hs_par != 0);
    a.R = R;
    a.c = c;
    a.dti = T(1.0 / (dt.factor * (*dt.src)));
    a.zc = zc;
    a.wrap = wrap;
    if (FORCE) {  // the coefficients k_lf_apply uses (lf_apply_block)
        a.xf = 1;
        a.xA = T(S->lf[4]);
        for (int q = 0; q < 3; q++) a.xm[q] = T(S->lf[1 + q]);
        a.xdt = T(dt.factor * (*dt.src));
    }
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    double pw = 0.0;
    const double s = hsmac_sweep_block<T, COLORING, TY, TX, FORCE>(a, bx, by, bz, reinterpret_cast<T *>(hs_lds_raw),
                                                                   red_lds, &pw);
    const long blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (threadIdx.x == 0) {
        part[blk] = s;
        if (FORCE) ppart[blk] = pw;
    }
}

// Two fused sweeps, RA then RB, current -> other buffer set (phase functions
// and schedule in lcs_hsmac_tile2.h).  The div^2 sums of the two sweeps go to
// their own partial slots.
template <typename T, int COLORING, int TY, int TX>
__device__ inline void hsmac_sweep2_block(const Hs2Args<T> &a, int bx, int by, int bz, T *lds, double *red_lds,
                                          double &sa, double &sb)
{
    using Tile = HsTile2<T, COLORING, TY, TX>;
    typename Tile::Thread th;
    Tile::init(a, bx, by, bz, threadIdx.x, th);
    Tile::prologue(a, lds, th);
    __syncthreads();
    const int q0 = Tile::first_step(th), q1 = Tile::last_step(th);
    for (int q = q0; q <= q1; q += 2) {
        Tile::step_1(a, lds, q, th, 0);
        __syncthreads();
        Tile::step_2(a, lds, q, th, 0);
        __syncthreads();
        if (q + 1 > q1) break;
        Tile::step_1(a, lds, q + 1, th, 1);
        __syncthreads();
        Tile::step_2(a, lds, q + 1, th, 1);
        __syncthreads();
    }
    sa = block_sum(th.accA, red_lds);
    sb = block_sum(th.accB, red_lds);
}

template <typename T>
__device__ inline void hs2_bind(Hs2Args<T> &a, const BufSets<T> &bs, bool par)
{
    a.Us = par ? bs.b[0] : bs.a[0];
    a.Vs = par ? bs.b[1] : bs.a[1];
    a.Ws = par ? bs.b[2] : bs.a[2];
    a.Ps = par ? bs.b[3] : bs.a[3];
    a.Ud = par ? bs.a[0] : bs.b[0];
    a.Vd = par ? bs.a[1] : bs.b[1];
    a.Wd = par ? bs.a[2] : bs.b[2];
    a.Pd = par ? bs.a[3] : bs.b[3];
}

template <typename T, int COLORING, int TY, int TX>
__global__ void __launch_bounds__(TX *TY) k_hsmac_sweep2(Geom g, BufSets<T> bs, Box3 RA, Box3 RB, HsCoef<T> c,
                                                        DtRef dt, DevScalars *S, double *partA, double *partB, int zc)
{
    extern __shared__ __align__(16) unsigned char hs_lds_raw[];
    __shared__ double red_lds[16];
    if (S->hs_done) return;
    Hs2Args<T> a;
    a.g = g;
    hs2_bind(a, bs, S->hs_par != 0);
    a.RA = RA;
    a.RB = RB;
    a.c = c;
    a.dti = T(1.0 / (dt.factor * (*dt.src)));
    a.zc = zc;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    double sa, sb;
    hsmac_sweep2_block<T, COLORING, TY, TX>(a, bx, by, bz, reinterpret_cast<T *>(hs_lds_raw), red_lds, sa, sb);
    const int L = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (threadIdx.x == 0) {
        partA[L] = sa;
        partB[L] = sb;
    }
}

struct CheckArgs {
    int npass;
    int cnt[kMaxSweepPasses];
    double len3, dxmin, epsp;
    int itrmin;
};

__global__ void k_hsmac_check(const double *part, CheckArgs a, DevScalars *S)
{
    __shared__ double lds[16];
    if (S->hs_done) return;  // uniform across the block
    double acc = 0.0;
    for (int q = 0; q < a.npass; q++) {
        const double *pp = part + (long)q * kMaxPartBlocks;
        for (int m = threadIdx.x; m < a.cnt[q]; m += blockDim.x) acc += pp[m];
    }
    double s = block_sum(acc, lds);
    if (threadIdx.x == 0) {
        double div = sqrt(s / a.len3);
        int itr = S->hs_itr;
        S->hs_div = div;
        if (div * a.dxmin < a.epsp && a.itrmin <= itr)
            S->hs_done = 1;
        else
            S->hs_itr = itr + 1;
    }
}

__global__ void k_hsmac_begin(DevScalars *S)
{
    S->hs_done = 0;
    S->hs_itr = 0;
    S->hs_div = 0.0;
    S->hs_par = 0;
}

__global__ void k_hsmac_flip(DevScalars *S)
{
    if (!S->hs_done) S->hs_par ^= 1;
}

__global__ void k_hsmac_end(DevScalars *S, int slot)
{
    if (slot >= 0 && slot < 2) {
        S->itr[slot] = S->hs_itr;
        S->div[slot] = S->hs_div;
    }
}

__global__ void k_mean_to(const double *part, int nblk, double len, double *dst);

template <typename T>
struct HsmacRun {
    Box &b;
    BufSets<T> bs;
    HsCoef<T> c;
    DtRef dt;
    int coloring;
    CheckArgs ca;
    int npass = 0;
    std::vector<double> sweep_bytes;  // algorithmic bytes of each fused sweep launched in this call
    std::vector<int> sweep_first;     // for each launch of sweep_bytes: index of its first sweep
    std::vector<const char *> sweep_name;
    int sweeps = 0;                   // sweeps enqueued so far in this call
    bool force_next = false;          // apply the deferred linear forcing in the next fused sweep
    bool wrap_next = false;           // the next fused sweep loads ghost U, V, W from their images (in place of a fill)

    // one colour over boxes r0 and r1 (see k_hsmac_pass for when two boxes may share a launch)
    int pass(const Box3 &r0, const Box3 &r1, int color)
    {
        const int n0 = pass_count(r0), n1 = pass_count(r1);
        if (n0 + n1 == 0) {
            ca.cnt[npass++] = 0;
            return LCS_OK;
        }
        int blocks = grid_for((long)n0 + n1, 256, 4096);
        double *part = b.part + (long)npass * kMaxPartBlocks;
        // algorithmic: half of the 8 element passes (U,V,W,P read+write) per cell-sweep
        auto cells = [](const Box3 &r) {
            return r.hi[0] < r.lo[0] || r.hi[1] < r.lo[1] || r.hi[2] < r.lo[2]
                       ? 0L
                       : (long)(r.hi[0] - r.lo[0] + 1) * (r.hi[1] - r.lo[1] + 1) * (r.hi[2] - r.lo[2] + 1);
        };
        ProfScope ps(b, "hsmac_pass", 4.0 * sizeof(T) * (cells(r0) + cells(r1)));
        if (coloring == LCS_COLORING_RED_BLACK)
            hipLaunchKernelGGL((k_hsmac_pass<T, LCS_COLORING_RED_BLACK>), dim3(blocks), dim3(256), 0, b.stream,
                               b.g, bs, r0, r1, n0, n1, color, c, dt, b.S, part);
        else
            hipLaunchKernelGGL((k_hsmac_pass<T, LCS_COLORING_RED_BLACK_FAST>), dim3(blocks), dim3(256), 0,
                               b.stream, b.g, bs, r0, r1, n0, n1, color, c, dt, b.S, part);
        LCS_HIP(hipGetLastError());
        ca.cnt[npass++] = blocks;
        return LCS_OK;
    }
    // A box alone runs its x slabs as the pair from kXfMinRows rows per slab:
    // below that a slab has a few patches, whose rows then all go through the
    // few CUs that run them, and the two colour passes spread over many CUs
    // are faster (20-25 us a group at 32^3-128^3; 256^3 and up gain).  In a
    // batch the patches of all boxes fill the device, so every box uses it.
    static constexpr long kXfMinRows = 1L << 15;
    // both colours over the x- and x+ slabs, through the compact buffer
    int xface(const Box3 &r0, const Box3 &r1)
    {
        using X = XFaceOf<T, LCS_COLORING_RED_BLACK>;
        const Geom &g = b.g;
        const size_t need = (size_t)kXfVals * SlabAxes<0>::rows(r0) * sizeof(T);
        if (b.xf_bytes < need) {
            if (b.xf) {
                LCS_HIP(hipStreamSynchronize(b.stream));
                LCS_HIP(hipFree(b.xf));
                b.bytes -= b.xf_bytes;
                b.xf = nullptr;
                b.xf_bytes = 0;
            }
            LCS_HIP(hipMalloc(&b.xf, need));
            b.xf_bytes = need;
            b.bytes += need;
        }
        const int blocks = std::min(X::blocks(r0), kMaxPartBlocks);
        double *part = b.part + (long)npass * kMaxPartBlocks;
        const double cells = (double)(r0.hi[0] - r0.lo[0] + r1.hi[0] - r1.lo[0] + 2) * (g.ny - 1) * (g.nz - 1);
        const size_t lds = X::LDS_ELEMS * sizeof(T);
        static_assert(X::LDS_ELEMS * sizeof(T) + 128 <= 160 * 1024, "x-face patch exceeds the LDS of a CU");
        {
            ProfScope ps(b, "hsmac_xface", 4.0 * sizeof(T) * cells);
            const bool rb = coloring == LCS_COLORING_RED_BLACK;
            const void *k = rb ? (const void *)k_hsmac_xface<T, LCS_COLORING_RED_BLACK>
                               : (const void *)k_hsmac_xface<T, LCS_COLORING_RED_BLACK_FAST>;
            int rc = kernel_setup(k, X::NT, lds, nullptr);
            if (rc) return rc;
            if (rb)
                hipLaunchKernelGGL((k_hsmac_xface<T, LCS_COLORING_RED_BLACK>), dim3(blocks), dim3(X::NT), lds,
                                   b.stream, g, bs, r0, r1, c, dt, b.S, (T *)b.xf, part);
            else
                hipLaunchKernelGGL((k_hsmac_xface<T, LCS_COLORING_RED_BLACK_FAST>), dim3(blocks), dim3(X::NT), lds,
                                   b.stream, g, bs, r0, r1, c, dt, b.S, (T *)b.xf, part);
            LCS_HIP(hipGetLastError());
        }
        ca.cnt[npass++] = blocks;
        ProfScope ps(b, "hsmac_xface_commit", 4.0 * sizeof(T) * cells);
        hipLaunchKernelGGL(k_hsmac_xface_commit<T>, dim3(grid_for(SlabAxes<0>::rows(r0), 256, 8192)), dim3(256), 0, b.stream, g,
                           bs, r0, r1, b.S, (const T *)b.xf);
        LCS_HIP(hipGetLastError());
        return LCS_OK;
    }
    // one fused red+black sweep over a full box, then the parity flip
    template <int COLORING, int TY, int TX>
    int fused_launch(const Box3 &r)
    {
        using Tile = HsTile<T, COLORING, TY, TX>;
        static_assert(Tile::LDS_ELEMS * sizeof(T) <= 160 * 1024, "tile exceeds the LDS of a CU");
        const Geom &g = b.g;
        const int gx = (g.nx + 2 * g.h + Tile::TX - 1) / Tile::TX, gy = (g.ny + 2 * g.h + TY - 1) / TY;
        const int zplanes = g.nz + 2 * g.h;
        const long cols = (long)gx * gy;
        const size_t lds = Tile::LDS_ELEMS * sizeof(T);
        int per_cu = 1;  // resident workgroups per CU
        int rc = kernel_setup((const void *)k_hsmac_sweep<T, COLORING, TY, TX>, Tile::NT, lds, &per_cu);
        if (rc) return rc;
        // z-chunks: a chunk re-reads five halo planes and the last round of
        // workgroups may be partial; pick the count with the fewest
        // plane-steps on the busiest CU slot, among chunks of at most
        // kSweepMaxChunk planes where the grid allows.  Workgroups finish at
        // different times, so a launch ends about one chunk after most of
        // its work: at 1024^3 9 chunks of 115 planes take 9.46 ms a sweep,
        // against 9.71 ms for the 3 chunks of 344 planes the count alone
        // picks; at 256^3 and 512^3 the cap changes nothing.
        constexpr int kSweepMaxChunk = 128;
        const long slots = (long)b.cus * per_cu;
        int n0 = (zplanes + kSweepMaxChunk - 1) / kSweepMaxChunk;
        if (cols * n0 > kMaxPartBlocks) n0 = 1;
        int nzc = n0;
        double best = 1e300;
        for (int n = n0; n <= 64 && n <= zplanes; n++) {
            if (cols * n > kMaxPartBlocks) break;
            const long zc_n = (zplanes + n - 1) / n;
            const double cost = (double)((cols * n + slots - 1) / slots) * (double)(zc_n + 5);
            if (cost < best * 0.999) {
                best = cost;
                nzc = n;
            }
        }
        if (cols * nzc > kMaxPartBlocks) return fail(LCS_ERR_INVALID, "grid too large for the fused HSMAC sweep");
        const int zc = (zplanes + nzc - 1) / nzc;
        nzc = (zplanes + zc - 1) / zc;
        double *part = b.part + (long)npass * kMaxPartBlocks;
        long cells = (long)(r.hi[0] - r.lo[0] + 1) * (r.hi[1] - r.lo[1] + 1) * (r.hi[2] - r.lo[2] + 1);
        ProfScope ps(b, "hsmac_sweep", 8.0 * sizeof(T) * cells);
        sweep_bytes.push_back(8.0 * sizeof(T) * cells);
        sweep_first.push_back(sweeps++);
        sweep_name.push_back("hsmac_sweep");
        double *ppart = force_next ? b.part + kPowerSlot : nullptr;
        const int wrap = wrap_next ? 1 : 0;
        wrap_next = false;
        if (force_next) {
            rc = kernel_setup((const void *)k_hsmac_sweep<T, COLORING, TY, TX, true>, Tile::NT, lds, nullptr);
            if (rc) return rc;
            hipLaunchKernelGGL((k_hsmac_sweep<T, COLORING, TY, TX, true>), dim3(gx, gy, nzc), dim3(Tile::NT), lds,
                               b.stream, g, bs, r, c, dt, b.S, part, zc, ppart, wrap);
        } else {
            hipLaunchKernelGGL((k_hsmac_sweep<T, COLORING, TY, TX>), dim3(gx, gy, nzc), dim3(Tile::NT), lds,
                               b.stream, g, bs, r, c, dt, b.S, part, zc, ppart, wrap);
        }
        LCS_HIP(hipGetLastError());
        if (force_next) {  // state.p: mean cell power (lf.jl:103)
            force_next = false;
            hipLaunchKernelGGL(k_mean_to, dim3(1), dim3(256), 0, b.stream, ppart, (int)(cols * nzc),
                               (double)g.cells(), &b.S->p);
            LCS_HIP(hipGetLastError());
        }
        ca.cnt[npass++] = (int)(cols * nzc);
        hipLaunchKernelGGL(k_hsmac_flip, dim3(1), dim3(1), 0, b.stream, b.S);
        LCS_HIP(hipGetLastError());
        return LCS_OK;
    }
    template <int COLORING>
    int fused_shape(const Box3 &r, int tx, int ty)
    {
        if (tx == 64 && ty == 16) return fused_launch<COLORING, 16, 64>(r);
        if (tx == 64 && ty == 8) return fused_launch<COLORING, 8, 64>(r);
        if (tx == 32 && ty == 16) return fused_launch<COLORING, 16, 32>(r);
        if (tx == 32 && ty == 8) return fused_launch<COLORING, 8, 32>(r);
        if constexpr (sizeof(T) == 4) {
            if (tx == 128 && ty == 8) return fused_launch<COLORING, 8, 128>(r);
            if (tx == 128 && ty == 16) return fused_launch<COLORING, 16, 128>(r);
        }
        return fail(LCS_ERR_INVALID, "unsupported sweep tile shape");
    }
    // one fused sweep with the box's tile shape
    int fused(const Box3 &r)
    {
        const TileShapes t = b.tiles.sweep_tx ? b.tiles : tiles_default(b.g, b.esize);
        return coloring == LCS_COLORING_RED_BLACK ? fused_shape<LCS_COLORING_RED_BLACK>(r, t.sweep_tx, t.sweep_ty)
                                                  : fused_shape<LCS_COLORING_RED_BLACK_FAST>(r, t.sweep_tx, t.sweep_ty);
    }
    // Topologies.faces order z-, z+, y-, y+, x-, x+ (lib/topologies/src/halo.jl:349-366)
    int sweep_core_faces(const Box3 &d)
    {
        const Geom &g = b.g;
        Box3 core = {{2, 2, 2}, {g.nx, g.ny, g.nz}};
        int rc = fused(core);
        if (rc) return rc;
        Box3 f[6] = {
            {{d.lo[0], d.lo[1], d.lo[2]}, {d.hi[0], d.hi[1], 1}},
            {{d.lo[0], d.lo[1], g.nz + 1}, {d.hi[0], d.hi[1], d.hi[2]}},
            {{d.lo[0], d.lo[1], 2}, {d.hi[0], 1, g.nz}},
            {{d.lo[0], g.ny + 1, 2}, {d.hi[0], d.hi[1], g.nz}},
            {{d.lo[0], 2, 2}, {1, g.ny, g.nz}},
            {{g.nx + 1, 2, 2}, {d.hi[0], g.ny, g.nz}},
        };
        // Opposite slabs of one axis share no face for n >= 2, so the
        // reference's order (z- red, z- black, z+ red, z+ black, y- ...) gives
        // the same result as (z- and z+ red), (z- and z+ black), (y- ...).
        const bool pair = g.nx >= 2 && g.ny >= 2 && g.nz >= 2;
        for (int q = 0; q < 6; q += 2) {
            if (q == 4 && xface_fits(g, f + 4) && (long)(g.ny - 1) * (g.nz - 1) >= kXfMinRows) {
                if ((rc = xface(f[4], f[5]))) return rc;
            } else if (pair) {
                if ((rc = pass(f[q], f[q + 1], 0)) || (rc = pass(f[q], f[q + 1], 1))) return rc;
            } else {
                const Box3 none = {{1, 1, 1}, {0, 0, 0}};
                for (int s = 0; s < 2; s++)
                    if ((rc = pass(f[q + s], none, 0)) || (rc = pass(f[q + s], none, 1))) return rc;
            }
        }
        return LCS_OK;
    }
    // two fused sweeps RA then RB; partial slots npass (RA) and npass+1 (RB)
    template <int COLORING, int TY, int TX>
    int fused2_launch(const Box3 &ra, const Box3 &rb)
    {
        using Tile = HsTile2<T, COLORING, TY, TX>;
        static_assert(Tile::LDS_ELEMS * sizeof(T) + 256 <= 160 * 1024, "tile exceeds the LDS of a CU");
        const Geom &g = b.g;
        const int gx = (g.nx + 2 * g.h + TX - 1) / TX, gy = (g.ny + 2 * g.h + TY - 1) / TY;
        const int zplanes = g.nz + 2 * g.h;
        const long cols = (long)gx * gy;
        const size_t lds = Tile::LDS_ELEMS * sizeof(T);
        int per_cu = 1;
        int rc = kernel_setup((const void *)k_hsmac_sweep2<T, COLORING, TY, TX>, Tile::NT, lds, &per_cu);
        if (rc) return rc;
        // z-chunks as for one sweep; a chunk here re-reads about nine halo planes
        const long slots = (long)b.cus * per_cu;
        int nzc = 1;
        double best = 1e300;
        for (int n = 1; n <= 64 && n <= zplanes; n++) {
            if (cols * n > kMaxPartBlocks) break;
            const long zc_n = (zplanes + n - 1) / n;
            const double cost = (double)((cols * n + slots - 1) / slots) * (double)(zc_n + 9);
            if (cost < best * 0.999) {
                best = cost;
                nzc = n;
            }
        }
        if (cols * nzc > kMaxPartBlocks || npass + 2 > kMaxSweepPasses)
            return fail(LCS_ERR_INVALID, "grid too large for the two-sweep HSMAC kernel");
        const int zc = (zplanes + nzc - 1) / nzc;
        nzc = (zplanes + zc - 1) / zc;
        double *pa = b.part + (long)npass * kMaxPartBlocks, *pb = pa + kMaxPartBlocks;
        auto cells = [](const Box3 &r) {
            return (double)(r.hi[0] - r.lo[0] + 1) * (r.hi[1] - r.lo[1] + 1) * (r.hi[2] - r.lo[2] + 1);
        };
        const double bytes = 8.0 * sizeof(T) * (cells(ra) + cells(rb));
        ProfScope ps(b, "hsmac_sweep2", bytes);
        sweep_bytes.push_back(bytes);
        sweep_first.push_back(sweeps);
        sweep_name.push_back("hsmac_sweep2");
        sweeps += 2;
        hipLaunchKernelGGL((k_hsmac_sweep2<T, COLORING, TY, TX>), dim3(gx, gy, nzc), dim3(Tile::NT), lds, b.stream,
                           g, bs, ra, rb, c, dt, b.S, pa, pb, zc);
        LCS_HIP(hipGetLastError());
        ca.cnt[npass++] = (int)(cols * nzc);
        ca.cnt[npass++] = (int)(cols * nzc);
        hipLaunchKernelGGL(k_hsmac_flip, dim3(1), dim3(1), 0, b.stream, b.S);
        LCS_HIP(hipGetLastError());
        return LCS_OK;
    }
    // the d2 and d3 sweeps of a group as one pass, if this box has a two-sweep shape
    bool has_fused2() const
    {
        const TileShapes t = b.tiles.sweep_tx ? b.tiles : tiles_default(b.g, b.esize);
        return sizeof(T) == 4 && t.sweep2_tx > 0;
    }
    template <int COLORING>
    int fused2_shape(const Box3 &ra, const Box3 &rb, int tx, int ty)
    {
        if constexpr (sizeof(T) == 4) {
            if (tx == 32 && ty == 24) return fused2_launch<COLORING, 24, 32>(ra, rb);
            if (tx == 64 && ty == 12) return fused2_launch<COLORING, 12, 64>(ra, rb);
            if (tx == 32 && ty == 16) return fused2_launch<COLORING, 16, 32>(ra, rb);
        }
        return fail(LCS_ERR_INVALID, "unsupported two-sweep tile shape");
    }
    int fused2(const Box3 &ra, const Box3 &rb)
    {
        const TileShapes t = b.tiles.sweep_tx ? b.tiles : tiles_default(b.g, b.esize);
        return coloring == LCS_COLORING_RED_BLACK
                   ? fused2_shape<LCS_COLORING_RED_BLACK>(ra, rb, t.sweep2_tx, t.sweep2_ty)
                   : fused2_shape<LCS_COLORING_RED_BLACK_FAST>(ra, rb, t.sweep2_tx, t.sweep2_ty);
    }
    // one convergence check over partial slots [first, first + count)
    int check_slots(int first, int count)
    {
        CheckArgs c2 = ca;
        c2.npass = count;
        for (int i = 0; i < count; i++) c2.cnt[i] = ca.cnt[first + i];
        ProfScope ps(b, "hsmac_check", 0.0);
        hipLaunchKernelGGL(k_hsmac_check, dim3(1), dim3(256), 0, b.stream, b.part + (long)first * kMaxPartBlocks, c2,
                           b.S);
        LCS_HIP(hipGetLastError());
        return LCS_OK;
    }
    int check()
    {
        ca.npass = npass;
        ProfScope ps(b, "hsmac_check", 0.0);
        hipLaunchKernelGGL(k_hsmac_check, dim3(1), dim3(256), 0, b.stream, b.part, ca, b.S);
        LCS_HIP(hipGetLastError());
        npass = 0;
        return LCS_OK;
    }
    int fill(int width, bool withP)
    {
        if (b.esize == 8)
            return ghost_fill_t<double>(b, (void *const *)bs.a, (void *const *)bs.b, withP ? 4 : 3, width, 1, b.S);
        return ghost_fill_t<float>(b, (void *const *)bs.a, (void *const *)bs.b, withP ? 4 : 3, width, 1, b.S);
    }
    int poll(bool &done)
    {
        LCS_HIP(hipMemcpyAsync(b.S_host, b.S, sizeof(DevScalars), hipMemcpyDeviceToHost, b.stream));
        LCS_HIP(hipStreamSynchronize(b.stream));
        done = b.S_host->hs_done != 0;
        return LCS_OK;
    }
};

bool hsmac_group_wraps(const Geom &g, const lcs_hsmac_params &p)
{
    return p.time_blocking && g.nx >= g.h && g.ny >= g.h && g.nz >= g.h;
}

template <typename T>
static int hsmac_t(Box &b, void **Us, DtRef dt, const lcs_hsmac_params &prm, int slot, bool lf_deferred)
{
    const Geom &g = b.g;
    if (g.h < 3 && prm.time_blocking) return fail(LCS_ERR_INVALID, "time-blocked HSMAC needs halo_size >= 3");
    for (int q = 0; q < 4; q++)
        if (!b.spare[q]) {
            LCS_HIP(hipMalloc(&b.spare[q], (size_t)g.len * sizeof(T)));
            b.bytes += (size_t)g.len * sizeof(T);
        }
    double dx = spacing(g.nx), dy = spacing(g.ny), dz = spacing(g.nz);
    T dxi = T(1.0 / dx), dyi = T(1.0 / dy), dzi = T(1.0 / dz);
    BufSets<T> bs;
    void **Pp = &b.f[LCS_FIELD_P];
    for (int q = 0; q < 3; q++) bs.a[q] = (T *)Us[q];
    bs.a[3] = (T *)*Pp;
    for (int q = 0; q < 4; q++) bs.b[q] = (T *)b.spare[q];
    HsmacRun<T> run{b, bs, {}, dt, prm.coloring, {}};
    run.c.dxi = dxi;
    run.c.dyi = dyi;
    run.c.dzi = dzi;
    run.c.dl = T(1) / (T(2) * ((dxi * dxi + dyi * dyi) + dzi * dzi));
    run.c.omega = T(prm.omega);
    run.ca.len3 = (double)(g.nx + 1) * (double)(g.ny + 1) * (double)(g.nz + 1);
    run.ca.dxmin = std::min(std::min(dx, dy), dz);
    run.ca.epsp = prm.epsp;
    run.ca.itrmin = prm.itrmin;
    run.force_next = lf_deferred;  // the first sweep of the call applies the linear forcing

    hipLaunchKernelGGL(k_hsmac_begin, dim3(1), dim3(1), 0, b.stream, b.S);
    LCS_HIP(hipGetLastError());
    int rc;
    bool done = false, polled_last = false;
    const Box3 d3 = {{1, 1, 1}, {g.nx + 1, g.ny + 1, g.nz + 1}};
    if (!prm.time_blocking) {
        // _couple_naive!: hsmac.jl:77-129
        for (int s = 0; s < prm.itrmax && !done; s++) {
            if ((rc = run.fill(1, !prm.overlap))) return rc;
            rc = prm.overlap ? run.sweep_core_faces(d3) : run.fused(d3);
            if (rc) return rc;
            if ((rc = run.check())) return rc;
            polled_last = false;
            if (s >= prm.itrmin) {
                if ((rc = run.poll(done))) return rc;
                polled_last = true;
            }
        }
    } else {
        // _couple_time_blocking!: hsmac.jl:131-201
        const Box3 d1 = {{-1, -1, -1}, {g.nx + 3, g.ny + 3, g.nz + 3}};
        const Box3 d2 = {{0, 0, 0}, {g.nx + 2, g.ny + 2, g.nz + 2}};
        // The group's FullHalo fill of U, V, W folds into the loads of its
        // first sweep (HsArgs::wrap) where every ghost has one image.
        const bool wrap = hsmac_group_wraps(g, prm);
        for (int grp = 0; 3 * grp < prm.itrmax && !done; grp++) {
            if (wrap)
                run.wrap_next = true;
            else if ((rc = run.fill(g.h, false)))
                return rc;
            rc = prm.overlap ? run.sweep_core_faces(d1) : run.fused(d1);
            if (rc) return rc;
            if ((rc = run.check())) return rc;
            if (3 * grp + 2 < prm.itrmin && run.has_fused2()) {
                // neither the d2 nor the d3 sweep can end the iteration: one pass
                if ((rc = run.fused2(d2, d3)) || (rc = run.check_slots(0, 1)) || (rc = run.check_slots(1, 1)))
                    return rc;
                run.npass = 0;
            } else {
                if ((rc = run.fused(d2))) return rc;
                if ((rc = run.check())) return rc;
                if ((rc = run.fused(d3))) return rc;
                if ((rc = run.check())) return rc;
            }
            polled_last = false;
            if (3 * grp + 2 >= prm.itrmin) {
                if ((rc = run.poll(done))) return rc;
                polled_last = true;
            }
        }
    }
    hipLaunchKernelGGL(k_hsmac_end, dim3(1), dim3(1), 0, b.stream, b.S, slot);
    LCS_HIP(hipGetLastError());
    if (!polled_last && (rc = run.poll(done))) return rc;
    if (done) {
        // every sweep has exactly one fused launch; those after the converged
        // sweep returned at once
        const int executed = b.S_host->hs_itr + 1;
        for (size_t q = 0; q < run.sweep_bytes.size(); q++)
            if (run.sweep_first[q] >= executed) prof_uncount(b, run.sweep_name[q], 1, run.sweep_bytes[q]);
    }
    if (b.S_host->hs_par) {
        // the result lives in the spare set: swap it in
        for (int q = 0; q < 3; q++) std::swap(Us[q], b.spare[q]);
        std::swap(*Pp, b.spare[3]);
    }
    return LCS_OK;
}

int launch_hsmac(Box &b, void **Us, DtRef dt, const lcs_hsmac_params &prm, int slot, bool lf_deferred)
{
    if (prm.coloring != LCS_COLORING_RED_BLACK_FAST && prm.coloring != LCS_COLORING_RED_BLACK)
        return fail(LCS_ERR_INVALID, "unsupported coloring");
    if (prm.itrmax < 0 || prm.itrmin < 0) return fail(LCS_ERR_INVALID, "itrmin/itrmax must be >= 0");
    if (lf_deferred && prm.itrmax < 1) return fail(LCS_ERR_INVALID, "deferred forcing needs a sweep");
    return b.esize == 8 ? hsmac_t<double>(b, Us, dt, prm, slot, lf_deferred)
                        : hsmac_t<float>(b, Us, dt, prm, slot, lf_deferred);
}

// ---------------------------------------------------------------------------
// CFL time step (src/flows/timestep.jl:18-65): FaceHalo(h) fill, then max of
// |interp_o4| per axis, then dt on the device.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void cfl_block(const Geom &g, const T *__restrict__ U, const T *__restrict__ V,
                                 const T *__restrict__ W, int b, int nb, double m[3])
{
    const long sy = g.sx, sz = g.sx * g.sy;
    m[0] = m[1] = m[2] = 0.0;
    for_rows(g, b, nb, [&](int, int, int, long p) {
        T uc = interp_o4(U[p - 2], U[p - 1], U[p], U[p + 1]);
        T vc = interp_o4(V[p - 2 * sy], V[p - sy], V[p], V[p + sy]);
        T wc = interp_o4(W[p - 2 * sz], W[p - sz], W[p], W[p + sz]);
        m[0] = fmax(m[0], fabs((double)uc));
        m[1] = fmax(m[1], fabs((double)vc));
        m[2] = fmax(m[2], fabs((double)wc));
    });
}

__device__ inline int xcd_block()
{
    const int nb = gridDim.x;
    const int per_x = nb / kXcds, rem = nb % kXcds, x = blockIdx.x % kXcds, i0 = blockIdx.x / kXcds;
    return x < rem ? x * (per_x + 1) + i0 : rem * (per_x + 1) + (x - rem) * per_x + i0;
}

// Columns of a 64 x PY patch march up a chunk of zc planes, each keeping its
// last four W values in registers, so the z-stencil reads W once instead of
// four times (a row-walking block cannot keep four planes in L2 at 1024^2
// cells a plane).  U and V neighbours are loads of the same workgroup (L1
// and L2 hits) except the patch's own ring: with 16-row patches V reads 19
// rows per 16 from HBM (4-row patches read 7 per 4, and the counters showed
// the y-neighbour patch's rows did not stay in L2 for them: 1.67x the three
// arrays at 1024^3).  The maxima do not depend on the order, so dt is the
// row-walking kernel's to the bit.
constexpr int kCflPX = 64;
inline int cfl_patch_rows(const Geom &g) { return g.ny >= 64 ? 16 : 4; }

template <typename T, int PY>
__global__ void __launch_bounds__(kCflPX *PY) k_cfl(Geom g, const T *__restrict__ U, const T *__restrict__ V,
                                                   const T *__restrict__ W, double *part, int zc)
{
    __shared__ double lds[16];
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int x = 1 + bx * kCflPX + (int)threadIdx.x % kCflPX, y = 1 + by * PY + (int)threadIdx.x / kCflPX;
    const int z0 = 1 + bz * zc, z1 = min(g.nz, z0 + zc - 1);
    double m[3] = {0.0, 0.0, 0.0};
    if (x <= g.nx && y <= g.ny && z0 <= z1) {
        const long sy = g.sx, sz = g.sx * g.sy;
        long p = g.at(x, y, z0);
        T w0 = W[p - 2 * sz], w1 = W[p - sz], w2 = W[p];
        for (int z = z0; z <= z1; z++, p += sz) {
            const T w3 = W[p + sz];
            const T uc = interp_o4(U[p - 2], U[p - 1], U[p], U[p + 1]);
            const T vc = interp_o4(V[p - 2 * sy], V[p - sy], V[p], V[p + sy]);
            const T wc = interp_o4(w0, w1, w2, w3);
            m[0] = fmax(m[0], fabs((double)uc));
            m[1] = fmax(m[1], fabs((double)vc));
            m[2] = fmax(m[2], fabs((double)wc));
            w0 = w1;
            w1 = w2;
            w2 = w3;
        }
    }
    const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    for (int a = 0; a < 3; a++) {
        double s = block_max(m[a], lds);
        if (threadIdx.x == 0) part[(long)a * kMaxPartBlocks + blk] = s;
    }
}

// grid of the column-marching row kernels (CFL, forcing means): at most
// 4096 workgroups, z-chunks of at least 8 planes
struct ColGrid {
    int gx, gy, gz, zc, py;
    int blocks() const { return gx * gy * gz; }
};
inline ColGrid col_grid(const Geom &g)
{
    ColGrid c;
    c.py = cfl_patch_rows(g);
    c.gx = (g.nx + kCflPX - 1) / kCflPX;
    c.gy = (g.ny + c.py - 1) / c.py;
    c.gz = std::max(1, std::min((g.nz + 7) / 8, 4096 / (c.gx * c.gy)));
    c.zc = (g.nz + c.gz - 1) / c.gz;
    c.gz = (g.nz + c.zc - 1) / c.zc;
    return c;
}

// dt from the per-axis maxima (src/flows/timestep.jl:56-64)
__device__ inline void cfl_dt(DevScalars *S, const double mm[3], double cfl, double dtmax, double dx, double dy,
                              double dz, double Re)
{
    double umax = mm[0] + 1e-10, vmax = mm[1] + 1e-10, wmax = mm[2] + 1e-10;
    S->umax[0] = umax;
    S->umax[1] = vmax;
    S->umax[2] = wmax;
    double dt = dtmax;
    dt = fmin(dt, cfl * dx / umax);
    dt = fmin(dt, cfl * dy / vmax);
    dt = fmin(dt, cfl * dz / wmax);
    dt = fmin(dt, 0.5 * ((dx / 2) * (dx / 2)) / (1 / Re));
    S->dt_flow = dt;
    S->dt = dt;
}

__global__ void k_cfl_finalize(const double *part, int nblk, double cfl, double dtmax, double dx, double dy,
                               double dz, double Re, DevScalars *S)
{
    __shared__ double lds[16];
    double mm[3];
    for (int a = 0; a < 3; a++) {
        double m = 0.0;
        for (int q = threadIdx.x; q < nblk; q += blockDim.x) m = fmax(m, part[(long)a * kMaxPartBlocks + q]);
        mm[a] = block_max(m, lds);
    }
    if (threadIdx.x == 0) cfl_dt(S, mm, cfl, dtmax, dx, dy, dz, Re);
}

int launch_cfl(Box &b, double cfl, double dtmax, bool fill)
{
    void *arr[3] = {b.f[0], b.f[1], b.f[2]};
    int rc = fill ? launch_ghost_fill(b, arr, 3, b.g.h, 0, nullptr) : LCS_OK;
    if (rc) return rc;
    const ColGrid c = col_grid(b.g);
    const int blocks = c.blocks();
    if (blocks > kMaxPartBlocks) return fail(LCS_ERR_INVALID, "grid too large for the CFL reduction");
    ProfScope ps(b, "cfl", 3.0 * b.esize * b.g.cells());
    const dim3 grid(c.gx, c.gy, c.gz);
#define CFL_LAUNCH(T, PY)                                                                                      \
    hipLaunchKernelGGL((k_cfl<T, PY>), grid, dim3(kCflPX * PY), 0, b.stream, b.g, (const T *)b.f[0],          \
                       (const T *)b.f[1], (const T *)b.f[2], b.part, c.zc)
    if (b.esize == 8) {
        if (c.py == 16) CFL_LAUNCH(double, 16); else CFL_LAUNCH(double, 4);
    } else {
        if (c.py == 16) CFL_LAUNCH(float, 16); else CFL_LAUNCH(float, 4);
    }
#undef CFL_LAUNCH
    LCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_cfl_finalize, dim3(1), dim3(256), 0, b.stream, b.part, blocks, cfl, dtmax,
                       spacing(b.g.nx), spacing(b.g.ny), spacing(b.g.nz), b.Re(), b.S);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// ---------------------------------------------------------------------------
// Constant-power linear forcing (src/flows/forcings/lf.jl:27-106)
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void lf_means_block(const Geom &g, const T *__restrict__ U, const T *__restrict__ V,
                                      const T *__restrict__ W, int b, int nb, double s[4])
{
    const long sy = g.sx, sz = g.sx * g.sy;
    s[0] = s[1] = s[2] = s[3] = 0.0;
    for_rows(g, b, nb, [&](int, int, int, long p) {
        T u = interp_o2(U[p - 1], U[p]);
        T v = interp_o2(V[p - sy], V[p]);
        T w = interp_o2(W[p - sz], W[p]);
        T u2 = (u * u + v * v) + w * w;
        s[0] += (double)u2;
        s[1] += (double)U[p];
        s[2] += (double)V[p];
        s[3] += (double)W[p];
    });
}

// The means of one box march columns of a 64 x PY patch up a chunk of
// planes like k_cfl (W of the plane below stays in a register), so each of
// U, V, W is read about once from HBM; a row-walking pass read V and W twice
// (the row below and the plane below were no longer in L2).  Sums are per
// workgroup in a fixed order, combined in block order by k_lf_finalize.
template <typename T, int PY>
__global__ void __launch_bounds__(kCflPX *PY) k_lf_means(Geom g, const T *__restrict__ U, const T *__restrict__ V,
                                                        const T *__restrict__ W, double *part, int zc)
{
    __shared__ double lds[16];
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int x = 1 + bx * kCflPX + (int)threadIdx.x % kCflPX, y = 1 + by * PY + (int)threadIdx.x / kCflPX;
    const int z0 = 1 + bz * zc, z1 = min(g.nz, z0 + zc - 1);
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (x <= g.nx && y <= g.ny && z0 <= z1) {
        const long sy = g.sx, sz = g.sx * g.sy;
        long p = g.at(x, y, z0);
        T wm = W[p - sz];
        for (int z = z0; z <= z1; z++, p += sz) {
            const T u0 = U[p], v0 = V[p], w0 = W[p];
            const T u = interp_o2(U[p - 1], u0);
            const T v = interp_o2(V[p - sy], v0);
            const T w = interp_o2(wm, w0);
            const T u2 = (u * u + v * v) + w * w;
            s[0] += (double)u2;
            s[1] += (double)u0;
            s[2] += (double)v0;
            s[3] += (double)w0;
            wm = w0;
        }
    }
    const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    for (int a = 0; a < 4; a++) {
        double r = block_sum(s[a], lds);
        if (threadIdx.x == 0) part[(long)a * kMaxPartBlocks + blk] = r;
    }
}

// e and A = P0/(2e) from the four means (src/flows/forcings/lf.jl:56-64)
__device__ inline void lf_coef(DevScalars *S, const double m[4], double len, double power)
{
    double u2 = m[0] / len, u = m[1] / len, v = m[2] / len, w = m[3] / len;
    double e = 0.5 * (((u2 - u * u) - v * v) - w * w);
    S->lf[0] = u2;
    S->lf[1] = u;
    S->lf[2] = v;
    S->lf[3] = w;
    S->lf[4] = power / (2 * e);
}

__global__ void k_lf_finalize(const double *part, int nblk, double len, double power, DevScalars *S)
{
    __shared__ double lds[16];
    double m[4];
    for (int a = 0; a < 4; a++) {
        double s = 0.0;
        for (int q = threadIdx.x; q < nblk; q += blockDim.x) s += part[(long)a * kMaxPartBlocks + q];
        m[a] = block_sum(s, lds);
    }
    if (threadIdx.x == 0) lf_coef(S, m, len, power);
}

template <typename T>
__device__ inline double lf_apply_block(const Geom &g, T *U, T *V, T *W, const DevScalars *S, T rdt, int b, int nb)
{
    const T A = T(S->lf[4]), um = T(S->lf[1]), vm = T(S->lf[2]), wm = T(S->lf[3]);
    double sp = 0.0;
    for_rows(g, b, nb, [&](int, int, int, long p) {
        T u = U[p], v = V[p], w = W[p];
        T usrc = A * (u - um), vsrc = A * (v - vm), wsrc = A * (w - wm);
        T pw = (u * usrc + v * vsrc) + w * wsrc;
        U[p] = u + usrc * rdt;
        V[p] = v + vsrc * rdt;
        W[p] = w + wsrc * rdt;
        sp += (double)pw;
    });
    return sp;
}

template <typename T>
__global__ void __launch_bounds__(256) k_lf_apply(Geom g, T *U, T *V, T *W, const DevScalars *S, DtRef dt,
                                                 double *part)
{
    __shared__ double lds[16];
    const T rdt = T(dt.factor * (*dt.src));
    double r = block_sum(lf_apply_block(g, U, V, W, S, rdt, xcd_block(), gridDim.x), lds);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void k_mean_to(const double *part, int nblk, double len, double *dst)
{
    __shared__ double lds[16];
    double s = 0.0;
    for (int q = threadIdx.x; q < nblk; q += blockDim.x) s += part[q];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) *dst = s / len;
}

// ---------------------------------------------------------------------------
// Flow statistics (Flows.stat, src/flows/stat.jl:67-155): one pass over the
// interior for the summaries of U, V, W and the means of u2, 2 S_ij S_ij and
// the powers of dudx; the integral length (stat.jl:157-182) in its own pass.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T diff1_o4_half(T a, T b, T c, T d)  // src/diff.jl:13-15
{
    return T(1.0 / 24.0) * (((a - T(27) * b) + T(27) * c) - d);
}

constexpr int kStatQ = 14;  // U, V, W (sum, max, min), then u2, s2, d2, d3, d4 sums

template <typename T>
__global__ void __launch_bounds__(256) k_flow_stat(Geom g, const T *__restrict__ U, const T *__restrict__ V,
                                                  const T *__restrict__ W, T dxi, T dyi, T dzi, double *part)
{
    __shared__ double lds[16];
    const long sy = g.sx, sz = g.sx * g.sy;
    double s[kStatQ];
    for (int a = 0; a < 3; a++) {
        s[3 * a] = 0.0;
        s[3 * a + 1] = -INFINITY;
        s[3 * a + 2] = INFINITY;
    }
    for (int a = 9; a < kStatQ; a++) s[a] = 0.0;
    for_rows(g, xcd_block(), gridDim.x, [&](int, int, int, long p) {
        const T u = U[p], v = V[p], w = W[p];
        const T u2 = (u * u + v * v) + w * w;
        const T dudx = dxi * diff1_o4_half(U[p - 2], U[p - 1], U[p], U[p + 1]);
        const T dvdy = dyi * diff1_o4_half(V[p - 2 * sy], V[p - sy], V[p], V[p + sy]);
        const T dwdz = dzi * diff1_o4_half(W[p - 2 * sz], W[p - sz], W[p], W[p + sz]);
        const T dudy = dyi * diff1_o4_half(U[p - 2 * sy], U[p - sy], U[p], U[p + sy]);
        const T dvdx = dxi * diff1_o4_half(V[p - sy - 1], V[p - sy], V[p - sy + 1], V[p - sy + 2]);
        const T dvdz = dzi * diff1_o4_half(V[p - sy - sz], V[p - sy], V[p - sy + sz], V[p - sy + 2 * sz]);
        const T dwdy = dyi * diff1_o4_half(W[p - 2 * sy], W[p - sy], W[p], W[p + sy]);
        const T dwdx = dxi * diff1_o4_half(W[p - sz - 1], W[p - sz], W[p - sz + 1], W[p - sz + 2]);
        const T dudz = dzi * diff1_o4_half(U[p - 2 * sz], U[p - sz], U[p], U[p + sz]);
        const T s11 = T(0.5) * (dudx + dudx), s22 = T(0.5) * (dvdy + dvdy), s33 = T(0.5) * (dwdz + dwdz);
        const T s12 = T(0.5) * (dudy + dvdx), s13 = T(0.5) * (dudz + dwdx), s23 = T(0.5) * (dvdz + dwdy);
        const T s2 = T(2) * (((s11 * s11 + s22 * s22) + s33 * s33) + T(2) * ((s12 * s12 + s13 * s13) + s23 * s23));
        const T d2 = dudx * dudx;
        const T vals[3] = {u, v, w};
        for (int a = 0; a < 3; a++) {
            s[3 * a] += (double)vals[a];
            s[3 * a + 1] = fmax(s[3 * a + 1], (double)vals[a]);
            s[3 * a + 2] = fmin(s[3 * a + 2], (double)vals[a]);
        }
        s[9] += (double)u2;
        s[10] += (double)s2;
        s[11] += (double)d2;
        s[12] += (double)(d2 * dudx);
        s[13] += (double)(d2 * d2);
    });
    for (int a = 0; a < kStatQ; a++) {
        const int kind = a < 9 ? a % 3 : 0;  // 0 sum, 1 max, 2 min
        double r = kind == 0 ? block_sum(s[a], lds) : kind == 1 ? block_max(s[a], lds) : -block_max(-s[a], lds);
        if (threadIdx.x == 0) part[(long)a * kMaxPartBlocks + blockIdx.x] = r;
    }
}

__global__ void k_flow_stat_fin(const double *part, int nblk, double len, double *out)
{
    __shared__ double lds[16];
    for (int a = 0; a < kStatQ; a++) {
        const int kind = a < 9 ? a % 3 : 0;
        double r = kind == 0 ? 0.0 : kind == 1 ? -INFINITY : INFINITY;
        for (int q = threadIdx.x; q < nblk; q += blockDim.x) {
            const double x = part[(long)a * kMaxPartBlocks + q];
            r = kind == 0 ? r + x : kind == 1 ? fmax(r, x) : fmin(r, x);
        }
        r = kind == 0 ? block_sum(r, lds) : kind == 1 ? block_max(r, lds) : -block_max(-r, lds);
        if (threadIdx.x == 0) out[a] = kind == 0 ? r / len : r;
    }
}

// Integral length (stat.jl:157-182).  The reference forms R = irfft(|rfft(U)|^2)
// along x, i.e. the circular autocorrelation R[r] = sum_i U[i] U[i+r], and
// sums lags 0..n/2 up to the first negative one: Ls = acc dx / R[0].  Here
// each block takes a run of rows; a row sits twice in LDS (so i+r needs no
// wrap) and each thread forms one lag directly, 256 lags at a time, stopping
// at the first chunk that holds a negative lag.
template <typename T>
__global__ void __launch_bounds__(256) k_integral_length(Geom g, const T *__restrict__ U, double dx, double *part)
{
    extern __shared__ double row[];  // 2 nx
    __shared__ double lag[256];
    __shared__ double lds[16];
    __shared__ int stop;
    const int nx = g.nx, mid = nx / 2 + 1, rows = g.ny * g.nz;
    const int per = (rows + gridDim.x - 1) / gridDim.x, r0 = blockIdx.x * per, r1 = min(rows, r0 + per);
    double sum_ls = 0.0;
    for (int r = r0; r < r1; r++) {
        const long base = g.at(1, r % g.ny + 1, r / g.ny + 1);
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const double v = (double)U[base + i];
            row[i] = v;
            row[i + nx] = v;
        }
        __syncthreads();
        double acc = 0.0, rzero = 0.0;
        for (int l0 = 0; l0 < mid; l0 += 256) {
            const int l = l0 + (int)threadIdx.x;
            double c = 0.0;
            if (l < mid)
                for (int i = 0; i < nx; i++) c += row[i] * row[i + l];
            lag[threadIdx.x] = c;
            __syncthreads();
            if (threadIdx.x == 0) {
                int brk = 0;
                for (int q = 0; q < 256 && l0 + q < mid; q++) {
                    if (l0 + q == 0) rzero = lag[0];
                    if (lag[q] < 0) {
                        brk = 1;
                        break;
                    }
                    acc += lag[q];
                }
                stop = brk;
            }
            __syncthreads();
            if (stop) break;
        }
        if (threadIdx.x == 0) sum_ls += acc * dx / rzero;
        __syncthreads();  // the row is rewritten next
    }
    double s = block_sum(threadIdx.x == 0 ? sum_ls : 0.0, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

int launch_flow_stat(Box &b, int integral_length, double *out_dev)
{
    void *arr[3] = {b.f[0], b.f[1], b.f[2]};
    int rc = launch_ghost_fill(b, arr, 3, b.g.h, 1, nullptr);  // FullHalo on Us (stat.jl:79)
    if (rc) return rc;
    const Geom &g = b.g;
    const int blocks = grid_rows(g);
    const double dxi = 1.0 / spacing(g.nx), dyi = 1.0 / spacing(g.ny), dzi = 1.0 / spacing(g.nz);
    {
        ProfScope ps(b, "flow_stat", 3.0 * b.esize * (double)g.cells());
        if (b.esize == 8)
            hipLaunchKernelGGL(k_flow_stat<double>, dim3(blocks), dim3(256), 0, b.stream, g, (const double *)b.f[0],
                               (const double *)b.f[1], (const double *)b.f[2], dxi, dyi, dzi, b.part);
        else
            hipLaunchKernelGGL(k_flow_stat<float>, dim3(blocks), dim3(256), 0, b.stream, g, (const float *)b.f[0],
                               (const float *)b.f[1], (const float *)b.f[2], (float)dxi, (float)dyi, (float)dzi,
                               b.part);
        LCS_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_flow_stat_fin, dim3(1), dim3(256), 0, b.stream, b.part, blocks, (double)g.cells(), out_dev);
    LCS_HIP(hipGetLastError());
    if (!integral_length) return LCS_OK;
    if (g.nx > 4096) return fail(LCS_ERR_INVALID, "integral length needs nx <= 4096");
    double *part = b.part + (long)kStatQ * kMaxPartBlocks;
    const size_t shm = sizeof(double) * 2 * (size_t)g.nx;
    {
        ProfScope ps(b, "integral_length", (double)b.esize * (double)g.cells());
        if (b.esize == 8)
            hipLaunchKernelGGL(k_integral_length<double>, dim3(blocks), dim3(256), shm, b.stream, g,
                               (const double *)b.f[0], spacing(g.nx), part);
        else
            hipLaunchKernelGGL(k_integral_length<float>, dim3(blocks), dim3(256), shm, b.stream, g,
                               (const float *)b.f[0], spacing(g.nx), part);
        LCS_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mean_to, dim3(1), dim3(256), 0, b.stream, part, blocks, (double)g.ny * g.nz, out_dev + kStatQ);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// The means and A of the linear forcing into S->lf (lf.jl:40-66).
int launch_force_linear_coef(Box &b, double power, bool face_only)
{
    void *arr[3] = {b.f[0], b.f[1], b.f[2]};
    int rc = face_only ? launch_ghost_fill(b, arr, 3, 1, 0, nullptr) : launch_ghost_fill(b, arr, 3, b.g.h, 1, nullptr);
    if (rc) return rc;
    const ColGrid c = col_grid(b.g);
    const int blocks = c.blocks();
    if (blocks > kMaxPartBlocks) return fail(LCS_ERR_INVALID, "grid too large for the forcing reduction");
    double len = (double)b.g.cells();
    {
    ProfScope ps(b, "lf_means", 3.0 * b.esize * len);
    const dim3 grid(c.gx, c.gy, c.gz);
#define LF_LAUNCH(T, PY)                                                                                       \
    hipLaunchKernelGGL((k_lf_means<T, PY>), grid, dim3(kCflPX * PY), 0, b.stream, b.g, (const T *)b.f[0],     \
                       (const T *)b.f[1], (const T *)b.f[2], b.part, c.zc)
    if (b.esize == 8) {
        if (c.py == 16) LF_LAUNCH(double, 16); else LF_LAUNCH(double, 4);
    } else {
        if (c.py == 16) LF_LAUNCH(float, 16); else LF_LAUNCH(float, 4);
    }
#undef LF_LAUNCH
    LCS_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_lf_finalize, dim3(1), dim3(256), 0, b.stream, b.part, blocks, len, power, b.S);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

int launch_force_linear(Box &b, double power, DtRef dt)
{
    int rc = launch_force_linear_coef(b, power);
    if (rc) return rc;
    const int blocks = grid_rows(b.g);
    const double len = (double)b.g.cells();
    {
    ProfScope ps(b, "lf_apply", 6.0 * b.esize * len);
    if (b.esize == 8)
        hipLaunchKernelGGL(k_lf_apply<double>, dim3(blocks), dim3(256), 0, b.stream, b.g, (double *)b.f[0],
                           (double *)b.f[1], (double *)b.f[2], b.S, dt, b.part);
    else
        hipLaunchKernelGGL(k_lf_apply<float>, dim3(blocks), dim3(256), 0, b.stream, b.g, (float *)b.f[0],
                           (float *)b.f[1], (float *)b.f[2], b.S, dt, b.part);
    LCS_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mean_to, dim3(1), dim3(256), 0, b.stream, b.part, blocks, len, &b.S->p);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// ---------------------------------------------------------------------------
// rc-energy forcing (src/flows/forcings/rcf.jl, lib/fouriertools)
// rcf buffer layout (doubles): [0, 3c^3) coarse grids U,V,W;
// then hat: 6 doubles per crop slot; then crop: 2*3*ncrop.
// ---------------------------------------------------------------------------
// Box mean over (n/c)^3 blocks (ext/AMDGPUExt.jl:18-41): one block per coarse
// cell and component, deterministic tree sum, times the reciprocal weight.
// (the body of one block, shared with the batched kernel)
template <typename T>
__device__ inline void boxmean_block(const Geom &g, const T *A, int c, int cell, int comp, double *coarse,
                                     double *lds)
{
    int ic = cell % c, jc = (cell / c) % c, kc = cell / (c * c);
    int kx = g.nx / c, ky = g.ny / c, kz = g.nz / c;
    long nb = (long)kx * ky * kz;
    double s = 0.0;
    // element t = di + kx (dj + ky dk), t = tid, tid + B, ...: the indices
    // advance by B in mixed radix (kx, ky) with carries, not divisions
    const int B = blockDim.x;
    int di = threadIdx.x % kx, dj = (threadIdx.x / kx) % ky, dk = threadIdx.x / (kx * ky);
    const int bx = B % kx, by = (B / kx) % ky, bz = B / (kx * ky);
    for (long t = threadIdx.x; t < nb; t += B) {
        s += (double)A[g.at(ic * kx + di + 1, jc * ky + dj + 1, kc * kz + dk + 1)];
        di += bx;
        int carry = di >= kx ? 1 : 0;
        di -= carry * kx;
        dj += by + carry;
        carry = dj >= ky ? 1 : 0;
        dj -= carry * ky;
        dk += bz + carry;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) {
        T w = T(1) / T((double)kx * ky * kz);
        coarse[(long)comp * c * c * c + cell] = (double)(T(s) * w);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_boxmean(Geom g, const T *U, const T *V, const T *W, int c,
                                                double *coarse)
{
    __shared__ double lds[16];
    const int comp = blockIdx.y;
    boxmean_block<T>(g, comp == 0 ? U : (comp == 1 ? V : W), c, blockIdx.x, comp, coarse, lds);
}

__host__ __device__ inline int fftidx(int k, int n) { return k >= 0 ? k : n + k; }
__host__ __device__ inline int rcf_crop_slots_dev(int kmax) { return (kmax + 1) * (2 * kmax + 1) * (2 * kmax + 1); }

// Real DFT (FFTW sign, unnormalised) of each coarse grid at the wavevector
// whose rfft-array index is (kx, fftindex(ky,c), fftindex(kz,c)); one block
// per crop slot (kx in 0:kmax, ky,kz in -kmax:kmax).
__device__ inline void rcf_dft_block(const double *coarse, int c, int kmax, int slot, double *hat, double *lds)
{
    int nk = 2 * kmax + 1;
    int kx = slot % (kmax + 1);
    int ky = (slot / (kmax + 1)) % nk - kmax;
    int kz = slot / ((kmax + 1) * nk) - kmax;
    int fx = kx, fy = fftidx(ky, c), fz = fftidx(kz, c);
    long den = (long)c * c * c;
    long n3 = den;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (long t = threadIdx.x; t < n3; t += blockDim.x) {
        int x = (int)(t % c), y = (int)((t / c) % c), z = (int)(t / ((long)c * c));
        long num = (long)fx * x * c * c + (long)fy * y * c * c + (long)fz * z * c * c;
        long m = num % den;
        double sn, cs;
        sincospi(-2.0 * (double)m / (double)den, &sn, &cs);
        for (int a = 0; a < 3; a++) {
            double v = coarse[(long)a * n3 + t];
            acc[2 * a] += v * cs;
            acc[2 * a + 1] += v * sn;
        }
    }
    for (int q = 0; q < 6; q++) {
        double s = block_sum(acc[q], lds);
        if (threadIdx.x == 0) hat[6L * slot + q] = s;
    }
}

__global__ void __launch_bounds__(256) k_rcf_dft(const double *coarse, int c, int kmax, double *hat)
{
    __shared__ double lds[16];
    rcf_dft_block(coarse, c, kmax, blockIdx.x, hat, lds);
}

__device__ inline int kcount_dev(int k0)
{
    int c = 0;
    double lo = (k0 - 0.5) * (k0 - 0.5), hi = (k0 + 0.5) * (k0 + 0.5);
    for (int kz = -k0; kz <= k0; kz++)
        for (int ky = -k0; ky <= k0; ky++)
            for (int kx = -k0; kx <= k0; kx++) {
                int r2 = kx * kx + ky * ky + kz * kz;
                if (lo <= r2 && r2 < hi) c++;
            }
    return c;
}

// enespe + shell rescaling + crop (rcf.jl:49-67, FourierTools.jl:103-157) on
// one thread: a few hundred operations.  Aliased wavevectors (c < 2kmax+1)
// share one rfft entry, as in the reference, via rep[].
__device__ inline void rcf_spectrum_one(double *hat, int c, int kmin, int kmax, double e10, double *crop,
                                        DevScalars *S)
{
    int nk = 2 * kmax + 1;
    int ncrop = (kmax + 1) * nk * nk;
    auto slot_of = [&](int kx, int ky, int kz) { return kx + (kmax + 1) * ((ky + kmax) + nk * (kz + kmax)); };
    auto rep_of = [&](int kx, int ky, int kz) {
        int fy = fftidx(ky, c), fz = fftidx(kz, c);
        for (int z = -kmax; z <= kmax; z++)
            for (int y = -kmax; y <= kmax; y++)
                if (fftidx(y, c) == fy && fftidx(z, c) == fz) return slot_of(kx, y, z);
        return slot_of(kx, ky, kz);
    };
    double norm = (double)((c / 2 + 1) * 2 - 2) * c * c;
    for (int k0 = kmin; k0 <= kmax; k0++) {
        double e = 0;
        double lo = (k0 - 0.5) * (k0 - 0.5), hi = (k0 + 0.5) * (k0 + 0.5);
        for (int kz = -k0; kz <= k0; kz++)
            for (int ky = -k0; ky <= k0; ky++)
                for (int kx = 0; kx <= k0; kx++) {
                    int r2 = kx * kx + ky * ky + kz * kz;
                    if (!(lo <= r2 && r2 < hi)) continue;
                    const double *hq = hat + 6L * rep_of(kx, ky, kz);
                    double w = (kx == 0) ? 1 : 2;
                    double s = 0;
                    for (int a = 0; a < 3; a++) {
                        double re = hq[2 * a] / norm, im = hq[2 * a + 1] / norm;
                        double ab = re * re + im * im;
                        s = (a == 0) ? ab : s + ab;
                    }
                    e += w * s;
                }
        double en = e / 2;
        if (k0 - kmin < 8) S->rcf_en[k0 - kmin] = en;
        double c1 = (double)kcount_dev(1), ck = (double)kcount_dev(k0);
        double A = k0 + 0.5, B = k0 - 0.5;
        double en0 = e10 * pow((double)k0, -5.0 / 3.0) * (1.5 * 1.5 * 1.5 - 0.5 * 0.5 * 0.5) /
                     ((A * A * A) - (B * B * B)) * ck / c1;
        double r = sqrt(en0 / fmax(en, 1e-8)) - 1.0;
        for (int kz = -kmax; kz <= kmax; kz++)
            for (int ky = -kmax; ky <= kmax; ky++)
                for (int kx = 0; kx <= kmax; kx++) {
                    int r2 = kx * kx + ky * ky + kz * kz;
                    if (!(lo <= r2 && r2 < hi)) continue;
                    double *hq = hat + 6L * rep_of(kx, ky, kz);
                    for (int a = 0; a < 6; a++) hq[a] *= r;
                }
    }
    for (int kz = -kmax; kz <= kmax; kz++)
        for (int ky = -kmax; ky <= kmax; ky++)
            for (int kx = 0; kx <= kmax; kx++) {
                const double *hq = hat + 6L * rep_of(kx, ky, kz);
                int o = kx + (kmax + 1) * (fftidx(ky, nk) + nk * fftidx(kz, nk));
                for (int a = 0; a < 3; a++) {
                    crop[2 * ((long)a * ncrop + o)] = hq[2 * a] / norm;
                    crop[2 * ((long)a * ncrop + o) + 1] = hq[2 * a + 1] / norm;
                }
            }
    S->e10 = e10;
}

__global__ void k_rcf_spectrum(double *hat, int c, int kmin, int kmax, double e10, double *crop, DevScalars *S)
{
    if (threadIdx.x == 0) rcf_spectrum_one(hat, c, kmin, kmax, e10, crop, S);
}

struct Cx {
    double re, im;
};
__device__ inline Cx cx_mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline Cx cx_pow(Cx x, int n)
{
    if (n == 0) return {1.0, 0.0};
    if (n < 0) {
        Cx r = cx_pow(x, -n);
        return {r.re, -r.im};
    }
    Cx r = x;
    for (int m = 1; m < n; m++) r = cx_mul(x, r);
    return r;
}

constexpr int kMaxRcfModes = 128;
constexpr int kMaxRcfKx = 8;  // kx in 0..kmax; the mode cap keeps kmax <= 4
struct RcfModes {
    int n;
    int kmin;
    int kof[8], kn[8];
    signed char k[kMaxRcfModes][3];
    // the modes grouped by kx (kset order within a group): group kx is
    // byx[xof[kx] .. xof[kx + 1])
    unsigned char byx[kMaxRcfModes];
    int xof[kMaxRcfKx + 1];
};

// per-block scratch of the apply: each (row, mode)'s A_hat(k) e^{i(ky y + kz z)}
// for the rows in hand, then their sums over the modes of one kx
constexpr int kRcfCellsPerPass = 1024;  // 4 cells per thread of a 256-thread block
constexpr int kRcfRowSlots = 256;       // rows x modes at once
constexpr int kRcfMaxRows = 16;
struct RcfRowLds {
    double a[kMaxRcfModes][6];  // A_hat(k) of each mode, three components
    double t[kRcfRowSlots][6];
    double b[kRcfMaxRows][3][kMaxRcfKx][2];
};

// rows a block takes at once: up to 1024 cells, within the LDS slots
__device__ inline int rcf_rows_per_group(const Geom &g, int nmodes)
{
    int R = kRcfCellsPerPass / g.nx;
    R = min(R, kRcfRowSlots / max(nmodes, 1));
    return max(1, min(R, kRcfMaxRows));
}

// Sparse inverse DFT + apply (rcf.jl:74-98; sirdft FourierTools.jl:165-208):
// src(x) = sum over the forced modes of 2 Re(A_hat(k) e^{ik.x}), with the
// phases cispi from the global index raised to integer powers, U += src,
// power = u.src/dt.  The reference evaluates every mode at every cell; here
// a block takes R x-rows (j, k) at a time, folds the y and z phases into the
// amplitudes once per row, B(kx) = sum_{ky,kz} A_hat(k) e^{i(ky y + kz z)},
// and a cell sums kmax + 1 terms 2 Re(B(kx) e^{i kx x}): the same sum, its
// terms grouped by kx.  The rows' velocities are loaded before the row sums
// are formed, so their latency hides behind that work.  Row groups blk,
// blk + nblk, ... of the box; returns the thread's part of the forcing power.
template <typename T>
__device__ inline double rcf_apply_part(const Geom &g, T *U, T *V, T *W, int c, int kmax, const RcfModes &md,
                                        const double *crop, double dti, int blk, int nblk, RcfRowLds &L)
{
    constexpr int Q = kRcfCellsPerPass / 256;
    const int nk = 2 * kmax + 1, ncrop = (kmax + 1) * nk * nk, nkx = kmax + 1;
    const double phi = (double)(g.nx / c) / 2.0;
    const double nxi = 1.0 / g.nx, nyi = 1.0 / g.ny, nzi = 1.0 / g.nz;
    const long rows = (long)g.ny * g.nz;
    const int R = rcf_rows_per_group(g, md.n);
    const long groups = (rows + R - 1) / R;
    for (int m = threadIdx.x; m < md.n; m += blockDim.x) {
        const int ky = md.k[m][1], kz = md.k[m][2];
        const int o = md.k[m][0] + (kmax + 1) * (fftidx(ky, nk) + nk * fftidx(kz, nk));
        for (int a = 0; a < 3; a++) {
            L.a[m][2 * a] = crop[2 * ((long)a * ncrop + o)];
            L.a[m][2 * a + 1] = crop[2 * ((long)a * ncrop + o) + 1];
        }
    }
    __syncthreads();
    double sp = 0.0;
    for (long gi = blk; gi < groups; gi += nblk) {  // uniform over the block: every barrier is reached by all
        const long r0 = gi * R;
        const int nr = (int)min((long)R, rows - r0);
        const int ncell = nr * g.nx;  // more than one pass only when R = 1
        const int j0 = (int)(r0 % g.ny), k0 = (int)(r0 / g.ny);  // row r0 at (j0 + 1, k0 + 1)
        for (int x0 = 0; x0 < ncell; x0 += kRcfCellsPerPass) {
            long p[Q];
            int rr[Q], ii[Q];
            T u[Q], v[Q], w[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int e = x0 + threadIdx.x + 256 * q;
                rr[q] = R == 1 ? 0 : e / g.nx;
                ii[q] = e - rr[q] * g.nx + 1;
                if (e < ncell) {
                    const int jr = j0 + rr[q];  // 32-bit: row r0 + rr in (j, k)
                    const int j = R == 1 ? j0 : jr % g.ny, k = R == 1 ? k0 : k0 + jr / g.ny;
                    p[q] = g.at(ii[q], j + 1, k + 1);
                    u[q] = U[p[q]];
                    v[q] = V[p[q]];
                    w[q] = W[p[q]];
                }
            }
            if (x0 == 0) {
                for (int it = threadIdx.x; it < nr * md.n; it += blockDim.x) {
                    const int q = it / md.n, m = it - q * md.n;
                    const int jr = j0 + q, j = jr % g.ny + 1, k = k0 + jr / g.ny + 1;
                    Cx ej, ek;
                    sincospi(2 * ((double)k - phi) * nzi, &ek.im, &ek.re);
                    sincospi(2 * ((double)j - phi) * nyi, &ej.im, &ej.re);
                    const Cx yz = cx_mul(cx_pow(ej, md.k[m][1]), cx_pow(ek, md.k[m][2]));
                    for (int a = 0; a < 3; a++) {
                        const Cx A = {L.a[m][2 * a], L.a[m][2 * a + 1]};
                        const Cx t = cx_mul(A, yz);
                        L.t[it][2 * a] = t.re;
                        L.t[it][2 * a + 1] = t.im;
                    }
                }
                __syncthreads();  // (also: every thread is past the previous group's reads of L.b)
                for (int it = threadIdx.x; it < nr * 3 * nkx; it += blockDim.x) {
                    const int q = it / (3 * nkx), a = (it / nkx) % 3, kx = it % nkx;
                    const double(*t)[6] = L.t + q * md.n;
                    double re = 0.0, im = 0.0;
                    for (int s = md.xof[kx]; s < md.xof[kx + 1]; s++) {
                        re += t[md.byx[s]][2 * a];
                        im += t[md.byx[s]][2 * a + 1];
                    }
                    L.b[q][a][kx][0] = re;
                    L.b[q][a][kx][1] = im;
                }
                __syncthreads();
            }
#pragma unroll
            for (int q = 0; q < Q; q++) {
                if (x0 + (int)threadIdx.x + 256 * q >= ncell) continue;
                Cx ei;
                sincospi(2 * ((double)ii[q] - phi) * nxi, &ei.im, &ei.re);
                const double(*bq)[kMaxRcfKx][2] = L.b[rr[q]];
                double src[3] = {0, 0, 0};
                Cx e = {1.0, 0.0};
                for (int kx = 0; kx <= kmax; kx++) {
                    if (kx) e = cx_mul(ei, e);
                    for (int a = 0; a < 3; a++) src[a] += 2 * (bq[a][kx][0] * e.re - bq[a][kx][1] * e.im);
                }
                const T usrc = T(src[0]), vsrc = T(src[1]), wsrc = T(src[2]);
                const T pw = ((u[q] * usrc + v[q] * vsrc) + w[q] * wsrc) * T(dti);
                U[p[q]] = u[q] + usrc;
                V[p[q]] = v[q] + vsrc;
                W[p[q]] = w[q] + wsrc;
                sp += (double)pw;
            }
        }
    }
    return sp;
}

template <typename T>
__global__ void __launch_bounds__(256) k_rcf_apply(Geom g, T *U, T *V, T *W, int c, int kmax, RcfModes md,
                                                  const double *crop, DtRef dt, double *part)
{
    __shared__ double lds[16];
    __shared__ RcfRowLds rl;
    __shared__ RcfModes mds;  // the mode table in LDS, off the vector-memory queue of the field loads
    for (int q = threadIdx.x; q < (int)(sizeof(RcfModes) / sizeof(int)); q += blockDim.x)
        ((int *)&mds)[q] = ((const int *)&md)[q];
    __syncthreads();
    const double dti = 1.0 / (dt.factor * (*dt.src));
    double r = block_sum(rcf_apply_part(g, U, V, W, c, kmax, mds, crop, dti, blockIdx.x, gridDim.x, rl), lds);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

static int kset_host(int k0, signed char (*out)[3], int cap)
{
    int n = 0;
    double lo = (k0 - 0.5) * (k0 - 0.5), hi = (k0 + 0.5) * (k0 + 0.5);
    for (int kz = -k0; kz <= k0; kz++)
        for (int ky = -k0; ky <= k0; ky++)
            for (int kx = 0; kx <= k0; kx++) {
                int r2 = kx * kx + ky * ky + kz * kz;
                if (!(lo <= r2 && r2 < hi)) continue;
                int v[3] = {kx, ky, kz};
                int sgn = 0;
                for (int a = 0; a < 3 && sgn == 0; a++) sgn = v[a] > 0 ? 1 : (v[a] < 0 ? -1 : 0);
                if (sgn < 0)
                    for (int a = 0; a < 3; a++) v[a] = -v[a];
                bool dup = false;
                for (int m = 0; m < n && !dup; m++) dup = out[m][0] == v[0] && out[m][1] == v[1] && out[m][2] == v[2];
                if (dup) continue;
                if (n >= cap) return -1;
                for (int a = 0; a < 3; a++) out[n][a] = (signed char)v[a];
                n++;
            }
    return n;
}

size_t rcf_scratch_doubles(int c, int kmax)
{
    size_t nk = 2 * (size_t)kmax + 1, ncrop = ((size_t)kmax + 1) * nk * nk;
    return (size_t)3 * c * c * c + 6 * ncrop + 6 * ncrop;
}

static int rcf_buffers(Box &b, int c, int kmax)
{
    size_t need = sizeof(double) * rcf_scratch_doubles(c, kmax);
    if (need > b.rcf_bytes) {
        if (b.rcf) LCS_HIP(hipFree(b.rcf));
        b.rcf = nullptr;
        LCS_HIP(hipMalloc(&b.rcf, need));
        b.bytes += need - b.rcf_bytes;
        b.rcf_bytes = need;
    }
    return LCS_OK;
}

int launch_boxmean(Box &b, int c, double *coarse_dev)
{
    const Geom &g = b.g;
    if (c < 1 || g.nx % c || g.ny % c || g.nz % c)
        return fail(LCS_ERR_INVALID, "coarse_size must divide the grid dims");
    dim3 grid(c * c * c, 3);
    ProfScope ps(b, "boxmean", 3.0 * b.esize * g.cells());
    if (b.esize == 8)
        hipLaunchKernelGGL(k_boxmean<double>, grid, dim3(256), 0, b.stream, g, (const double *)b.f[0],
                           (const double *)b.f[1], (const double *)b.f[2], c, coarse_dev);
    else
        hipLaunchKernelGGL(k_boxmean<float>, grid, dim3(256), 0, b.stream, g, (const float *)b.f[0],
                           (const float *)b.f[1], (const float *)b.f[2], c, coarse_dev);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

static int rcf_check(int kmin, int kmax, int c)
{
    if (kmin < 1 || kmax < kmin) return fail(LCS_ERR_INVALID, "rc-energy needs 1 <= kmin <= kmax");
    if (kmax > c / 2) return fail(LCS_ERR_INVALID, "rc-energy needs kmax <= coarse_size/2");
    if (kmax - kmin + 1 > 8) return fail(LCS_ERR_INVALID, "at most 8 forced shells");
    if (kmax >= kMaxRcfKx) return fail(LCS_ERR_INVALID, "rc-energy needs kmax <= 7");
    return LCS_OK;
}

// the forced wavevectors of shells kmin..kmax, shell by shell in kset order
static int rcf_modes_host(int kmin, int kmax, RcfModes &md)
{
    memset(&md, 0, sizeof(md));
    md.kmin = kmin;
    for (int K = kmin; K <= kmax; K++) {
        md.kof[K - kmin] = md.n;
        int n = kset_host(K, md.k + md.n, kMaxRcfModes - md.n);
        if (n < 0) return fail(LCS_ERR_INVALID, "too many forced modes");
        md.kn[K - kmin] = n;
        md.n += n;
    }
    int s = 0;
    for (int kx = 0; kx < kMaxRcfKx; kx++) {
        md.xof[kx] = s;
        for (int m = 0; m < md.n; m++)
            if (md.k[m][0] == kx) md.byx[s++] = (unsigned char)m;
    }
    md.xof[kMaxRcfKx] = s;
    if (s != md.n) return fail(LCS_ERR_INVALID, "forced wavevector outside kx in 0..7");
    return LCS_OK;
}

int rcf_apply_blocks(const Box &b) { return grid_rows(b.g); }
int rcf_crop_slots(int kmax) { return (kmax + 1) * (2 * kmax + 1) * (2 * kmax + 1); }

int rcf_prepare(Box &b, int kmin, int kmax, int c)
{
    int rc = rcf_check(kmin, kmax, c);
    if (rc) return rc;
    const Geom &g = b.g;
    if (c < 1 || g.nx % c || g.ny % c || g.nz % c)
        return fail(LCS_ERR_INVALID, "coarse_size must divide the grid dims");
    if ((rc = rcf_buffers(b, c, kmax))) return rc;
    if (b.rcf_modes && b.rcf_key[0] == kmin && b.rcf_key[1] == kmax && b.rcf_key[2] == c) return LCS_OK;
    RcfModes md;
    if ((rc = rcf_modes_host(kmin, kmax, md))) return rc;
    if (!b.rcf_modes) {
        LCS_HIP(hipMalloc(&b.rcf_modes, sizeof(RcfModes)));
        b.bytes += sizeof(RcfModes);
    }
    LCS_HIP(hipMemcpy(b.rcf_modes, &md, sizeof(RcfModes), hipMemcpyHostToDevice));
    b.rcf_key[0] = kmin;
    b.rcf_key[1] = kmax;
    b.rcf_key[2] = c;
    return LCS_OK;
}

// ---- batched rc-energy forcing: the per-box work of launch_force_rcf with the
// box, its parameters and buffers taken from the batch's records ----
template <typename T>
__global__ void __launch_bounds__(256) k_b_boxmean(const BoxRec *tab, const WorkItem *work)
{
    __shared__ double lds[16];
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    const int c = B.rcf_c, n3 = c * c * c, comp = w.blk / n3;
    boxmean_block<T>(B.g, (const T *)B.f[comp], c, w.blk % n3, comp, B.rcf_buf, lds);
}

__global__ void __launch_bounds__(256) k_b_rcf_dft(const BoxRec *tab, const WorkItem *work)
{
    __shared__ double lds[16];
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    const int c = B.rcf_c;
    rcf_dft_block(B.rcf_buf, c, B.rcf_kmax, w.blk, B.rcf_buf + 3L * c * c * c, lds);
}

__global__ void k_b_rcf_spectrum(const BoxRec *tab, const int *boxes)
{
    const BoxRec &B = tab[boxes[blockIdx.x]];
    if (b_halted(B)) return;
    const int c = B.rcf_c;
    double *hat = B.rcf_buf + 3L * c * c * c;
    if (threadIdx.x == 0)
        rcf_spectrum_one(hat, c, B.rcf_kmin, B.rcf_kmax, B.rcf_energy, hat + 6L * rcf_crop_slots_dev(B.rcf_kmax), B.S);
}

template <typename T>
__global__ void __launch_bounds__(256) k_b_rcf_apply(const BoxRec *tab, const WorkItem *work)
{
    __shared__ double lds[16];
    __shared__ RcfRowLds rl;
    __shared__ RcfModes md;  // a copy the field stores cannot alias, off the vector-memory queue
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    static_assert(sizeof(RcfModes) % sizeof(int) == 0, "RcfModes copied as ints");
    for (int q = threadIdx.x; q < (int)(sizeof(RcfModes) / sizeof(int)); q += blockDim.x)
        ((int *)&md)[q] = ((const int *)B.rcf_modes)[q];
    __syncthreads();
    const Geom g = B.g;
    const int c = B.rcf_c, kmax = B.rcf_kmax, nblk = B.rcf_blocks;
    const double *crop = B.rcf_buf + 3L * c * c * c + 6L * rcf_crop_slots_dev(kmax);
    const double dti = 1.0 / (1.0 * B.S->dt);
    T *U = (T *)B.f[0], *V = (T *)B.f[1], *W = (T *)B.f[2];
    double *part = B.part;
    const double r = block_sum(rcf_apply_part(g, U, V, W, c, kmax, md, crop, dti, w.blk, nblk, rl), lds);
    if (threadIdx.x == 0) part[w.blk] = r;
}

// p = mean over the cells of the power partials (k_mean_to per box)
__global__ void k_b_rcf_power(const BoxRec *tab, const int *boxes)
{
    __shared__ double lds[16];
    const BoxRec &B = tab[boxes[blockIdx.x]];
    if (b_halted(B)) return;
    double s = 0.0;
    for (int q = threadIdx.x; q < B.rcf_blocks; q += blockDim.x) s += B.part[q];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) B.S->p = s / (double)B.g.cells();
}

template <typename T>
static void b_rcf_t(hipStream_t s, const BoxRec *tab, const WorkItem *mean, int nmean, const WorkItem *dft, int ndft,
                    const int *boxes, int nbox, const WorkItem *apply, int napply)
{
    hipLaunchKernelGGL(k_b_boxmean<T>, dim3(nmean), dim3(256), 0, s, tab, mean);
    hipLaunchKernelGGL(k_b_rcf_dft, dim3(ndft), dim3(256), 0, s, tab, dft);
    hipLaunchKernelGGL(k_b_rcf_spectrum, dim3(nbox), dim3(64), 0, s, tab, boxes);
    hipLaunchKernelGGL(k_b_rcf_apply<T>, dim3(napply), dim3(256), 0, s, tab, apply);
    hipLaunchKernelGGL(k_b_rcf_power, dim3(nbox), dim3(256), 0, s, tab, boxes);
}

int launch_b_rcf(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *mean, int nmean, const WorkItem *dft,
                 int ndft, const int *boxes, int nbox, const WorkItem *apply, int napply)
{
    if (nbox == 0) return LCS_OK;
    if (esize == 8)
        b_rcf_t<double>(s, tab, mean, nmean, dft, ndft, boxes, nbox, apply, napply);
    else
        b_rcf_t<float>(s, tab, mean, nmean, dft, ndft, boxes, nbox, apply, napply);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

int launch_force_rcf(Box &b, double energy, int kmin, int kmax, int c, DtRef dt, double *scratch)
{
    int rc = rcf_check(kmin, kmax, c);
    if (rc) return rc;
    if (!scratch && (rc = rcf_buffers(b, c, kmax))) return rc;
    RcfModes md;
    if ((rc = rcf_modes_host(kmin, kmax, md))) return rc;
    int nk = 2 * kmax + 1, ncrop = (kmax + 1) * nk * nk;
    double *coarse = scratch ? scratch : b.rcf;
    double *hat = coarse + 3L * c * c * c;
    double *crop = hat + 6L * ncrop;
    if ((rc = launch_boxmean(b, c, coarse))) return rc;
    hipLaunchKernelGGL(k_rcf_dft, dim3(ncrop), dim3(256), 0, b.stream, coarse, c, kmax, hat);
    LCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rcf_spectrum, dim3(1), dim3(64), 0, b.stream, hat, c, kmin, kmax, energy, crop, b.S);
    LCS_HIP(hipGetLastError());
    const int blocks = rcf_apply_blocks(b);
    ProfScope ps(b, "rcf_apply", 6.0 * b.esize * b.g.cells());
    if (b.esize == 8)
        hipLaunchKernelGGL(k_rcf_apply<double>, dim3(blocks), dim3(256), 0, b.stream, b.g, (double *)b.f[0],
                           (double *)b.f[1], (double *)b.f[2], c, kmax, md, (const double *)crop, dt, b.part);
    else
        hipLaunchKernelGGL(k_rcf_apply<float>, dim3(blocks), dim3(256), 0, b.stream, b.g, (float *)b.f[0],
                           (float *)b.f[1], (float *)b.f[2], c, kmax, md, (const double *)crop, dt, b.part);
    LCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mean_to, dim3(1), dim3(256), 0, b.stream, b.part, blocks, (double)b.g.cells(), &b.S->p);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// ---------------------------------------------------------------------------
// Spectrum initialiser (builder-defined input, mirrors the oracle's
// o_spectrum_modes/o_init_spectrum): separable evaluation of
//   F_c(x) = sum_m 2 Re(a_{m,c} e^{i k_m . x})
// on the staggered positions, in three 1-D passes (x, then y, then z).
// ---------------------------------------------------------------------------
static int spectrum_modes_host(uint64_t seed, double k0, int K, std::vector<int> &kt, std::vector<double> &at)
{
    for (int kz = -K; kz <= K; kz++)
        for (int ky = -K; ky <= K; ky++)
            for (int kx = 0; kx <= K; kx++) {
                int canonical = (kx > 0) || (kx == 0 && ky > 0) || (kx == 0 && ky == 0 && kz > 0);
                int r2 = kx * kx + ky * ky + kz * kz;
                if (!canonical || r2 == 0 || r2 > K * K) continue;
                uint64_t idx = (uint64_t)(kx + (K + 1) * ((ky + K) + (2 * K + 1) * (kz + K)));
                double a[6];
                for (int c = 0; c < 6; c++) a[c] = 2.0 * rand_u01(seed, (uint64_t)c, idx) - 1.0;
                double kk = std::sqrt((double)r2);
                double amp = kk * std::exp(-(kk / k0) * (kk / k0));
                double kv[3] = {(double)kx, (double)ky, (double)kz};
                for (int part = 0; part < 2; part++) {
                    double d = 0;
                    for (int c = 0; c < 3; c++) d += kv[c] * a[2 * c + part];
                    for (int c = 0; c < 3; c++) a[2 * c + part] = amp * (a[2 * c + part] - kv[c] * d / r2);
                }
                kt.push_back(kx);
                kt.push_back(ky);
                kt.push_back(kz);
                for (int c = 0; c < 6; c++) at.push_back(a[c]);
            }
    return (int)(kt.size() / 3);
}

// pass 1: G1[(ky,kz)][x] = sum_kx a(kx,ky,kz) e^{i kx x}; table A is dense
// over (kx in 0..K, ky, kz in -K..K) for one component (re, im).
__global__ void k_spec_x(const double *A, int K, int nx, double xoff, double dx, double *G1)
{
    int nk = 2 * K + 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)nk * nk * nx;
    if (t >= total) return;
    int x = (int)(t % nx);
    int yz = (int)(t / nx);
    double xp = ((double)(x + 1) - xoff) * dx;
    double sr = 0, si = 0;
    for (int kx = 0; kx <= K; kx++) {
        const double *a = A + 2L * (kx + (K + 1) * (long)yz);
        double sn, cs;
        sincos(kx * xp, &sn, &cs);
        sr += a[0] * cs - a[1] * sn;
        si += a[0] * sn + a[1] * cs;
    }
    G1[2 * t] = sr;
    G1[2 * t + 1] = si;
}

// pass 2: G2[kz][y][x] = sum_ky G1[(ky,kz)][x] e^{i ky y}
__global__ void k_spec_y(const double *G1, int K, int nx, int ny, double yoff, double dy, double *G2)
{
    int nk = 2 * K + 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)nk * ny * nx;
    if (t >= total) return;
    int x = (int)(t % nx);
    int y = (int)((t / nx) % ny);
    int kz = (int)(t / ((long)nx * ny));
    double yp = ((double)(y + 1) - yoff) * dy;
    double sr = 0, si = 0;
    for (int ky = -K; ky <= K; ky++) {
        const double *a = G1 + 2L * (x + (long)nx * ((ky + K) + nk * (long)kz));
        double sn, cs;
        sincos(ky * yp, &sn, &cs);
        sr += a[0] * cs - a[1] * sn;
        si += a[0] * sn + a[1] * cs;
    }
    G2[2 * t] = sr;
    G2[2 * t + 1] = si;
}

// pass 3: F(x,y,z) = sum_kz 2 Re(G2[kz][y][x] e^{i kz z}); writes the field
// interior and accumulates sum F^2 partials.
template <typename T>
__global__ void __launch_bounds__(256) k_spec_z(Geom g, const double *G2, int K, double zoff, double dz, T *F,
                                               double *part)
{
    __shared__ double lds[16];
    double ss = 0;
    int nk = 2 * K + 1;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < g.cells(); t += (long)gridDim.x * blockDim.x) {
        int x = (int)(t % g.nx);
        long q = t / g.nx;
        int y = (int)(q % g.ny);
        int z = (int)(q / g.ny);
        double zp = ((double)(z + 1) - zoff) * dz;
        double s = 0;
        for (int kz = -K; kz <= K; kz++) {
            const double *a = G2 + 2L * (x + (long)g.nx * (y + (long)g.ny * (kz + K)));
            double sn, cs;
            sincos(kz * zp, &sn, &cs);
            s += 2.0 * (a[0] * cs - a[1] * sn);
        }
        F[g.at(x + 1, y + 1, z + 1)] = T(s);
        ss += s * s;
    }
    (void)nk;
    double r = block_sum(ss, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

template <typename T>
__global__ void k_scale_interior(Geom g, T *U, T *V, T *W, double scale)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.cells()) return;
    int i = (int)(t % g.nx) + 1;
    long q = t / g.nx;
    int j = (int)(q % g.ny) + 1;
    int k = (int)(q / g.ny) + 1;
    long p = g.at(i, j, k);
    U[p] = T((double)U[p] * scale);
    V[p] = T((double)V[p] * scale);
    W[p] = T((double)W[p] * scale);
}

template <typename T>
static int init_spectrum_t(Box &b, uint64_t seed, double k0, double uprime)
{
    const Geom &g = b.g;
    int nmin = std::min(g.nx, std::min(g.ny, g.nz));
    int K = (int)std::ceil(3.0 * k0);
    if (K > nmin / 2 - 1) K = nmin / 2 - 1;
    if (K < 1) return fail(LCS_ERR_INVALID, "grid too small for the spectrum initialiser");
    std::vector<int> kt;
    std::vector<double> at;
    int nm = spectrum_modes_host(seed, k0, K, kt, at);
    int nk = 2 * K + 1;
    size_t dense = (size_t)(K + 1) * nk * nk;
    std::vector<double> A(2 * dense);
    int nmax = std::max(g.nx, std::max(g.ny, g.nz));
    size_t g1n = 2 * (size_t)nk * nk * nmax;
    size_t g2n = 2 * (size_t)nk * g.ny * g.nx;
    double *dA = nullptr, *dG1 = nullptr, *dG2 = nullptr;
    LCS_HIP(hipMalloc(&dA, sizeof(double) * A.size()));
    LCS_HIP(hipMalloc(&dG1, sizeof(double) * g1n));
    LCS_HIP(hipMalloc(&dG2, sizeof(double) * g2n));
    for (int a = 0; a < 4; a++) LCS_HIP(hipMemsetAsync(b.f[LCS_FIELD_U + (a == 3 ? LCS_FIELD_P : a)], 0, g.len * sizeof(T), b.stream));
    double d[3] = {spacing(g.nx), spacing(g.ny), spacing(g.nz)};
    int blocks = grid_for(g.cells(), 256);
    std::vector<double> sums(kMaxPartBlocks);
    double total = 0.0;
    for (int comp = 0; comp < 3; comp++) {
        std::fill(A.begin(), A.end(), 0.0);
        for (int m = 0; m < nm; m++) {
            int kx = kt[3 * m], ky = kt[3 * m + 1], kz = kt[3 * m + 2];
            size_t q = (size_t)kx + (K + 1) * ((size_t)(ky + K) + nk * (size_t)(kz + K));
            A[2 * q] = at[6 * m + 2 * comp];
            A[2 * q + 1] = at[6 * m + 2 * comp + 1];
        }
        LCS_HIP(hipMemcpyAsync(dA, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice, b.stream));
        // staggering: component c sits on the face at index*d along axis c,
        // at (index - 1/2)*d along the others
        double off[3] = {0.5, 0.5, 0.5};
        off[comp] = 0.0;
        long t1 = (long)nk * nk * g.nx;
        hipLaunchKernelGGL(k_spec_x, dim3((t1 + 255) / 256), dim3(256), 0, b.stream, dA, K, g.nx, off[0], d[0], dG1);
        long t2 = (long)nk * g.ny * g.nx;
        hipLaunchKernelGGL(k_spec_y, dim3((t2 + 255) / 256), dim3(256), 0, b.stream, dG1, K, g.nx, g.ny, off[1], d[1],
                           dG2);
        hipLaunchKernelGGL(k_spec_z<T>, dim3(blocks), dim3(256), 0, b.stream, g, dG2, K, off[2], d[2],
                           (T *)b.f[LCS_FIELD_U + comp], b.part);
        LCS_HIP(hipGetLastError());
        LCS_HIP(hipMemcpyAsync(sums.data(), b.part, sizeof(double) * blocks, hipMemcpyDeviceToHost, b.stream));
        LCS_HIP(hipStreamSynchronize(b.stream));
        for (int q = 0; q < blocks; q++) total += sums[q];
    }
    double up = std::sqrt(total / (double)g.cells() / 3.0);
    double scale = uprime / up;
    hipLaunchKernelGGL(k_scale_interior<T>, dim3((g.cells() + 255) / 256), dim3(256), 0, b.stream, g,
                       (T *)b.f[0], (T *)b.f[1], (T *)b.f[2], scale);
    LCS_HIP(hipGetLastError());
    LCS_HIP(hipStreamSynchronize(b.stream));
    LCS_HIP(hipFree(dA));
    LCS_HIP(hipFree(dG1));
    LCS_HIP(hipFree(dG2));
    return LCS_OK;
}

int launch_init_spectrum(Box &b, uint64_t seed, double k0, double uprime)
{
    return b.esize == 8 ? init_spectrum_t<double>(b, seed, k0, uprime) : init_spectrum_t<float>(b, seed, k0, uprime);
}

// ===========================================================================
// Batched kernels (lcs_batch_*): one launch covers every box of a batch.  A
// work table maps each workgroup to (box, block of that box) or (box, tile);
// pointers, sizes and coefficients come from the box's BoxRec at run time.
// The per-box arithmetic is the single-box kernels' own device code.
// ===========================================================================
// arrays of a stage set: 0 -> (U, V, W, P), 1 -> (U2, V2, W2, P)
__device__ inline void *set_array(const BoxRec &B, int set, int q) { return q < 3 ? B.f[3 * set + q] : B.f[6]; }

template <typename T>
__global__ void __launch_bounds__(256) k_b_ghost_fill(const BoxRec *tab, const WorkItem *work, int set, int narr,
                                                     int r, int full, int hsmac)
{
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    if (hsmac && B.S->hs_done) return;
    const bool alt = hsmac && B.S->hs_par;
    const Geom g = B.g;
    long cz, cy, cx;
    ghost_counts(g, r, full, cz, cy, cx);
    const long t = (long)w.blk * blockDim.x + threadIdx.x;
    if (t >= cz + cy + cx) return;
    long d, s;
    ghost_cell(g, r, full, (int)cz, (int)cy, (int)t, d, s);
    for (int q = 0; q < narr; q++) {
        T *A = (T *)(alt ? B.spare[q] : set_array(B, set, q));
        A[d] = A[s];
    }
}

int launch_b_ghost_fill(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *work, int nwork, int set,
                        int narr, int r, int full, int hsmac)
{
    if (nwork == 0) return LCS_OK;
    if (esize == 8)
        hipLaunchKernelGGL(k_b_ghost_fill<double>, dim3(nwork), dim3(256), 0, s, tab, work, set, narr, r, full, hsmac);
    else
        hipLaunchKernelGGL(k_b_ghost_fill<float>, dim3(nwork), dim3(256), 0, s, tab, work, set, narr, r, full, hsmac);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// CFL dt per box: rows work item blk of B.row_blocks, then one block per box
template <typename T>
__global__ void __launch_bounds__(256) k_b_cfl(const BoxRec *tab, const WorkItem *work)
{
    __shared__ double lds[16];
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    double m[3];
    cfl_block(B.g, (const T *)B.f[0], (const T *)B.f[1], (const T *)B.f[2], w.blk, B.row_blocks, m);
    for (int a = 0; a < 3; a++) {
        double v = block_max(m[a], lds);
        if (threadIdx.x == 0) B.part[(long)a * kMaxPartBlocks + w.blk] = v;
    }
}

__global__ void k_b_cfl_fin(const BoxRec *tab, double cfl, double dtmax)
{
    __shared__ double lds[16];
    const BoxRec &B = tab[blockIdx.x];
    if (b_halted(B)) return;
    double mm[3];
    for (int a = 0; a < 3; a++) {
        double m = 0.0;
        for (int q = threadIdx.x; q < B.row_blocks; q += blockDim.x) m = fmax(m, B.part[(long)a * kMaxPartBlocks + q]);
        mm[a] = block_max(m, lds);
    }
    if (threadIdx.x == 0) cfl_dt(B.S, mm, cfl, dtmax, B.dx, B.dy, B.dz, B.Re);
}

int launch_b_cfl(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *rows, int nrows, int nbox, double cfl,
                 double dtmax)
{
    if (esize == 8)
        hipLaunchKernelGGL(k_b_cfl<double>, dim3(nrows), dim3(256), 0, s, tab, rows);
    else
        hipLaunchKernelGGL(k_b_cfl<float>, dim3(nrows), dim3(256), 0, s, tab, rows);
    hipLaunchKernelGGL(k_b_cfl_fin, dim3(nbox), dim3(256), 0, s, tab, cfl, dtmax);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// linear-power forcing of the boxes listed in `boxes` (their rows in `rows`)
template <typename T>
__global__ void __launch_bounds__(256) k_b_lf_means(const BoxRec *tab, const WorkItem *work)
{
    __shared__ double lds[16];
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    double s4[4];
    lf_means_block(B.g, (const T *)B.f[0], (const T *)B.f[1], (const T *)B.f[2], w.blk, B.row_blocks, s4);
    for (int a = 0; a < 4; a++) {
        double v = block_sum(s4[a], lds);
        if (threadIdx.x == 0) B.part[(long)a * kMaxPartBlocks + w.blk] = v;
    }
}

__global__ void k_b_lf_fin(const BoxRec *tab, const int *boxes)
{
    __shared__ double lds[16];
    const BoxRec &B = tab[boxes[blockIdx.x]];
    if (b_halted(B)) return;
    double m[4];
    for (int a = 0; a < 4; a++) {
        double v = 0.0;
        for (int q = threadIdx.x; q < B.row_blocks; q += blockDim.x) v += B.part[(long)a * kMaxPartBlocks + q];
        m[a] = block_sum(v, lds);
    }
    if (threadIdx.x == 0) lf_coef(B.S, m, (double)B.g.cells(), B.power);
}

template <typename T>
__global__ void __launch_bounds__(256) k_b_lf_apply(const BoxRec *tab, const WorkItem *work)
{
    __shared__ double lds[16];
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    const T rdt = T(B.S->dt);  // stage-2 dt
    double v = block_sum(lf_apply_block(B.g, (T *)B.f[0], (T *)B.f[1], (T *)B.f[2], B.S, rdt, w.blk, B.row_blocks),
                         lds);
    if (threadIdx.x == 0) B.part[w.blk] = v;
}

// state.p of the boxes whose first stage-2 sweep applied the forcing: the
// mean of the per-tile power sums (lf.jl:103)
__global__ void k_b_lf_power(const BoxRec *tab, const int *boxes)
{
    __shared__ double lds[16];
    const BoxRec &B = tab[boxes[blockIdx.x]];
    if (b_halted(B)) return;
    double v = 0.0;
    for (int q = threadIdx.x; q < B.hs_tiles; q += blockDim.x) v += B.part[kPowerSlot + q];
    v = block_sum(v, lds);
    if (threadIdx.x == 0) B.S->p = v / (double)B.g.cells();
}

int launch_b_lf_power(hipStream_t s, const BoxRec *tab, const int *boxes, int nbox)
{
    if (nbox == 0) return LCS_OK;
    hipLaunchKernelGGL(k_b_lf_power, dim3(nbox), dim3(256), 0, s, tab, boxes);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

__global__ void k_b_lf_mean(const BoxRec *tab, const int *boxes)
{
    __shared__ double lds[16];
    const BoxRec &B = tab[boxes[blockIdx.x]];
    if (b_halted(B)) return;
    double v = 0.0;
    for (int q = threadIdx.x; q < B.row_blocks; q += blockDim.x) v += B.part[q];
    v = block_sum(v, lds);
    if (threadIdx.x == 0) B.S->p = v / (double)B.g.cells();
}

int launch_b_lf(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *rows, int nrows, const int *boxes,
                int nbox, bool apply)
{
    if (nbox == 0) return LCS_OK;
    if (esize == 8)
        hipLaunchKernelGGL(k_b_lf_means<double>, dim3(nrows), dim3(256), 0, s, tab, rows);
    else
        hipLaunchKernelGGL(k_b_lf_means<float>, dim3(nrows), dim3(256), 0, s, tab, rows);
    hipLaunchKernelGGL(k_b_lf_fin, dim3(nbox), dim3(256), 0, s, tab, boxes);
    LCS_HIP(hipGetLastError());
    if (!apply) return LCS_OK;
    if (esize == 8)
        hipLaunchKernelGGL(k_b_lf_apply<double>, dim3(nrows), dim3(256), 0, s, tab, rows);
    else
        hipLaunchKernelGGL(k_b_lf_apply<float>, dim3(nrows), dim3(256), 0, s, tab, rows);
    hipLaunchKernelGGL(k_b_lf_mean, dim3(nbox), dim3(256), 0, s, tab, boxes);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// fused RHS + update of one stage for every box
template <typename T, int TX, int TY>
__global__ void __launch_bounds__(TX *TY) k_b_rhs_tile(const BoxRec *tab, const TileItem *tiles, int stage,
                                                        int spare_out)
{
    extern __shared__ __align__(16) unsigned char rhs_lds_raw[];
    const TileItem t = tiles[blockIdx.x];
    const BoxRec &B = tab[t.box];
    if (b_halted(B)) return;
    RhsArgs<T> a;
    a.g = B.g;
    const int in = stage == 1 ? 0 : 3;
    a.U = (const T *)B.f[in];
    a.V = (const T *)B.f[in + 1];
    a.W = (const T *)B.f[in + 2];
    a.P = (const T *)B.f[6];
    for (int q = 0; q < 3; q++) {
        a.out[q] = spare_out ? (T *)B.spare[q] : (T *)B.f[(stage == 1 ? 3 : 0) + q];
        a.base[q] = stage == 1 ? nullptr : (const T *)B.f[q];  // stage 1: base == input
        a.di[q] = T(B.rhs_di[q]);
        a.di2[q] = T(B.rhs_di2[q]);
    }
    a.Rei = T(B.rhs_Rei);
    a.zc = B.rhs_zc;
    const T rdt = T((stage == 1 ? 0.5 : 1.0) * B.S->dt);
    rhs_tile_block<T, TX, TY, true>(a, t.bx, t.by, t.bz, rdt, reinterpret_cast<T *>(rhs_lds_raw));
}

template <typename T, int TX, int TY>
static int b_rhs_t(hipStream_t s, const BoxRec *tab, const TileItem *tiles, int ntiles, int stage, int spare_out)
{
    using Tile = RhsTile<T, TX, TY>;
    const size_t lds = Tile::LDS_ELEMS * sizeof(T);
    static_assert(Tile::LDS_ELEMS * sizeof(T) <= 160 * 1024, "RHS tile exceeds the LDS of a CU");
    int rc = kernel_setup((const void *)k_b_rhs_tile<T, TX, TY>, TX * TY, lds, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL((k_b_rhs_tile<T, TX, TY>), dim3(ntiles), dim3(TX * TY), lds, s, tab, tiles, stage, spare_out);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

int launch_b_rhs(hipStream_t s, int esize, int tx, int ty, const BoxRec *tab, const TileItem *tiles, int ntiles,
                 int stage, bool spare_out)
{
    if (ntiles == 0) return LCS_OK;
    const int so = spare_out ? 1 : 0;
    if (esize == 8) {
        if (tx == 32 && ty == 4) return b_rhs_t<double, 32, 4>(s, tab, tiles, ntiles, stage, so);
        if (tx == 32 && ty == 8) return b_rhs_t<double, 32, 8>(s, tab, tiles, ntiles, stage, so);
        if (tx == 64 && ty == 8) return b_rhs_t<double, 64, 8>(s, tab, tiles, ntiles, stage, so);
    } else {
        if (tx == 32 && ty == 8) return b_rhs_t<float, 32, 8>(s, tab, tiles, ntiles, stage, so);
        if (tx == 64 && ty == 8) return b_rhs_t<float, 64, 8>(s, tab, tiles, ntiles, stage, so);
        if (tx == 64 && ty == 16) return b_rhs_t<float, 64, 16>(s, tab, tiles, ntiles, stage, so);
        if (tx == 32 && ty == 16) return b_rhs_t<float, 32, 16>(s, tab, tiles, ntiles, stage, so);
    }
    return fail(LCS_ERR_INVALID, "unsupported batched RHS tile shape");
}

template <typename T>
__device__ inline void hs_coef(const BoxRec &B, HsCoef<T> &c)
{
    c.dxi = T(B.hs_dxi);
    c.dyi = T(B.hs_dyi);
    c.dzi = T(B.hs_dzi);
    c.dl = T(B.hs_dl);
    c.omega = T(B.hs_omega);
}

// one tile of the batched sweep (a no-op for a halted or converged box)
template <typename T, int COLORING, int TY, int TX, bool FORCE>
__device__ __forceinline__ void b_sweep_tile(const BoxRec *tab, const TileItem &t, int set, int kind, int slot, double dtfac,
                                    int wrap, T *lds, double *red_lds)
{
    const BoxRec &B = tab[t.box];
    if (b_halted(B)) return;
    if (B.S->hs_done) return;
    BufSets<T> bs;
    for (int q = 0; q < 4; q++) {
        bs.a[q] = (T *)set_array(B, set, q);
        bs.b[q] = (T *)B.spare[q];
    }
    HsArgs<T> a;
    a.g = B.g;
    hs_bind(a, bs, B.S->hs_par != 0);
    a.R = hs_region(B.g, kind);
    hs_coef(B, a.c);
    a.dti = T(1.0 / (dtfac * B.S->dt));
    a.zc = B.hs_zc;
    a.wrap = wrap;
    if (FORCE && B.lf_force) {  // the coefficients k_b_lf_apply uses (lf_apply_block)
        a.xf = 1;
        a.xA = T(B.S->lf[4]);
        for (int q = 0; q < 3; q++) a.xm[q] = T(B.S->lf[1 + q]);
        a.xdt = T(dtfac * B.S->dt);
    }
    double pw = 0.0;
    const double s = hsmac_sweep_block<T, COLORING, TY, TX, FORCE>(a, t.bx, t.by, t.bz, lds, red_lds, &pw);
    const int gx = (B.g.nx + 2 * B.g.h + TX - 1) / TX, gy = (B.g.ny + 2 * B.g.h + TY - 1) / TY;
    const long tix = t.bx + gx * (t.by + gy * t.bz);
    if (threadIdx.x == 0) {
        B.part[(long)slot * kMaxPartBlocks + tix] = s;
        if (FORCE && B.lf_force) B.part[kPowerSlot + tix] = pw;
    }
}

// One workgroup per tile of the list.
template <typename T, int COLORING, int TY, int TX, bool FORCE = false>
__global__ void __launch_bounds__((HsTile<T, COLORING, TY, TX>::NT))
    k_b_hsmac_sweep(const BoxRec *tab, const TileItem *tiles, int set, int kind, int slot, double dtfac, int wrap)
{
    extern __shared__ __align__(16) unsigned char hs_lds_raw[];
    __shared__ double red_lds[16];
    b_sweep_tile<T, COLORING, TY, TX, FORCE>(tab, tiles[blockIdx.x], set, kind, slot, dtfac, wrap,
                                             reinterpret_cast<T *>(hs_lds_raw), red_lds);
}

// The same sweep as a persistent grid that walks the tile list.  queue =
// null: workgroup b takes tiles b, b + G, b + 2G, ... (G = the grid).
// Otherwise (G a multiple of 8) tiles come from eight queues, one per XCD:
// the list puts item L on XCD L % 8, and workgroup b, which the hardware runs
// on XCD b % 8, takes items x, x + 8, x + 16, ... (x = b % 8) in the order
// its queue hands them out.  queue[0..7] count the items handed out,
// queue[8] the workgroups done; the last one done sets all nine back to zero
// for the next launch.  A tile's result does not depend on the workgroup that
// runs it.
template <typename T, int COLORING, int TY, int TX, bool FORCE = false>
__global__ void __launch_bounds__((HsTile<T, COLORING, TY, TX>::NT))
    k_b_hsmac_sweep_walk(const BoxRec *tab, const TileItem *tiles, int ntiles, int set, int kind, int slot,
                         double dtfac, int wrap, int *queue)
{
    extern __shared__ __align__(16) unsigned char hs_lds_raw[];
    __shared__ double red_lds[16];
    __shared__ int item;
    const int x = blockIdx.x % 8;
    for (int k = 0;; k++) {
        if (k) __syncthreads();  // the previous tile is done with the LDS, and every thread has read `item`
        if (threadIdx.x == 0) item = queue ? x + 8 * atomicAdd(&queue[x], 1) : (int)(blockIdx.x + k * gridDim.x);
        __syncthreads();
        const int i = item;
        if (i >= ntiles) break;  // uniform over the workgroup
        b_sweep_tile<T, COLORING, TY, TX, FORCE>(tab, tiles[i], set, kind, slot, dtfac, wrap,
                                                 reinterpret_cast<T *>(hs_lds_raw), red_lds);
    }
    if (queue && threadIdx.x == 0 && atomicAdd(&queue[8], 1) == (int)gridDim.x - 1)
        for (int q = 0; q < 9; q++) atomicExch(&queue[q], 0);
}

template <typename T, int COLORING, int TX, int TY, bool FORCE>
static int b_sweep_f(hipStream_t s, const BoxRec *tab, const TileItem *tiles, int ntiles, int set, int kind, int slot,
                     double dtfac, int wrap, int cus, int *queue)
{
    using Tile = HsTile<T, COLORING, TY, TX, FORCE>;
    const size_t lds = Tile::LDS_ELEMS * sizeof(T);
    if (cus > 0) {
        int per_cu = 1;
        const void *k = (const void *)k_b_hsmac_sweep_walk<T, COLORING, TY, TX, FORCE>;
        int rc = kernel_setup(k, Tile::NT, lds, &per_cu);
        if (rc) return rc;
        long grid = (long)cus * per_cu;
        if (grid < ntiles) {
            int *q = nullptr;
            if (queue && grid >= 8) {  // per-XCD queues need every XCD's workgroups present
                grid -= grid % 8;
                q = queue;
            }
            hipLaunchKernelGGL((k_b_hsmac_sweep_walk<T, COLORING, TY, TX, FORCE>), dim3((int)grid), dim3(Tile::NT),
                               lds, s, tab, tiles, ntiles, set, kind, slot, dtfac, wrap, q);
            LCS_HIP(hipGetLastError());
            return LCS_OK;
        }
    }
    int rc = kernel_setup((const void *)k_b_hsmac_sweep<T, COLORING, TY, TX, FORCE>, Tile::NT, lds, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL((k_b_hsmac_sweep<T, COLORING, TY, TX, FORCE>), dim3(ntiles), dim3(Tile::NT), lds, s, tab,
                       tiles, set, kind, slot, dtfac, wrap);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

template <typename T, int COLORING, int TX, int TY>
static int b_sweep_t(hipStream_t s, const BoxRec *tab, const TileItem *tiles, int ntiles, int set, int kind, int slot,
                     double dtfac, bool force, int wrap, int cus, int *queue)
{
    return force ? b_sweep_f<T, COLORING, TX, TY, true>(s, tab, tiles, ntiles, set, kind, slot, dtfac, wrap, cus, queue)
                 : b_sweep_f<T, COLORING, TX, TY, false>(s, tab, tiles, ntiles, set, kind, slot, dtfac, wrap, cus,
                                                          queue);
}

template <typename T, int TX, int TY>
static int b_sweep_c(hipStream_t s, int coloring, const BoxRec *tab, const TileItem *tiles, int ntiles, int set,
                     int kind, int slot, double dtfac, bool force, int wrap, int cus, int *queue)
{
    return coloring == LCS_COLORING_RED_BLACK
               ? b_sweep_t<T, LCS_COLORING_RED_BLACK, TX, TY>(s, tab, tiles, ntiles, set, kind, slot, dtfac, force,
                                                              wrap, cus, queue)
               : b_sweep_t<T, LCS_COLORING_RED_BLACK_FAST, TX, TY>(s, tab, tiles, ntiles, set, kind, slot, dtfac,
                                                                    force, wrap, cus, queue);
}

template <typename T>
static int b_sweep_shape(hipStream_t s, int coloring, int tx, int ty, const BoxRec *tab, const TileItem *tiles,
                         int ntiles, int set, int kind, int slot, double dtfac, bool force, int wrap, int cus,
                         int *queue)
{
#define B_SHAPE(X, Y) \
    if (tx == X && ty == Y)                                                                                         \
        return b_sweep_c<T, X, Y>(s, coloring, tab, tiles, ntiles, set, kind, slot, dtfac, force, wrap, cus, queue)
    B_SHAPE(64, 8);
    B_SHAPE(32, 8);
    B_SHAPE(64, 16);
    B_SHAPE(32, 16);
#undef B_SHAPE
    return fail(LCS_ERR_INVALID, "unsupported sweep tile shape");
}

int launch_b_hsmac_sweep(hipStream_t s, int esize, int coloring, int tx, int ty, const BoxRec *tab,
                         const TileItem *tiles, int ntiles, int set, int kind, int slot, double dtfac, bool force,
                         bool wrap, int persist_cus, int *queue)
{
    if (ntiles == 0) return LCS_OK;
    const int w = wrap ? 1 : 0, c = persist_cus;
    return esize == 8 ? b_sweep_shape<double>(s, coloring, tx, ty, tab, tiles, ntiles, set, kind, slot, dtfac, force,
                                              w, c, queue)
                      : b_sweep_shape<float>(s, coloring, tx, ty, tab, tiles, ntiles, set, kind, slot, dtfac, force,
                                             w, c, queue);
}

// one colour over both face slabs of one axis, in place on the current set
template <typename T, int COLORING>
__global__ void __launch_bounds__(256) k_b_hsmac_faces(const BoxRec *tab, const WorkItem *work, int set, int domain,
                                                      int axis, int color, int slot, double dtfac)
{
    __shared__ double lds[16];
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    const DevScalars *S = B.S;
    if (S->hs_done) return;
    const bool alt = S->hs_par != 0;
    T *U = (T *)(alt ? B.spare[0] : set_array(B, set, 0)), *V = (T *)(alt ? B.spare[1] : set_array(B, set, 1));
    T *W = (T *)(alt ? B.spare[2] : set_array(B, set, 2)), *P = (T *)(alt ? B.spare[3] : set_array(B, set, 3));
    HsCoef<T> c;
    hs_coef(B, c);
    const T dti = T(1.0 / (dtfac * S->dt));
    const Geom g = B.g;
    const long sy = g.sx, sz = g.sx * g.sy;
    Box3 r0, r1;
    hs_faces(g, domain, axis, r0, r1);
    const int n0 = pass_count(r0), n1 = pass_count(r1);
    double acc = 0.0;
    const int t = w.blk * blockDim.x + threadIdx.x;
    if (t < n0 + n1) {
        const bool second = t >= n0;
        const Box3 r = second ? r1 : r0;
        const int u = second ? t - n0 : t;
        const int a0 = r.lo[0];
        const int hw = (r.hi[0] - r.lo[0] + 2) / 2;
        const int ny = r.hi[1] - r.lo[1] + 1;
        const int ih = u % hw, q = u / hw;
        const int j = r.lo[1] + q % ny, k = r.lo[2] + q / ny;
        const int i = a0 + 2 * ih + ((j + k - a0 + color) & 1);  // foraxes.jl:100-105
        if (i <= r.hi[0] && visited<COLORING>(r, i, j, k)) {
            const long p = g.at(i, j, k);
            T dudx = c.dxi * (-U[p - 1] + U[p]);
            T dvdy = c.dyi * (-V[p - sy] + V[p]);
            T dwdz = c.dzi * (-W[p - sz] + W[p]);
            T div = (dudx + dvdy) + dwdz;
            T dp = (-c.omega * div) * c.dl;
            P[p] += dp * dti;
            U[p - 1] -= dp * c.dxi;
            U[p] += dp * c.dxi;
            V[p - sy] -= dp * c.dyi;
            V[p] += dp * c.dyi;
            W[p - sz] -= dp * c.dzi;
            W[p] += dp * c.dzi;
            if (in_d3(g, i, j, k)) acc += (double)(div * div);
        }
    }
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) B.part[(long)slot * kMaxPartBlocks + w.blk - 0] = s;
}

int launch_b_hsmac_faces(hipStream_t s, int esize, int coloring, const BoxRec *tab, const WorkItem *work, int nwork,
                         int set, int domain, int axis, int color, int slot, double dtfac)
{
    if (nwork == 0) return LCS_OK;
#define B_FACES(T, C)                                                                                                \
    hipLaunchKernelGGL((k_b_hsmac_faces<T, C>), dim3(nwork), dim3(256), 0, s, tab, work, set, domain, axis, color, \
                       slot, dtfac)
    if (esize == 8) {
        if (coloring == LCS_COLORING_RED_BLACK)
            B_FACES(double, LCS_COLORING_RED_BLACK);
        else
            B_FACES(double, LCS_COLORING_RED_BLACK_FAST);
    } else {
        if (coloring == LCS_COLORING_RED_BLACK)
            B_FACES(float, LCS_COLORING_RED_BLACK);
        else
            B_FACES(float, LCS_COLORING_RED_BLACK_FAST);
    }
#undef B_FACES
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// both colours over the x face slabs of one box per work item (one patch)
template <typename T, int COLORING>
__global__ void __launch_bounds__(XfPatch<T>::N *XfPatch<T>::N)
    k_b_hsmac_xface(const BoxRec *tab, const WorkItem *work, int set, int domain, int slot, double dtfac)
{
    using X = XFaceOf<T, COLORING>;
    extern __shared__ __align__(16) unsigned char xf_lds_raw[];
    T *lds = reinterpret_cast<T *>(xf_lds_raw);
    __shared__ double red_lds[16];
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    const DevScalars *S = B.S;
    if (S->hs_done) return;
    const bool alt = S->hs_par != 0;
    XfArgs<T> a;
    a.g = B.g;
    a.U = (T *)(alt ? B.spare[0] : set_array(B, set, 0));
    a.V = (T *)(alt ? B.spare[1] : set_array(B, set, 1));
    a.W = (T *)(alt ? B.spare[2] : set_array(B, set, 2));
    a.P = (T *)(alt ? B.spare[3] : set_array(B, set, 3));
    hs_faces(a.g, domain, 2, a.R[0], a.R[1]);
    hs_coef(B, a.c);
    a.dti = T(1.0 / (dtfac * S->dt));
    a.buf = (T *)B.xf;
    typename X::Patch p;
    X::patch(a, w.blk, p);
    double acc = 0.0;
    X::load(a, p, lds, threadIdx.x);
    __syncthreads();
    X::red(a, p, lds, threadIdx.x, acc);
    __syncthreads();
    X::black(a, p, lds, threadIdx.x, acc);
    __syncthreads();
    X::finals(a, p, lds, threadIdx.x);
    const double s = block_sum(acc, red_lds);
    if (threadIdx.x == 0) B.part[(long)slot * kMaxPartBlocks + w.blk] = s;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_b_hsmac_xface_commit(const BoxRec *tab, const WorkItem *work, int set, int domain)
{
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    const DevScalars *S = B.S;
    if (S->hs_done) return;
    const bool alt = S->hs_par != 0;
    T *U = (T *)(alt ? B.spare[0] : set_array(B, set, 0)), *V = (T *)(alt ? B.spare[1] : set_array(B, set, 1));
    T *W = (T *)(alt ? B.spare[2] : set_array(B, set, 2)), *P = (T *)(alt ? B.spare[3] : set_array(B, set, 3));
    Box3 r0, r1;
    hs_faces(B.g, domain, 2, r0, r1);
    const long r = (long)w.blk * blockDim.x + threadIdx.x;
    if (r < SlabAxes<0>::rows(r0))
        XFaceOf<T, LCS_COLORING_RED_BLACK>::commit_row(B.g, r0, r1, U, V, W, P, (const T *)B.xf, r);
}

template <typename T, int COLORING>
static int b_xface_t(hipStream_t s, const BoxRec *tab, const WorkItem *work, int nwork, int set, int domain, int slot,
                     double dtfac)
{
    using X = XFaceOf<T, COLORING>;
    const size_t lds = X::LDS_ELEMS * sizeof(T);
    int rc = kernel_setup((const void *)k_b_hsmac_xface<T, COLORING>, X::NT, lds, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL((k_b_hsmac_xface<T, COLORING>), dim3(nwork), dim3(X::NT), lds, s, tab, work, set, domain, slot,
                       dtfac);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

int launch_b_hsmac_xface(hipStream_t s, int esize, int coloring, const BoxRec *tab, const WorkItem *work, int nwork,
                         int set, int domain, int slot, double dtfac)
{
    if (nwork == 0) return LCS_OK;
    const bool rb = coloring == LCS_COLORING_RED_BLACK;
    if (esize == 8)
        return rb ? b_xface_t<double, LCS_COLORING_RED_BLACK>(s, tab, work, nwork, set, domain, slot, dtfac)
                  : b_xface_t<double, LCS_COLORING_RED_BLACK_FAST>(s, tab, work, nwork, set, domain, slot, dtfac);
    return rb ? b_xface_t<float, LCS_COLORING_RED_BLACK>(s, tab, work, nwork, set, domain, slot, dtfac)
              : b_xface_t<float, LCS_COLORING_RED_BLACK_FAST>(s, tab, work, nwork, set, domain, slot, dtfac);
}

int launch_b_hsmac_xface_commit(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *work, int nwork, int set,
                                int domain)
{
    if (nwork == 0) return LCS_OK;
    if (esize == 8)
        hipLaunchKernelGGL(k_b_hsmac_xface_commit<double>, dim3(nwork), dim3(256), 0, s, tab, work, set, domain);
    else
        hipLaunchKernelGGL(k_b_hsmac_xface_commit<float>, dim3(nwork), dim3(256), 0, s, tab, work, set, domain);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// RMS-div test per box over the passes since the last check (hsmac.jl:118-128)
struct BCheckKinds {
    int k[kMaxSweepPasses];
};

__global__ void k_b_hsmac_check(const BoxRec *tab, BCheckKinds kinds, int npass, double epsp, int itrmin)
{
    __shared__ double lds[16];
    const BoxRec &B = tab[blockIdx.x];
    if (b_halted(B)) return;
    DevScalars *S = B.S;
    if (S->hs_done) return;
    double acc = 0.0;
    for (int q = 0; q < npass; q++) {
        const int kd = kinds.k[q];
        const int cnt = kd == 0 ? B.hs_tiles : B.face_blocks[kd - 1];
        const double *pp = B.part + (long)q * kMaxPartBlocks;
        for (int m = threadIdx.x; m < cnt; m += blockDim.x) acc += pp[m];
    }
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) {
        const Geom &g = B.g;
        const double len3 = (double)(g.nx + 1) * (double)(g.ny + 1) * (double)(g.nz + 1);
        const double dxmin = fmin(fmin(B.dx, B.dy), B.dz);
        const double div = sqrt(s / len3);
        const int itr = S->hs_itr;
        S->hs_div = div;
        if (div * dxmin < epsp && itrmin <= itr)
            S->hs_done = 1;
        else
            S->hs_itr = itr + 1;
    }
}

int launch_b_hsmac_check(hipStream_t s, const BoxRec *tab, int nbox, const int *kinds, int npass, double epsp,
                         int itrmin)
{
    if (npass > kMaxSweepPasses) return fail(LCS_ERR_INVALID, "too many passes per check");
    BCheckKinds k = {};
    for (int q = 0; q < npass; q++) k.k[q] = kinds[q];
    hipLaunchKernelGGL(k_b_hsmac_check, dim3(nbox), dim3(256), 0, s, tab, k, npass, epsp, itrmin);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// begin / flip / end(slot) / commit(set) / poll / gate(stage) / unhalt, one
// thread per box
__global__ void k_b_hsmac_ctl(BoxRec *tab, int nbox, int op, int arg, HsPoll *poll)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    BoxRec &B = tab[b];
    if (op == kHsUnhalt) {
        if (b == 0) *B.halt = 0;
        return;
    }
    if (b_halted(B)) return;
    DevScalars *S = B.S;
    if (op == kHsGate) {  // every thread that stores, stores the same value
        if (!S->hs_done) *B.halt = arg;
        return;
    }
    if (op == kHsBegin) {
        S->hs_done = 0;
        S->hs_itr = 0;
        S->hs_div = 0.0;
        S->hs_par = 0;
    } else if (op == kHsFlip) {
        if (!S->hs_done) S->hs_par ^= 1;
    } else if (op == kHsEnd) {
        S->itr[arg] = S->hs_itr;
        S->div[arg] = S->hs_div;
    } else if (op == kHsCommit) {
        if (S->hs_par) {  // the result lives in the spares: swap them in
            for (int q = 0; q < 4; q++) {
                const int f = q < 3 ? 3 * arg + q : 6;
                void *t = B.f[f];
                B.f[f] = B.spare[q];
                B.spare[q] = t;
            }
        }
    } else if (op == kHsPoll) {
        poll[b].done = S->hs_done;
        poll[b].par = S->hs_par;
        poll[b].itr = S->hs_itr;
        poll[b].div = S->hs_div;
    }
}

int launch_b_hsmac_ctl(hipStream_t s, BoxRec *tab, int nbox, int op, int arg, HsPoll *poll)
{
    hipLaunchKernelGGL(k_b_hsmac_ctl, dim3((nbox + 63) / 64), dim3(64), 0, s, tab, nbox, op, arg, poll);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// evolve!(state) per box, and a copy of every box's scalars for the logs
__global__ void k_b_state(const BoxRec *tab, int nbox, DevScalars *gather)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    if (b_halted(tab[b])) return;
    DevScalars *S = tab[b].S;
    S->step += 1;
    S->t += S->dt;
    if (gather) gather[b] = *S;
}

int launch_b_state(hipStream_t s, const BoxRec *tab, int nbox, DevScalars *gather)
{
    hipLaunchKernelGGL(k_b_state, dim3((nbox + 63) / 64), dim3(64), 0, s, tab, nbox, gather);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

}  // namespace lcs
