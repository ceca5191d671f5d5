// Particle kernels of the MI355X backend: periodic wrap, trilinear
// interpolation + Stokes/Schiller-Naumann drag, RK update (separate phases and
// a fused per-stage kernel), particle dt, and cell re-binning with a
// hand-written stable LSD radix sort (wave64 ballot ranking).
#include <algorithm>
#include <cmath>

#include "lcs_internal.h"

namespace lcs {

static int pgrid(long n, int block = 256)
{
    long g = (n + block - 1) / block;
    return (int)std::max(1L, g);
}

// ---------------------------------------------------------------------------
// apply_periodic_boundary!: src/particles/boundary.jl:81-120
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void wrap_axis(T &x, T &x2)
{
    const T L = T(kDomainLength), O = T(kDomainOrigin), E = T(kDomainOrigin + kDomainLength);
    T x0 = x, s = x2;
    if (x0 < O) {
        x = x0 + L;
        x2 = s + L;
    }
    if (x0 >= E) {
        x = x0 - L;
        x2 = s - L;
    }
}

template <typename T>
__global__ void k_p_boundary(long n, T *xs, T *ys, T *zs, T *xs2, T *ys2, T *zs2)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    T x = xs[p], y = ys[p], z = zs[p], x2 = xs2[p], y2 = ys2[p], z2 = zs2[p];
    wrap_axis(x, x2);
    wrap_axis(y, y2);
    wrap_axis(z, z2);
    xs[p] = x;
    ys[p] = y;
    zs[p] = z;
    xs2[p] = x2;
    ys2[p] = y2;
    zs2[p] = z2;
}

// ---------------------------------------------------------------------------
// Interpolation (src/particles/interp.jl:62-126)
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T trilinear(const Geom &g, const T *__restrict__ A, int i, int j, int k, T r1, T r2, T r3)
{
    long p = g.at(i, j, k);
    const long sy = g.sx, sz = g.sx * g.sy;
    T c0 = A[p], c1 = A[p + 1], c2 = A[p + 1 + sy], c3 = A[p + sy];
    T c4 = A[p + sz], c5 = A[p + 1 + sz], c6 = A[p + 1 + sy + sz], c7 = A[p + sy + sz];
    const T o = T(1);
    T s = (((o - r1) * (o - r2)) * (o - r3)) * c0;
    s = s + ((r1 * (o - r2)) * (o - r3)) * c1;
    s = s + ((r1 * r2) * (o - r3)) * c2;
    s = s + (((o - r1) * r2) * (o - r3)) * c3;
    s = s + (((o - r1) * (o - r2)) * r3) * c4;
    s = s + ((r1 * (o - r2)) * r3) * c5;
    s = s + ((r1 * r2) * r3) * c6;
    s = s + (((o - r1) * r2) * r3) * c7;
    return s;
}

// unsafe_floor (src/utils/fast.jl:49-52): Int32(floor(x))
template <typename T>
__device__ inline int ufloor(T x)
{
    return (int)floor(x);
}

template <typename T>
__device__ inline T interp_axis(const Geom &g, const T *__restrict__ A, int axis, T x, T y, T z, T dx, T dy, T dz)
{
    const T h5 = T(0.5);
    // IEEE divisions (never reciprocal multiplies): cell indices must be bit-exact
    T qx = x / dx, qy = y / dy, qz = z / dz;
    int i, j, k;
    T r1, r2, r3;
    if (axis == 0) {  // xmapping
        i = ufloor(qx);
        j = ufloor(qy + h5);
        k = ufloor(qz + h5);
        r1 = qx - T(i);
        r2 = (qy - T(j)) + h5;
        r3 = (qz - T(k)) + h5;
    } else if (axis == 1) {  // ymapping
        j = ufloor(qy);
        k = ufloor(qz + h5);
        i = ufloor(qx + h5);
        r2 = qy - T(j);
        r3 = (qz - T(k)) + h5;
        r1 = (qx - T(i)) + h5;
    } else {  // zmapping
        k = ufloor(qz);
        i = ufloor(qx + h5);
        j = ufloor(qy + h5);
        r3 = qz - T(k);
        r2 = (qy - T(j)) + h5;
        r1 = (qx - T(i)) + h5;
    }
    // A position outside the ghost-covered range (the reference leaves this
    // undefined) yields NaN instead of an out-of-bounds load.
    if (i < 1 - g.h || i > g.nx + g.h - 2 || j < 1 - g.h || j > g.ny + g.h - 2 || k < 1 - g.h ||
        k > g.nz + g.h - 2)
        return T(NAN);
    return trilinear(g, A, i, j, k, r1, r2, r3);
}

template <typename T>
struct MotionCoef {
    T dx, dy, dz, Re, g, kap;  // taup = kap * d^2 (src/particles/config.jl:41)
    int drag;
};

// motion! per particle (src/particles/motion.jl:62-81)
template <typename T>
__device__ inline void motion_one(const Geom &g, const T *U, const T *V, const T *W, const MotionCoef<T> &c, T x,
                                  T y, T z, T up, T vp, T wp, T diam, T &du, T &dv, T &dw)
{
    x = x - T(kDomainOrigin);
    y = y - T(kDomainOrigin);
    z = z - T(kDomainOrigin);
    T uf = interp_axis(g, U, 0, x, y, z, c.dx, c.dy, c.dz);
    T vf = interp_axis(g, V, 1, x, y, z, c.dx, c.dy, c.dz);
    T wf = interp_axis(g, W, 2, x, y, z, c.dx, c.dy, c.dz);
    T a = uf - up, b = vf - vp, e = wf - wp;
    T vrel = sqrt((a * a + b * b) + e * e);
    T Rep = (diam * vrel) * c.Re;
    T f = c.drag ? (T(1.0) + T(0.15) * T(pow((double)Rep, 0.687))) : T(1.0);
    T taup = c.kap * (diam * diam);
    T taui = T(1) / taup;
    du = ((f * (uf - up)) * taui) - c.g;
    dv = (f * (vf - vp)) * taui;
    dw = (f * (wf - wp)) * taui;
}

template <typename T>
__global__ void k_p_motion(Geom g, const T *U, const T *V, const T *W, MotionCoef<T> c, long n, const T *xs,
                           const T *ys, const T *zs, const T *us, const T *vs, const T *ws, const T *diams,
                           T *dudts, T *dvdts, T *dwdts)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    T du, dv, dw;
    motion_one(g, U, V, W, c, xs[p], ys[p], zs[p], us[p], vs[p], ws[p], diams[p], du, dv, dw);
    dudts[p] = du;
    dvdts[p] = dv;
    dwdts[p] = dw;
}

// update! (src/particles/update.jl:48-62)
template <typename T>
__global__ void k_p_update(long n, DtRef dt, T *xs, T *ys, T *zs, const T *xs0, const T *ys0, const T *zs0, T *us,
                           T *vs, T *ws, const T *us0, const T *vs0, const T *ws0, const T *dudts, const T *dvdts,
                           const T *dwdts)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    T rdt = T(dt.factor * (*dt.src));
    T x0 = xs0[p], y0 = ys0[p], z0 = zs0[p];
    T u0 = us0[p], v0 = vs0[p], w0 = ws0[p];
    T a = dudts[p], b = dvdts[p], e = dwdts[p];
    xs[p] = x0 + rdt * u0;
    ys[p] = y0 + rdt * v0;
    zs[p] = z0 + rdt * w0;
    us[p] = u0 + rdt * a;
    vs[p] = v0 + rdt * b;
    ws[p] = w0 + rdt * e;
}

// boundary! + motion! + update! of one RK stage in one pass (src/evolve.jl:83-91,
// 110-118).  Stage 1: wrap xss (shadow xss2), forces from Us at xss/uss,
// xss2 = xss + dt uss, uss2 = uss + dt f.  Stage 2: wrap xss2 (shadow xss),
// forces from Us2 at xss2/uss2, xss = xss + dt uss, uss = uss + dt f.
template <typename T>
__device__ inline void p_stage_one(const Geom &g, const T *U, const T *V, const T *W, const MotionCoef<T> &c, long p,
                                   int stage, T rdt, T *xs, T *ys, T *zs, T *xs2, T *ys2, T *zs2, T *us, T *vs,
                                   T *ws, T *us2, T *vs2, T *ws2, const T *diams, T *dudts, T *dvdts, T *dwdts)
{
    T x = xs[p], y = ys[p], z = zs[p], x2 = xs2[p], y2 = ys2[p], z2 = zs2[p];
    T du, dv, dw;
    if (stage == 1) {
        wrap_axis(x, x2);
        wrap_axis(y, y2);
        wrap_axis(z, z2);
        T up = us[p], vp = vs[p], wp = ws[p];
        motion_one(g, U, V, W, c, x, y, z, up, vp, wp, diams[p], du, dv, dw);
        xs[p] = x;
        ys[p] = y;
        zs[p] = z;
        xs2[p] = x + rdt * up;
        ys2[p] = y + rdt * vp;
        zs2[p] = z + rdt * wp;
        us2[p] = up + rdt * du;
        vs2[p] = vp + rdt * dv;
        ws2[p] = wp + rdt * dw;
    } else {
        wrap_axis(x2, x);
        wrap_axis(y2, y);
        wrap_axis(z2, z);
        T up = us2[p], vp = vs2[p], wp = ws2[p];
        motion_one(g, U, V, W, c, x2, y2, z2, up, vp, wp, diams[p], du, dv, dw);
        xs2[p] = x2;
        ys2[p] = y2;
        zs2[p] = z2;
        T u0 = us[p], v0 = vs[p], w0 = ws[p];
        xs[p] = x + rdt * u0;
        ys[p] = y + rdt * v0;
        zs[p] = z + rdt * w0;
        us[p] = u0 + rdt * du;
        vs[p] = v0 + rdt * dv;
        ws[p] = w0 + rdt * dw;
    }
    dudts[p] = du;
    dvdts[p] = dv;
    dwdts[p] = dw;
}

template <typename T>
__global__ void k_p_stage(Geom g, const T *U, const T *V, const T *W, MotionCoef<T> c, long n, int stage,
                          DtRef dt, T *xs, T *ys, T *zs, T *xs2, T *ys2, T *zs2, T *us, T *vs, T *ws, T *us2,
                          T *vs2, T *ws2, const T *diams, T *dudts, T *dvdts, T *dwdts)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    p_stage_one(g, U, V, W, c, p, stage, T(dt.factor * (*dt.src)), xs, ys, zs, xs2, ys2, zs2, us, vs, ws, us2, vs2,
                ws2, diams, dudts, dvdts, dwdts);
}

template <typename T>
static MotionCoef<T> motion_coef(const Box &b, const ParticleProfile &pp)
{
    MotionCoef<T> c;
    c.dx = T(spacing(b.g.nx));
    c.dy = T(spacing(b.g.ny));
    c.dz = T(spacing(b.g.nz));
    c.Re = T(b.Re());
    c.g = T(pp.desc.gravity);
    c.kap = T(((1.0 / 18.0) * pp.desc.density_ratio) * b.Re());
    c.drag = pp.desc.drag;
    return c;
}

#define PR(T, q) ((T *)pp.props[q])

template <typename T>
static int p_boundary_t(Box &b, ParticleProfile &pp, int stage)
{
    long n = pp.desc.nvalid;
    if (n == 0) return LCS_OK;
    int a = stage == 1 ? LCS_PROP_X : LCS_PROP_X2, s = stage == 1 ? LCS_PROP_X2 : LCS_PROP_X;
    hipLaunchKernelGGL(k_p_boundary<T>, dim3(pgrid(n)), dim3(256), 0, b.stream, n, PR(T, a), PR(T, a + 1),
                       PR(T, a + 2), PR(T, s), PR(T, s + 1), PR(T, s + 2));
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

template <typename T>
static int p_motion_t(Box &b, ParticleProfile &pp, int stage)
{
    long n = pp.desc.nvalid;
    if (n == 0) return LCS_OK;
    void *const *Us = stage == 1 ? &b.f[LCS_FIELD_U] : &b.f[LCS_FIELD_U2];
    int x = stage == 1 ? LCS_PROP_X : LCS_PROP_X2, u = stage == 1 ? LCS_PROP_U : LCS_PROP_U2;
    hipLaunchKernelGGL(k_p_motion<T>, dim3(pgrid(n)), dim3(256), 0, b.stream, b.g, (const T *)Us[0],
                       (const T *)Us[1], (const T *)Us[2], motion_coef<T>(b, pp), n, PR(T, x), PR(T, x + 1),
                       PR(T, x + 2), PR(T, u), PR(T, u + 1), PR(T, u + 2), PR(T, LCS_PROP_DIAM),
                       PR(T, LCS_PROP_DUDT), PR(T, LCS_PROP_DVDT), PR(T, LCS_PROP_DWDT));
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

template <typename T>
static int p_update_t(Box &b, ParticleProfile &pp, int stage, DtRef dt)
{
    long n = pp.desc.nvalid;
    if (n == 0) return LCS_OK;
    // xss_update/xss0_update/uss_update/uss0_update: src/particles/update.jl:1-11
    int x = stage == 1 ? LCS_PROP_X2 : LCS_PROP_X, x0 = LCS_PROP_X;
    int u = stage == 1 ? LCS_PROP_U2 : LCS_PROP_U, u0 = LCS_PROP_U;
    hipLaunchKernelGGL(k_p_update<T>, dim3(pgrid(n)), dim3(256), 0, b.stream, n, dt, PR(T, x), PR(T, x + 1),
                       PR(T, x + 2), PR(T, x0), PR(T, x0 + 1), PR(T, x0 + 2), PR(T, u), PR(T, u + 1),
                       PR(T, u + 2), PR(T, u0), PR(T, u0 + 1), PR(T, u0 + 2), PR(T, LCS_PROP_DUDT),
                       PR(T, LCS_PROP_DVDT), PR(T, LCS_PROP_DWDT));
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

template <typename T>
static int p_stage_t(Box &b, ParticleProfile &pp, int stage, DtRef dt)
{
    long n = pp.desc.nvalid;
    if (n == 0) return LCS_OK;
    void *const *Us = stage == 1 ? &b.f[LCS_FIELD_U] : &b.f[LCS_FIELD_U2];
    // algorithmic per particle: stage 1 reads x,u,d writes x*,u* (13), stage 2 19
    ProfScope ps(b, "particles_stage", (double)n * sizeof(T) * (stage == 1 ? 13 : 19));
    hipLaunchKernelGGL(k_p_stage<T>, dim3(pgrid(n)), dim3(256), 0, b.stream, b.g, (const T *)Us[0],
                       (const T *)Us[1], (const T *)Us[2], motion_coef<T>(b, pp), n, stage, dt, PR(T, LCS_PROP_X),
                       PR(T, LCS_PROP_Y), PR(T, LCS_PROP_Z), PR(T, LCS_PROP_X2), PR(T, LCS_PROP_Y2),
                       PR(T, LCS_PROP_Z2), PR(T, LCS_PROP_U), PR(T, LCS_PROP_V), PR(T, LCS_PROP_W),
                       PR(T, LCS_PROP_U2), PR(T, LCS_PROP_V2), PR(T, LCS_PROP_W2), (const T *)PR(T, LCS_PROP_DIAM),
                       PR(T, LCS_PROP_DUDT), PR(T, LCS_PROP_DVDT), PR(T, LCS_PROP_DWDT));
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

int launch_p_boundary(Box &b, ParticleProfile &pp, int stage)
{
    return b.esize == 8 ? p_boundary_t<double>(b, pp, stage) : p_boundary_t<float>(b, pp, stage);
}
int launch_p_motion(Box &b, ParticleProfile &pp, int stage)
{
    return b.esize == 8 ? p_motion_t<double>(b, pp, stage) : p_motion_t<float>(b, pp, stage);
}
int launch_p_update(Box &b, ParticleProfile &pp, int stage, DtRef dt)
{
    return b.esize == 8 ? p_update_t<double>(b, pp, stage, dt) : p_update_t<float>(b, pp, stage, dt);
}
int launch_p_stage(Box &b, ParticleProfile &pp, int stage, DtRef dt)
{
    return b.esize == 8 ? p_stage_t<double>(b, pp, stage, dt) : p_stage_t<float>(b, pp, stage, dt);
}

// ---------------------------------------------------------------------------
// Particle dt (src/particles/timestep.jl:15-25): max diameter over all ntotal
// slots, then min(0.5 taup, dtmax); flow/particle minimum (src/timestep.jl:14-18).
// ---------------------------------------------------------------------------
__device__ inline double block_max_d(double v, double *lds)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) lds[w] = v;
    __syncthreads();
    double s = -INFINITY;
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        for (int q = 0; q < nw; q++) s = fmax(s, lds[q]);
    }
    __syncthreads();
    return s;
}

template <typename T>
__global__ void k_p_maxdiam(long n, const T *d, double *part)
{
    __shared__ double lds[16];
    double m = -INFINITY;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x)
        m = fmax(m, (double)d[p]);
    m = block_max_d(m, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void k_p_dt(const double *part, int nblk, double kap, double dtmax, double *dst)
{
    __shared__ double lds[16];
    double m = -INFINITY;
    for (int q = threadIdx.x; q < nblk; q += blockDim.x) m = fmax(m, part[q]);
    m = block_max_d(m, lds);
    if (threadIdx.x == 0) {
        double taup = kap * (m * m);
        *dst = fmin(0.5 * taup, dtmax);
    }
}

int launch_p_timestep(Box &b, ParticleProfile &pp, double dtmax, double *dtp_dev)
{
    long n = pp.desc.ntotal;
    int blocks = (int)std::min<long>(kMaxPartBlocks, std::max(1L, (n + 255) / 256));
    double *part = b.part + 3L * kMaxPartBlocks;  // keep the CFL partials intact
    ProfScope ps(b, "particles_maxdiam", (double)n * b.esize);
    if (b.esize == 8)
        hipLaunchKernelGGL(k_p_maxdiam<double>, dim3(blocks), dim3(256), 0, b.stream, n,
                           (const double *)pp.props[LCS_PROP_DIAM], part);
    else
        hipLaunchKernelGGL(k_p_maxdiam<float>, dim3(blocks), dim3(256), 0, b.stream, n,
                           (const float *)pp.props[LCS_PROP_DIAM], part);
    LCS_HIP(hipGetLastError());
    double kap = ((1.0 / 18.0) * pp.desc.density_ratio) * b.Re();
    hipLaunchKernelGGL(k_p_dt, dim3(1), dim3(256), 0, b.stream, part, blocks, kap, dtmax, dtp_dev);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

__global__ void k_finalize_dt(DevScalars *S, int np)
{
    double dt = S->dt_flow;
    for (int q = 0; q < np; q++) dt = fmin(dt, S->dtp[q]);
    S->dt = dt;
}

int launch_finalize_dt(Box &b, int nprofiles, double)
{
    hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(1), 0, b.stream, b.S, nprofiles);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// ---------------------------------------------------------------------------
// Cell re-binning: makeindex! (src/particles/makeindex.jl:29-67)
// ---------------------------------------------------------------------------
// hash = encode(hasher, floor((x - x0)/dxc) + 1) over axes 0:nc+1
// (src/utils/hasher.jl:15-38); keys are < 2^32 (hasher.jl:7-8).
template <typename T>
__global__ void k_hash(long n, const T *xs, const T *ys, const T *zs, T dxc0, T dxc1, T dxc2, int nc0, int nc1,
                       unsigned *keys, unsigned *vals)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const T *X[3] = {xs, ys, zs};
    const T dxc[3] = {dxc0, dxc1, dxc2};
    const int nc[2] = {nc0, nc1};
    int64_t idx = 1, stride = 1;
    for (int a = 0; a < 3; a++) {
        int c = (int)floor((X[a][p] - T(kDomainOrigin)) / dxc[a]);
        int64_t ci = (int64_t)c + 1;
        idx += ci * stride;
        if (a < 2) stride *= (nc[a] + 2);
    }
    keys[p] = (unsigned)idx;
    vals[p] = (unsigned)p;
}

constexpr int kSortBlock = 256;

__global__ void __launch_bounds__(kSortBlock) k_rs_hist(const unsigned *keys, long n, int shift, long chunk,
                                                      unsigned *hist, int nblk)
{
    __shared__ unsigned cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    long lo = (long)blockIdx.x * chunk, hi = std::min(n, lo + chunk);
    for (long p = lo + threadIdx.x; p < hi; p += kSortBlock) atomicAdd(&cnt[(keys[p] >> shift) & 255u], 1u);
    __syncthreads();
    hist[(long)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of each digit's row of per-block counts, hist[d * nblk + blk]:
// one workgroup per digit, coalesced chunks of 256 with a wave scan each.
// The row totals go to tot[d]; k_rs_scatter adds the digits' exclusive
// prefix of those totals.
__global__ void __launch_bounds__(256) k_rs_scan_rows(unsigned *hist, int nblk, unsigned *tot)
{
    __shared__ unsigned wsum[4];
    __shared__ unsigned carry;
    unsigned *row = hist + (long)blockIdx.x * nblk;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    for (int base = 0; base < nblk; base += 256) {
        const int q = base + (int)threadIdx.x;
        const unsigned v = q < nblk ? row[q] : 0u;
        unsigned x = v;  // inclusive scan within the wave
        for (int o = 1; o < 64; o <<= 1) {
            unsigned y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        unsigned pre = carry;
        for (int k = 0; k < w; k++) pre += wsum[k];
        if (q < nblk) row[q] = pre + x - v;
        __syncthreads();  // every thread has read carry and wsum
        if (threadIdx.x == 255) carry = pre + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// Stable scatter: items keep their input order within a digit.  Ranks inside
// a 256-item tile come from 8 ballots per wave (lanes with equal digit) plus
// per-wave digit counts in LDS.
__global__ void __launch_bounds__(kSortBlock) k_rs_scatter(const unsigned *kin, const unsigned *vin, unsigned *kout,
                                                         unsigned *vout, long n, int shift, long chunk,
                                                         const unsigned *hist, int nblk, const unsigned *tot)
{
    __shared__ unsigned run[256];
    __shared__ unsigned wc[4][256];
    __shared__ unsigned dsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    {
        // start of digit threadIdx.x: exclusive prefix of the digit totals
        const unsigned tv = tot[threadIdx.x];
        unsigned x = tv;
        for (int o = 1; o < 64; o <<= 1) {
            unsigned y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) dsum[w] = x;
        __syncthreads();
        unsigned pre = x - tv;
        for (int k = 0; k < w; k++) pre += dsum[k];
        run[threadIdx.x] = hist[(long)threadIdx.x * nblk + blockIdx.x] + pre;
    }
    long lo = (long)blockIdx.x * chunk, hi = std::min(n, lo + chunk);
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (long base = lo; base < hi; base += kSortBlock) {
        for (int q = 0; q < 4; q++) wc[q][threadIdx.x] = 0;
        __syncthreads();
        long p = base + threadIdx.x;
        bool valid = p < hi;
        unsigned key = valid ? kin[p] : 0u, val = valid ? vin[p] : 0u;
        unsigned d = (key >> shift) & 255u;
        unsigned long long m = __ballot(valid);
        for (int bit = 0; bit < 8; bit++) {
            unsigned long long bb = __ballot(valid && ((d >> bit) & 1u));
            m &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        unsigned rank = (unsigned)__popcll(m & lt);
        if (valid && rank == 0) wc[w][d] = (unsigned)__popcll(m);
        __syncthreads();
        unsigned pre = 0;
        for (int q = 0; q < w; q++) pre += wc[q][d];
        unsigned dst = run[d] + pre + rank;
        __syncthreads();
        if (valid) {
            kout[dst] = key;
            vout[dst] = val;
        }
        unsigned tsum = wc[0][threadIdx.x] + wc[1][threadIdx.x] + wc[2][threadIdx.x] + wc[3][threadIdx.x];
        run[threadIdx.x] += tsum;
        __syncthreads();
    }
}

template <typename E>
__global__ void k_permute(long n, long ntotal, const unsigned *perm, const E *src, E *dst)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ntotal) return;
    dst[p] = p < n ? src[perm[p]] : src[p];
}

__global__ void k_store_hashes(long n, const unsigned *keys, int64_t *hashes)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    hashes[p] = (int64_t)keys[p];
}

// starts[x] = searchsortedfirst(hashes, x), stops[x] = searchsortedlast(hashes, x)
// for every hash x (src/utils/fast.jl:8-40), 1-based.
__global__ void k_starts_stops(const int64_t *v, long n, long ncell, int64_t *starts, int64_t *stops)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int64_t x = c + 1;
    long lo = 1, len = n;
    while (len != 0) {
        long half = len >> 1;
        long m = lo + half;
        bool isless = v[m - 1] < x;
        lo = isless ? m + 1 : lo;
        len = isless ? len - (half + 1) : half;
    }
    starts[c] = lo;
    long l2 = 0, h2 = n + 1;
    while (l2 < h2 - 1) {
        long m = l2 + ((h2 - l2) >> 1);
        bool isless = x < v[m - 1];
        h2 = isless ? m : h2;
        l2 = isless ? l2 : m;
    }
    stops[c] = l2;
}

template <typename T>
static int makeindex_t(Box &b, ParticleProfile &pp, int stage)
{
    long n = pp.desc.nvalid, ntotal = pp.desc.ntotal;
    long ncell = (long)(pp.nc[0] + 2) * (pp.nc[1] + 2) * (pp.nc[2] + 2);
    int nblk = (int)std::min<long>(1024, std::max(1L, (n + 4 * kSortBlock - 1) / (4 * kSortBlock)));
    long chunk = (n + nblk - 1) / nblk;
    chunk = ((chunk + kSortBlock - 1) / kSortBlock) * kSortBlock;
    if (chunk == 0) chunk = kSortBlock;
    size_t need = sizeof(unsigned) * (4 * (size_t)std::max(1L, n) + 256 * (size_t)nblk + 256);
    if (need > pp.sort_tmp_bytes) {
        if (pp.sort_tmp) LCS_HIP(hipFree(pp.sort_tmp));
        pp.sort_tmp = nullptr;
        LCS_HIP(hipMalloc(&pp.sort_tmp, need));
        b.bytes += need - pp.sort_tmp_bytes;
        pp.sort_tmp_bytes = need;
    }
    unsigned *kA = (unsigned *)pp.sort_tmp, *vA = kA + std::max(1L, n), *kB = vA + std::max(1L, n),
             *vB = kB + std::max(1L, n), *hist = vB + std::max(1L, n), *tot = hist + 256L * nblk;
    int x = stage == 1 ? LCS_PROP_X : LCS_PROP_X2;
    double dxc[3] = {spacing(b.g.nx) * pp.desc.grid_ratio, spacing(b.g.ny) * pp.desc.grid_ratio,
                     spacing(b.g.nz) * pp.desc.grid_ratio};
    if (n > 0) {
        hipLaunchKernelGGL(k_hash<T>, dim3(pgrid(n)), dim3(256), 0, b.stream, n, (const T *)pp.props[x],
                           (const T *)pp.props[x + 1], (const T *)pp.props[x + 2], T(dxc[0]), T(dxc[1]), T(dxc[2]),
                           pp.nc[0], pp.nc[1], kA, vA);
        LCS_HIP(hipGetLastError());
        int bits = 0;
        while (bits < 32 && ((unsigned long long)ncell >> bits) != 0) bits++;
        for (int shift = 0; shift < bits; shift += 8) {
            hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(kSortBlock), 0, b.stream, kA, n, shift, chunk, hist, nblk);
            hipLaunchKernelGGL(k_rs_scan_rows, dim3(256), dim3(256), 0, b.stream, hist, nblk, tot);
            hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(kSortBlock), 0, b.stream, kA, vA, kB, vB, n, shift,
                               chunk, hist, nblk, tot);
            LCS_HIP(hipGetLastError());
            std::swap(kA, kB);
            std::swap(vA, vB);
        }
        hipLaunchKernelGGL(k_store_hashes, dim3(pgrid(n)), dim3(256), 0, b.stream, n, kA, (int64_t *)pp.hashes);
        LCS_HIP(hipGetLastError());
        // sortbyperm! of all 17 properties (src/utils/sort.jl:7-22): permute
        // into the staging buffer, then swap buffers
        for (int q = 0; q < LCS_N_PROPERTIES; q++) {
            if (q == LCS_PROP_ID)
                hipLaunchKernelGGL(k_permute<int64_t>, dim3(pgrid(ntotal)), dim3(256), 0, b.stream, n, ntotal, vA,
                                   (const int64_t *)pp.props[q], (int64_t *)prop_spare(pp, q));
            else
                hipLaunchKernelGGL(k_permute<T>, dim3(pgrid(ntotal)), dim3(256), 0, b.stream, n, ntotal, vA,
                                   (const T *)pp.props[q], (T *)prop_spare(pp, q));
            LCS_HIP(hipGetLastError());
            std::swap(pp.props[q], prop_spare(pp, q));
        }
        LCS_HIP(hipMemcpyAsync(pp.perm, vA, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, b.stream));
    }
    hipLaunchKernelGGL(k_starts_stops, dim3(pgrid(ncell)), dim3(256), 0, b.stream, (const int64_t *)pp.hashes, n,
                       ncell, (int64_t *)pp.starts, (int64_t *)pp.stops);
    LCS_HIP(hipGetLastError());
    pp.indexed = true;
    return LCS_OK;
}

// ---------------------------------------------------------------------------
// Storage order for locality.  Each particle is advanced independently, so the
// slot a particle occupies changes no result; sorting slots by the 8^3-cell
// tile of the position makes the trilinear gathers of k_p_stage hit the same
// few cache lines within a wave instead of 24 random lines per particle.  The
// logical order -- what upload/download and makeindex see -- is kept through
// the map loc (logical index -> slot).
// ---------------------------------------------------------------------------
constexpr int kLocTile = 8;

template <typename T>
__global__ void k_loc_key(long n, const T *xs, const T *ys, const T *zs, T idx0, T idx1, T idx2, int n0, int n1,
                          int n2, int t0, int t1, unsigned *keys, unsigned *vals)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const T x[3] = {xs[p], ys[p], zs[p]};
    const T idx[3] = {idx0, idx1, idx2};
    const int nn[3] = {n0, n1, n2};
    int t[3];
    for (int a = 0; a < 3; a++) {
        T f = floor((x[a] - T(kDomainOrigin)) * idx[a]);
        int c = f < T(0) ? 0 : (f >= T(nn[a]) ? nn[a] - 1 : (int)f);  // any position gets a tile
        t[a] = c / kLocTile;
    }
    keys[p] = (unsigned)(t[0] + t0 * (t[1] + t1 * t[2]));
    vals[p] = (unsigned)p;
}

// inv[slot_old] = slot_new for the sorted slots, identity past n
__global__ void k_loc_inverse(long n, long ntotal, const unsigned *sorted_vals, unsigned *inv)
{
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ntotal) return;
    if (j < n)
        inv[sorted_vals[j]] = (unsigned)j;
    else
        inv[j] = (unsigned)j;
}

// loc[i] = inv[loc[i]] (or inv[i] when the map was off)
__global__ void k_loc_compose(long ntotal, const unsigned *inv, unsigned *loc, int loc_on)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntotal) return;
    loc[i] = inv[loc_on ? loc[i] : (unsigned)i];
}

template <typename E>
__global__ void k_loc_move(long ntotal, const unsigned *loc, const E *src, E *dst, int scatter)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntotal) return;
    if (scatter)
        dst[loc[i]] = src[i];
    else
        dst[i] = src[loc[i]];
}

int launch_p_loc_move(Box &b, ParticleProfile &pp, const void *src, void *dst, size_t esize, bool scatter)
{
    long nt = pp.desc.ntotal;
    if (nt == 0) return LCS_OK;
    if (esize == 8)
        hipLaunchKernelGGL(k_loc_move<uint64_t>, dim3(pgrid(nt)), dim3(256), 0, b.stream, nt,
                           (const unsigned *)pp.loc, (const uint64_t *)src, (uint64_t *)dst, (int)scatter);
    else
        hipLaunchKernelGGL(k_loc_move<uint32_t>, dim3(pgrid(nt)), dim3(256), 0, b.stream, nt,
                           (const unsigned *)pp.loc, (const uint32_t *)src, (uint32_t *)dst, (int)scatter);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

int launch_p_unlocalize(Box &b, ParticleProfile &pp)
{
    if (!pp.loc_on) return LCS_OK;
    for (int q = 0; q < LCS_N_PROPERTIES; q++) {
        int rc = launch_p_loc_move(b, pp, pp.props[q], prop_spare(pp, q), q == LCS_PROP_ID ? 8 : b.esize, false);
        if (rc) return rc;
        std::swap(pp.props[q], prop_spare(pp, q));
    }
    pp.loc_on = false;
    return LCS_OK;
}

// radix sort of (keys, slot) pairs over n items in pp.sort_tmp; returns the
// buffers holding the sorted pairs
static int sort_pairs(Box &b, ParticleProfile &pp, long n, int bits, unsigned **keys, unsigned **vals,
                      unsigned **spare)
{
    int nblk = (int)std::min<long>(1024, std::max(1L, (n + 4 * kSortBlock - 1) / (4 * kSortBlock)));
    long chunk = (n + nblk - 1) / nblk;
    chunk = ((chunk + kSortBlock - 1) / kSortBlock) * kSortBlock;
    if (chunk == 0) chunk = kSortBlock;
    unsigned *kA = *keys, *vA = *vals, *kB = vA + std::max(1L, n), *vB = kB + std::max(1L, n),
             *hist = vB + std::max(1L, n), *tot = hist + 256L * nblk;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(kSortBlock), 0, b.stream, kA, n, shift, chunk, hist, nblk);
        hipLaunchKernelGGL(k_rs_scan_rows, dim3(256), dim3(256), 0, b.stream, hist, nblk, tot);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(kSortBlock), 0, b.stream, kA, vA, kB, vB, n, shift, chunk,
                           hist, nblk, tot);
        LCS_HIP(hipGetLastError());
        std::swap(kA, kB);
        std::swap(vA, vB);
    }
    *keys = kA;
    *vals = vA;
    *spare = kB;
    return LCS_OK;
}

template <typename T>
static int localize_t(Box &b, ParticleProfile &pp)
{
    long n = pp.desc.nvalid, nt = pp.desc.ntotal;
    if (n < 2) return LCS_OK;
    if (!pp.loc) {
        LCS_HIP(hipMalloc(&pp.loc, sizeof(unsigned) * (size_t)nt));
        LCS_HIP(hipMalloc(&pp.loc_tmp, sizeof(unsigned) * (size_t)nt));
        b.bytes += 2 * sizeof(unsigned) * (size_t)nt;
    }
    int nblk = (int)std::min<long>(1024, std::max(1L, (n + 4 * kSortBlock - 1) / (4 * kSortBlock)));
    size_t need = sizeof(unsigned) * (4 * (size_t)n + 256 * (size_t)nblk + 256);
    if (need > pp.sort_tmp_bytes) {
        if (pp.sort_tmp) LCS_HIP(hipFree(pp.sort_tmp));
        pp.sort_tmp = nullptr;
        LCS_HIP(hipMalloc(&pp.sort_tmp, need));
        b.bytes += need - pp.sort_tmp_bytes;
        pp.sort_tmp_bytes = need;
    }
    const int nn[3] = {b.g.nx, b.g.ny, b.g.nz};
    int nt3[3];
    long ntiles = 1;
    for (int a = 0; a < 3; a++) {
        nt3[a] = (nn[a] + kLocTile - 1) / kLocTile;
        ntiles *= nt3[a];
    }
    int bits = 0;
    while (bits < 32 && ((unsigned long long)(ntiles - 1) >> bits) != 0) bits++;
    unsigned *keys = (unsigned *)pp.sort_tmp, *vals = keys + n, *spare = nullptr;
    ProfScope ps(b, "particles_localize", (double)n * (17.0 * sizeof(T) + 16.0 * ((bits + 7) / 8)));
    hipLaunchKernelGGL(k_loc_key<T>, dim3(pgrid(n)), dim3(256), 0, b.stream, n, (const T *)pp.props[LCS_PROP_X],
                       (const T *)pp.props[LCS_PROP_Y], (const T *)pp.props[LCS_PROP_Z], T(1.0 / spacing(nn[0])),
                       T(1.0 / spacing(nn[1])), T(1.0 / spacing(nn[2])), nn[0], nn[1], nn[2], nt3[0], nt3[1], keys,
                       vals);
    LCS_HIP(hipGetLastError());
    int rc = sort_pairs(b, pp, n, bits, &keys, &vals, &spare);
    if (rc) return rc;
    for (int q = 0; q < LCS_N_PROPERTIES; q++) {
        if (q == LCS_PROP_ID)
            hipLaunchKernelGGL(k_permute<int64_t>, dim3(pgrid(nt)), dim3(256), 0, b.stream, n, nt, vals,
                               (const int64_t *)pp.props[q], (int64_t *)prop_spare(pp, q));
        else
            hipLaunchKernelGGL(k_permute<T>, dim3(pgrid(nt)), dim3(256), 0, b.stream, n, nt, vals,
                               (const T *)pp.props[q], (T *)prop_spare(pp, q));
        LCS_HIP(hipGetLastError());
        std::swap(pp.props[q], prop_spare(pp, q));
    }
    hipLaunchKernelGGL(k_loc_inverse, dim3(pgrid(nt)), dim3(256), 0, b.stream, n, nt, vals,
                       (unsigned *)pp.loc_tmp);
    hipLaunchKernelGGL(k_loc_compose, dim3(pgrid(nt)), dim3(256), 0, b.stream, nt, (const unsigned *)pp.loc_tmp,
                       (unsigned *)pp.loc, (int)pp.loc_on);
    LCS_HIP(hipGetLastError());
    pp.loc_on = true;
    (void)spare;
    return LCS_OK;
}

int launch_p_localize(Box &b, ParticleProfile &pp)
{
    return b.esize == 8 ? localize_t<double>(b, pp) : localize_t<float>(b, pp);
}

int launch_p_makeindex(Box &b, ParticleProfile &pp, int stage)
{
    // makeindex! sorts the logical order stably: restore it first
    int rc = launch_p_unlocalize(b, pp);
    if (rc) return rc;
    return b.esize == 8 ? makeindex_t<double>(b, pp, stage) : makeindex_t<float>(b, pp, stage);
}

// ---------------------------------------------------------------------------
// Builder-defined particle initialiser (mirrors the oracle's o_init_particles)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_p_init(long n, long ntotal, uint64_t seed, double st_min, double st_max, double taueta,
                         double density_ratio, double Re, T *xs, T *ys, T *zs, T *diams, int64_t *ids)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ntotal) return;
    if (p >= n) {
        ids[p] = kInvalidId;
        return;
    }
    xs[p] = T(kDomainLength * rand_u01(seed, 0, (uint64_t)p));
    ys[p] = T(kDomainLength * rand_u01(seed, 1, (uint64_t)p));
    zs[p] = T(kDomainLength * rand_u01(seed, 2, (uint64_t)p));
    double st = st_min;
    if (st_max != st_min) st = st_min * pow(st_max / st_min, rand_u01(seed, 3, (uint64_t)p));
    diams[p] = T(sqrt(18.0 * st * taueta / (density_ratio * Re)));
    ids[p] = (int64_t)p + 1;
}

template <typename T>
__global__ void k_p_fluid(Geom g, const T *U, const T *V, const T *W, long n, const T *xs, const T *ys, const T *zs,
                          T *us, T *vs, T *ws)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    T dx = T(spacing(g.nx)), dy = T(spacing(g.ny)), dz = T(spacing(g.nz));
    T x = xs[p] - T(kDomainOrigin), y = ys[p] - T(kDomainOrigin), z = zs[p] - T(kDomainOrigin);
    us[p] = interp_axis(g, U, 0, x, y, z, dx, dy, dz);
    vs[p] = interp_axis(g, V, 1, x, y, z, dx, dy, dz);
    ws[p] = interp_axis(g, W, 2, x, y, z, dx, dy, dz);
}

template <typename T>
static int p_init_t(Box &b, ParticleProfile &pp, uint64_t seed, double st_min, double st_max, double eps,
                    int zero_slip)
{
    long n = pp.desc.nvalid, ntotal = pp.desc.ntotal;
    for (int q = 1; q < LCS_N_PROPERTIES; q++)
        LCS_HIP(hipMemsetAsync(pp.props[q], 0, sizeof(T) * ntotal, b.stream));
    double taueta = std::sqrt(1.0 / (eps * b.Re()));
    hipLaunchKernelGGL(k_p_init<T>, dim3(pgrid(ntotal)), dim3(256), 0, b.stream, n, ntotal, seed, st_min, st_max,
                       taueta, pp.desc.density_ratio, b.Re(), PR(T, LCS_PROP_X), PR(T, LCS_PROP_Y), PR(T, LCS_PROP_Z),
                       PR(T, LCS_PROP_DIAM), (int64_t *)pp.props[LCS_PROP_ID]);
    LCS_HIP(hipGetLastError());
    if (zero_slip && n > 0) {
        hipLaunchKernelGGL(k_p_fluid<T>, dim3(pgrid(n)), dim3(256), 0, b.stream, b.g, (const T *)b.f[0],
                           (const T *)b.f[1], (const T *)b.f[2], n, PR(T, LCS_PROP_X), PR(T, LCS_PROP_Y),
                           PR(T, LCS_PROP_Z), PR(T, LCS_PROP_U), PR(T, LCS_PROP_V), PR(T, LCS_PROP_W));
        LCS_HIP(hipGetLastError());
    }
    return LCS_OK;
}

int launch_p_init(Box &b, ParticleProfile &pp, uint64_t seed, double st_min, double st_max, double eps,
                  int zero_slip)
{
    return b.esize == 8 ? p_init_t<double>(b, pp, seed, st_min, st_max, eps, zero_slip)
                        : p_init_t<float>(b, pp, seed, st_min, st_max, eps, zero_slip);
}

// ---------------------------------------------------------------------------
// Batched particle kernels (lcs_batch_*): work item (box, blk) covers 256
// particles of one profile; each box lists its profiles' blocks back to back.
// ---------------------------------------------------------------------------
__device__ inline int b_profile(const BoxRec &B, int blk, bool diam)
{
    for (int q = 0; q < B.nprof; q++) {
        const int b0 = diam ? B.prof[q].dblk0 : B.prof[q].blk0, nb = diam ? B.prof[q].dnblk : B.prof[q].nblk;
        if (blk >= b0 && blk < b0 + nb) return q;
    }
    return -1;
}

template <typename T>
__global__ void __launch_bounds__(256) k_b_p_stage(const BoxRec *tab, const WorkItem *work, int stage)
{
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    const int q = b_profile(B, w.blk, false);
    if (q < 0) return;
    const DevProfile &P = B.prof[q];
    const long p = (long)(w.blk - P.blk0) * blockDim.x + threadIdx.x;
    if (p >= P.nvalid) return;
    MotionCoef<T> c;
    c.dx = T(P.mdx);
    c.dy = T(P.mdy);
    c.dz = T(P.mdz);
    c.Re = T(P.mRe);
    c.g = T(P.mg);
    c.kap = T(P.mkap);
    c.drag = P.drag;
    const int in = stage == 1 ? 0 : 3;
    T *const *pr = (T *const *)P.props;
    p_stage_one(B.g, (const T *)B.f[in], (const T *)B.f[in + 1], (const T *)B.f[in + 2], c, p, stage,
                T((stage == 1 ? 0.5 : 1.0) * B.S->dt), pr[LCS_PROP_X], pr[LCS_PROP_Y], pr[LCS_PROP_Z],
                pr[LCS_PROP_X2], pr[LCS_PROP_Y2], pr[LCS_PROP_Z2], pr[LCS_PROP_U], pr[LCS_PROP_V], pr[LCS_PROP_W],
                pr[LCS_PROP_U2], pr[LCS_PROP_V2], pr[LCS_PROP_W2], (const T *)pr[LCS_PROP_DIAM], pr[LCS_PROP_DUDT],
                pr[LCS_PROP_DVDT], pr[LCS_PROP_DWDT]);
}

int launch_b_particles(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *work, int nwork, int stage)
{
    if (nwork == 0) return LCS_OK;
    if (esize == 8)
        hipLaunchKernelGGL(k_b_p_stage<double>, dim3(nwork), dim3(256), 0, s, tab, work, stage);
    else
        hipLaunchKernelGGL(k_b_p_stage<float>, dim3(nwork), dim3(256), 0, s, tab, work, stage);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

// max diameter over ntotal per profile (partials in slot 3 + profile), then
// the particle dt per profile and the step dt per box
template <typename T>
__global__ void __launch_bounds__(256) k_b_p_maxdiam(const BoxRec *tab, const WorkItem *work)
{
    __shared__ double lds[16];
    const WorkItem w = work[blockIdx.x];
    const BoxRec &B = tab[w.box];
    if (b_halted(B)) return;
    const int q = b_profile(B, w.blk, true);
    if (q < 0) return;
    const DevProfile &P = B.prof[q];
    const long p = (long)(w.blk - P.dblk0) * blockDim.x + threadIdx.x;
    double m = p < P.ntotal ? (double)((const T *)P.props[LCS_PROP_DIAM])[p] : -INFINITY;
    m = block_max_d(m, lds);
    if (threadIdx.x == 0) B.part[(3L + q) * kMaxPartBlocks + (w.blk - P.dblk0)] = m;
}

__global__ void k_b_p_dt(const BoxRec *tab, double dtmax)
{
    __shared__ double lds[16];
    const BoxRec &B = tab[blockIdx.x];
    if (b_halted(B)) return;
    DevScalars *S = B.S;
    for (int q = 0; q < B.nprof; q++) {
        const DevProfile &P = B.prof[q];
        double m = -INFINITY;
        for (int k = threadIdx.x; k < P.dnblk; k += blockDim.x) m = fmax(m, B.part[(3L + q) * kMaxPartBlocks + k]);
        m = block_max_d(m, lds);
        if (threadIdx.x == 0) S->dtp[q] = fmin(0.5 * (P.kap * (m * m)), dtmax);
    }
    if (threadIdx.x == 0) {
        double dt = S->dt_flow;
        for (int q = 0; q < B.nprof; q++) dt = fmin(dt, S->dtp[q]);
        S->dt = dt;
    }
}

int launch_b_particle_dt(hipStream_t s, int esize, const BoxRec *tab, const WorkItem *work, int nwork, int nbox,
                         double dtmax)
{
    if (nwork > 0) {
        if (esize == 8)
            hipLaunchKernelGGL(k_b_p_maxdiam<double>, dim3(nwork), dim3(256), 0, s, tab, work);
        else
            hipLaunchKernelGGL(k_b_p_maxdiam<float>, dim3(nwork), dim3(256), 0, s, tab, work);
    }
    hipLaunchKernelGGL(k_b_p_dt, dim3(nbox), dim3(256), 0, s, tab, dtmax);
    LCS_HIP(hipGetLastError());
    return LCS_OK;
}

}  // namespace lcs
