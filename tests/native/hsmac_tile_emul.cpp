// Host execution of the fused HSMAC sweep kernel's phase functions
// (lcs.jl_amd/csrc/lcs_hsmac_tile.h): every workgroup of the launch grid is
// run thread by thread, phase by phase, in the kernel's barrier order, so a
// test can compare the kernel's own code with the oracle's two colour passes.
#include <vector>

#include "lcs_hsmac_tile.h"

using namespace lcs;

// linear forcing applied on load (coefficients as k_hsmac_sweep sets them):
// null = none; else {A, mean U, mean V, mean W, dt}; power receives the sum
static const double *g_force = nullptr;
static double g_power = 0.0;
// U, V, W of ghost cells loaded from their periodic images (HsArgs::wrap)
static int g_wrap = 0;

template <typename T, int COLORING, int TY, int TX, bool FORCE>
static double run_t(int nx, int ny, int nz, int h, const int *lo, const int *hi, double dt, double omega,
                    void *const *src, void *const *dst, int zc)
{
    using Tile = HsTile<T, COLORING, TY, TX, FORCE>;
    HsArgs<T> a;
    a.g = make_geom(nx, ny, nz, h);
    a.Us = (const T *)src[0];
    a.Vs = (const T *)src[1];
    a.Ws = (const T *)src[2];
    a.Ps = (const T *)src[3];
    a.Ud = (T *)dst[0];
    a.Vd = (T *)dst[1];
    a.Wd = (T *)dst[2];
    a.Pd = (T *)dst[3];
    for (int q = 0; q < 3; q++) {
        a.R.lo[q] = lo[q];
        a.R.hi[q] = hi[q];
    }
    const T dxi = T(1.0 / spacing(nx)), dyi = T(1.0 / spacing(ny)), dzi = T(1.0 / spacing(nz));
    a.c.dxi = dxi;
    a.c.dyi = dyi;
    a.c.dzi = dzi;
    a.c.dl = T(1) / (T(2) * ((dxi * dxi + dyi * dyi) + dzi * dzi));
    a.c.omega = T(omega);
    a.dti = T(1.0 / dt);
    a.zc = zc;
    a.wrap = g_wrap;
    if (g_force) {
        a.xf = 1;
        a.xA = T(g_force[0]);
        for (int q = 0; q < 3; q++) a.xm[q] = T(g_force[1 + q]);
        a.xdt = T(g_force[4]);
    }
    std::vector<T> lds(Tile::LDS_ELEMS);
    std::vector<typename Tile::Thread> th(Tile::NT);
    const int gx = (nx + 2 * h + Tile::TX - 1) / Tile::TX, gy = (ny + 2 * h + TY - 1) / TY;
    const int gz = (nz + 2 * h + zc - 1) / zc;
    double s = 0.0;
    for (int bz = 0; bz < gz; bz++)
        for (int by = 0; by < gy; by++)
            for (int bx = 0; bx < gx; bx++)
                s += Tile::emulate_block(a, lds.data(), th.data(), bx, by, bz, g_force ? &g_power : nullptr);
    return s;
}

template <typename T, int COLORING, int TY, int TX>
static double run(int nx, int ny, int nz, int h, const int *lo, const int *hi, double dt, double omega,
                  void *const *src, void *const *dst, int zc)
{
    return g_force ? run_t<T, COLORING, TY, TX, true>(nx, ny, nz, h, lo, hi, dt, omega, src, dst, zc)
                   : run_t<T, COLORING, TY, TX, false>(nx, ny, nz, h, lo, hi, dt, omega, src, dst, zc);
}

extern "C" int hsmac_sweep_emul(int ty, int tx, int esize, int coloring, int nx, int ny, int nz, int h,
                                const int *lo, const int *hi, double dt, double omega, void *const *src,
                                void *const *dst, int zc, double *div2)
{
#define RUN(T, C, Y, X) *div2 = run<T, C, Y, X>(nx, ny, nz, h, lo, hi, dt, omega, src, dst, zc)
#define BY_SHAPE(T, C)                              \
    if (ty == 16 && tx == 64) RUN(T, C, 16, 64);    \
    else if (ty == 4 && tx == 64) RUN(T, C, 4, 64); \
    else if (ty == 8 && tx == 64) RUN(T, C, 8, 64); \
    else if (ty == 8 && tx == 32) RUN(T, C, 8, 32); \
    else if (ty == 4 && tx == 32) RUN(T, C, 4, 32); \
    else if (ty == 8 && tx == 128) RUN(T, C, 8, 128); \
    else if (ty == 16 && tx == 128) RUN(T, C, 16, 128); \
    else return -1;
    if (esize == 4) {
        if (coloring == LCS_COLORING_RED_BLACK) { BY_SHAPE(float, LCS_COLORING_RED_BLACK) }
        else { BY_SHAPE(float, LCS_COLORING_RED_BLACK_FAST) }
    } else {
        if (coloring == LCS_COLORING_RED_BLACK) { BY_SHAPE(double, LCS_COLORING_RED_BLACK) }
        else { BY_SHAPE(double, LCS_COLORING_RED_BLACK_FAST) }
    }
#undef BY_SHAPE
#undef RUN
    return 0;
}

// the same sweep loading U, V, W of ghost cells from their periodic images,
// optionally with the linear forcing (force null: none)
extern "C" int hsmac_sweep_wrap_emul(int ty, int tx, int esize, int coloring, int nx, int ny, int nz, int h,
                                     const int *lo, const int *hi, double dt, double omega, void *const *src,
                                     void *const *dst, int zc, const double *force, double *div2, double *power)
{
    g_wrap = 1;
    g_force = force;
    g_power = 0.0;
    const int rc = hsmac_sweep_emul(ty, tx, esize, coloring, nx, ny, nz, h, lo, hi, dt, omega, src, dst, zc, div2);
    *power = g_power;
    g_force = nullptr;
    g_wrap = 0;
    return rc;
}

// the same sweep with the linear forcing {A, mean U, V, W, dt} applied to
// U, V, W on load; *power receives the forcing power summed over the interior
extern "C" int hsmac_sweep_forced_emul(int ty, int tx, int esize, int coloring, int nx, int ny, int nz, int h,
                                       const int *lo, const int *hi, double dt, double omega, void *const *src,
                                       void *const *dst, int zc, const double *force, double *div2, double *power)
{
    g_force = force;
    g_power = 0.0;
    const int rc = hsmac_sweep_emul(ty, tx, esize, coloring, nx, ny, nz, h, lo, hi, dt, omega, src, dst, zc, div2);
    *power = g_power;
    g_force = nullptr;
    return rc;
}

// ---------------------------------------------------------------------------
// Both colour passes over the two x face slabs as k_hsmac_xface and
// k_hsmac_xface_commit run them (lcs_hsmac_xface.h): every workgroup of the
// compute launch, then every row of the commit launch.  lo/hi hold the two
// slabs (x- then x+), three entries each.  Returns -1 for slabs the form
// does not handle.
// ---------------------------------------------------------------------------
#include "lcs_hsmac_xface.h"

template <typename T, int COLORING, int PY, int TI = 0>
static int xface_t(int nx, int ny, int nz, int h, const int *lo, const int *hi, double dt, double omega,
                   void *const *f, double *div2)
{
    using X = XFace<T, COLORING, PY, PY, TI>;
    XfArgs<T> a;
    a.g = make_geom(nx, ny, nz, h);
    T *F[4];
    for (int q = 0; q < 4; q++) F[q] = (T *)f[q];
    a.U = F[0];
    a.V = F[1];
    a.W = F[2];
    a.P = F[3];
    for (int s = 0; s < 2; s++)
        for (int q = 0; q < 3; q++) {
            a.R[s].lo[q] = lo[3 * s + q];
            a.R[s].hi[q] = hi[3 * s + q];
        }
    if (!SlabAxes<TI>::fits(a.g, a.R)) return -1;
    const T dxi = T(1.0 / spacing(nx)), dyi = T(1.0 / spacing(ny)), dzi = T(1.0 / spacing(nz));
    a.c.dxi = dxi;
    a.c.dyi = dyi;
    a.c.dzi = dzi;
    a.c.dl = T(1) / (T(2) * ((dxi * dxi + dyi * dyi) + dzi * dzi));
    a.c.omega = T(omega);
    a.dti = T(1.0 / dt);
    const long rows = SlabAxes<TI>::rows(a.R[0]);
    std::vector<T> buf(kXfVals * rows);
    a.buf = buf.data();
    std::vector<T> lds(X::LDS_ELEMS);
    double s = 0.0;
    for (int blk = 0; blk < X::blocks(a.R[0]); blk++) s += X::emulate_block(a, lds.data(), blk);
    for (long r = 0; r < rows; r++) X::commit_row(a.g, a.R[0], a.R[1], F[0], F[1], F[2], F[3], (const T *)buf.data(), r);
    *div2 = s;
    return 0;
}

extern "C" int hsmac_xface_emul(int patch, int esize, int coloring, int nx, int ny, int nz, int h, const int *lo,
                                const int *hi, double dt, double omega, void *const *f, double *div2)
{
#define RUN(T, C, Y) return xface_t<T, C, Y>(nx, ny, nz, h, lo, hi, dt, omega, f, div2)
#define BY_PATCH(T, C)          \
    if (patch == 32) RUN(T, C, 32); \
    if (patch == 16) RUN(T, C, 16); \
    if (patch == 4) RUN(T, C, 4);   \
    return -2;
    if (esize == 4) {
        if (coloring == LCS_COLORING_RED_BLACK) { BY_PATCH(float, LCS_COLORING_RED_BLACK) }
        else { BY_PATCH(float, LCS_COLORING_RED_BLACK_FAST) }
    } else {
        if (coloring == LCS_COLORING_RED_BLACK) { BY_PATCH(double, LCS_COLORING_RED_BLACK) }
        else { BY_PATCH(double, LCS_COLORING_RED_BLACK_FAST) }
    }
#undef BY_PATCH
#undef RUN
}

// the slab pair across axis ti (0 x, 1 y, 2 z) with patch edge 16 or 4
extern "C" int hsmac_slab_pair_emul(int ti, int patch, int esize, int coloring, int nx, int ny, int nz, int h,
                                    const int *lo, const int *hi, double dt, double omega, void *const *f,
                                    double *div2)
{
#define RUN(T, C, Y, A) return xface_t<T, C, Y, A>(nx, ny, nz, h, lo, hi, dt, omega, f, div2)
#define BY_AXIS(T, C, Y)          \
    if (ti == 0) RUN(T, C, Y, 0); \
    if (ti == 1) RUN(T, C, Y, 1); \
    if (ti == 2) RUN(T, C, Y, 2);
#define BY_PATCH(T, C)                      \
    if (patch == 16) { BY_AXIS(T, C, 16) }  \
    if (patch == 4) { BY_AXIS(T, C, 4) }    \
    return -2;
    if (esize == 4) {
        if (coloring == LCS_COLORING_RED_BLACK) { BY_PATCH(float, LCS_COLORING_RED_BLACK) }
        else { BY_PATCH(float, LCS_COLORING_RED_BLACK_FAST) }
    } else {
        if (coloring == LCS_COLORING_RED_BLACK) { BY_PATCH(double, LCS_COLORING_RED_BLACK) }
        else { BY_PATCH(double, LCS_COLORING_RED_BLACK_FAST) }
    }
#undef BY_PATCH
#undef BY_AXIS
#undef RUN
}
