// Host execution of the two-sweep HSMAC kernel's phase functions
// (lcs.jl_amd/csrc/lcs_hsmac_tile2.h): every workgroup of the launch grid is
// run thread by thread, phase by phase, in the kernel's barrier order, so a
// test can compare the kernel's own code with the oracle's four colour passes.
#include <vector>

#include "lcs_hsmac_tile2.h"

using namespace lcs;

template <typename T, int COLORING, int TY, int TX>
static void run(int nx, int ny, int nz, int h, const int *loA, const int *hiA, const int *loB, const int *hiB,
                double dt, double omega, void *const *src, void *const *dst, int zc, double *sa, double *sb)
{
    using Tile = HsTile2<T, COLORING, TY, TX>;
    Hs2Args<T> a;
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
        a.RA.lo[q] = loA[q];
        a.RA.hi[q] = hiA[q];
        a.RB.lo[q] = loB[q];
        a.RB.hi[q] = hiB[q];
    }
    const T dxi = T(1.0 / spacing(nx)), dyi = T(1.0 / spacing(ny)), dzi = T(1.0 / spacing(nz));
    a.c.dxi = dxi;
    a.c.dyi = dyi;
    a.c.dzi = dzi;
    a.c.dl = T(1) / (T(2) * ((dxi * dxi + dyi * dyi) + dzi * dzi));
    a.c.omega = T(omega);
    a.dti = T(1.0 / dt);
    a.zc = zc;
    std::vector<T> lds(Tile::LDS_ELEMS, T(0));
    std::vector<typename Tile::Thread> th(Tile::NT);
    const int gx = (nx + 2 * h + TX - 1) / TX, gy = (ny + 2 * h + TY - 1) / TY;
    const int gz = (nz + 2 * h + zc - 1) / zc;
    *sa = *sb = 0.0;
    for (int bz = 0; bz < gz; bz++)
        for (int by = 0; by < gy; by++)
            for (int bx = 0; bx < gx; bx++) {
                double x = 0, y = 0;
                Tile::emulate_block(a, lds.data(), th.data(), bx, by, bz, x, y);
                *sa += x;
                *sb += y;
            }
}

extern "C" int hsmac_sweep2_emul(int ty, int tx, int esize, int coloring, int nx, int ny, int nz, int h,
                                 const int *loA, const int *hiA, const int *loB, const int *hiB, double dt,
                                 double omega, void *const *src, void *const *dst, int zc, double *sa, double *sb)
{
#define RUN(T, C, Y, X) run<T, C, Y, X>(nx, ny, nz, h, loA, hiA, loB, hiB, dt, omega, src, dst, zc, sa, sb)
#define BY_SHAPE(T, C)                                  \
    if (ty == 24 && tx == 32) RUN(T, C, 24, 32);        \
    else if (ty == 12 && tx == 64) RUN(T, C, 12, 64);   \
    else if (ty == 16 && tx == 32) RUN(T, C, 16, 32);   \
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
