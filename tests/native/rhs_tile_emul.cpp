// Host execution of the tiled RHS kernel's phase functions
// (lcs.jl_amd/csrc/lcs_rhs_tile.h): every workgroup of the launch grid is run
// thread by thread, phase by phase, in the kernel's barrier order, so a test
// can compare the kernel's own code with the oracle on the CPU.
#include <vector>

#include "lcs_rhs_tile.h"

using namespace lcs;

template <typename T, int TX, int TY>
static int run(int nx, int ny, int nz, int h, double Re, const void *U, const void *V, const void *W, const void *P,
               void *o0, void *o1, void *o2, const void *b0, const void *b1, const void *b2, int fused, double dt,
               int zc)
{
    using Tile = RhsTile<T, TX, TY>;
    RhsArgs<T> a;
    a.g = make_geom(nx, ny, nz, h);
    a.U = (const T *)U;
    a.V = (const T *)V;
    a.W = (const T *)W;
    a.P = (const T *)P;
    a.out[0] = (T *)o0;
    a.out[1] = (T *)o1;
    a.out[2] = (T *)o2;
    a.base[0] = (const T *)b0;
    a.base[1] = (const T *)b1;
    a.base[2] = (const T *)b2;
    const int n[3] = {nx, ny, nz};
    for (int c = 0; c < 3; c++) {
        a.di[c] = T(1.0 / spacing(n[c]));
        a.di2[c] = a.di[c] * a.di[c];
    }
    a.Rei = T(1.0 / Re);
    a.zc = zc;
    std::vector<T> lds(Tile::LDS_ELEMS);
    std::vector<typename Tile::Thread> th(Tile::NT);
    const int gx = (nx + TX - 1) / TX, gy = (ny + TY - 1) / TY, gz = (nz + zc - 1) / zc;
    for (int bz = 0; bz < gz; bz++)
        for (int by = 0; by < gy; by++)
            for (int bx = 0; bx < gx; bx++) {
                if (fused)
                    Tile::template emulate_block<true>(a, lds.data(), th.data(), bx, by, bz, T(dt));
                else
                    Tile::template emulate_block<false>(a, lds.data(), th.data(), bx, by, bz, T(dt));
            }
    return 0;
}

extern "C" int rhs_tile_emul(int shape, int esize, int nx, int ny, int nz, int h, double Re, const void *U,
                             const void *V, const void *W, const void *P, void *o0, void *o1, void *o2,
                             const void *b0, const void *b1, const void *b2, int fused, double dt, int zc)
{
#define RUN(T, X, Y) run<T, X, Y>(nx, ny, nz, h, Re, U, V, W, P, o0, o1, o2, b0, b1, b2, fused, dt, zc)
    if (esize == 4) {
        switch (shape) {
        case 0: return RUN(float, 64, 4);
        case 1: return RUN(float, 8, 4);
        case 2: return RUN(float, 16, 2);
        case 3: return RUN(float, 64, 16);
        case 4: return RUN(float, 32, 8);
        }
    } else {
        switch (shape) {
        case 0: return RUN(double, 64, 4);
        case 1: return RUN(double, 8, 4);
        case 2: return RUN(double, 16, 2);
        case 3: return RUN(double, 64, 16);
        case 4: return RUN(double, 32, 8);
        }
    }
#undef RUN
    return -1;
}
