// Tiled RHS (+ RK update) of src/flows/rhs.jl:59-291 and src/flows/update.jl:255-268.
//
// A workgroup owns a TX x TY column of interior cells and a chunk of z-planes
// and marches up in z.  U, V, W live in 7-plane LDS rings (planes z-3..z+3)
// over the tile plus a 3-cell in-plane halo.  Per output plane z:
//   fetch        global loads of plane z+4 into registers (latency hidden
//                behind the two phases below);
//   phase B      every in-plane flux J1/J3 of the three components (x-fluxes
//                at x in [X0-2, X0+TX], y-fluxes at y in [Y0-2, Y0+TY]) is
//                computed once into LDS;
//   phase C      each thread forms its cell's z-fluxes at z+1 (register
//                queues hold J1 at z-1..z+1 and J3 at z-2..z+1), the viscous
//                and pressure terms and the flux divergence, writes the
//                result, and stores the fetched plane z+4 into the ring slot
//                of plane z-3, which neither phase reads.
// Two barriers per plane: after B (J arrays complete) and after C (J arrays
// free, plane z+4 visible).  The arithmetic of each term is that of the
// per-cell kernel, operand for operand.
//
// Everything here is __host__ __device__ so that a host harness can execute
// the same code thread by thread, phase by phase, against the oracle.
#pragma once

#include "lcs_internal.h"

namespace lcs {

template <typename T>
struct RhsArgs {
    Geom g;
    const T *U, *V, *W, *P;  // velocity of the stage + pressure
    T *out[3];               // fused: updated velocity; else dU/dt (interior layout)
    const T *base[3];        // fused: velocity at the start of the step
    T di[3], di2[3], Rei;
    int zc;  // planes per chunk
};

template <typename T, int TX_, int TY_>
struct RhsTile {
    static constexpr int TX = TX_, TY = TY_, NT = TX * TY;
    static constexpr int LX = TX + 6, LY = TY + 6, NR = 7;
    static constexpr int FX = TX + 3, FY = TY + 3;  // x-flux columns, y-flux rows
    static constexpr int PLANE = LY * LX;
    static constexpr int RING = 3 * NR * PLANE;
    static constexpr int JX = 6 * TY * FX;  // [2c + (0: J1, 1: J3)][ty][fx]
    static constexpr int JY = 6 * FY * TX;  // [2c + ...][fy][tx]
    static constexpr int LDS_ELEMS = RING + JX + JY;
    static constexpr int PF = (PLANE + NT - 1) / NT;  // ring elements fetched per thread and array

    struct Thread {
        T j1[3][3];   // J1 z-flux of component c at z-1, z, z+1
        T j3[3][4];   // J3 z-flux at z-2 .. z+1
        T pf[3][PF];  // fetched plane
    };

    __host__ __device__ static int slot(int z) { return ((z % NR) + NR) % NR; }
    __host__ __device__ static T &R(T *lds, int a, int z, int ly, int lx)
    {
        return lds[(a * NR + slot(z)) * PLANE + ly * LX + lx];
    }
    __host__ __device__ static T &JXr(T *lds, int q, int ty, int fx) { return lds[RING + (q * TY + ty) * FX + fx]; }
    __host__ __device__ static T &JYr(T *lds, int q, int fy, int tx)
    {
        return lds[RING + JX + (q * FY + fy) * TX + tx];
    }

    __host__ __device__ static T i2(T a, T b) { return T(0.5) * (a + b); }
    __host__ __device__ static T i4(T a, T b, T c, T d) { return T(1.0 / 16.0) * (((-a + T(9) * b) + T(9) * c) - d); }
    __host__ __device__ static T d2(T a, T b, T c, T d, T e)
    {
        return T(1.0 / 12.0) * ((((-a + T(16) * b) - T(30) * c) + T(16) * d) - e);
    }
    __host__ __device__ static T sf(T j1a, T j1b, T j3a, T j3b)
    {
        return T(9.0 / 8.0) * (-j1a + j1b) - T((1.0 / 8.0) * (1.0 / 3.0)) * (-j3a + j3b);
    }

    // global loads of plane z of U, V, W over the tile + halo (zero outside the array)
    __host__ __device__ static void fetch(const RhsArgs<T> &a, int X0, int Y0, int z, int tid, Thread &th)
    {
        const Geom &g = a.g;
        const bool zin = z >= 1 - g.h && z <= g.nz + g.h;
#pragma unroll
        for (int r = 0; r < PF; r++) {
            const int e = tid + r * NT;
            const int lx = e % LX, ly = e / LX, x = X0 - 3 + lx, y = Y0 - 3 + ly;
            const bool in = e < PLANE && zin && x >= 1 - g.h && x <= g.nx + g.h && y >= 1 - g.h && y <= g.ny + g.h;
            const long p = in ? g.at(x, y, z) : 0;
            th.pf[0][r] = in ? a.U[p] : T(0);
            th.pf[1][r] = in ? a.V[p] : T(0);
            th.pf[2][r] = in ? a.W[p] : T(0);
        }
    }

    __host__ __device__ static void store(T *lds, int z, int tid, const Thread &th)
    {
        T *s = lds + slot(z) * PLANE;
#pragma unroll
        for (int r = 0; r < PF; r++) {
            const int e = tid + r * NT;
            if (e < PLANE) {
                s[e] = th.pf[0][r];
                s[NR * PLANE + e] = th.pf[1][r];
                s[2 * NR * PLANE + e] = th.pf[2][r];
            }
        }
    }

    // z-fluxes of the three components at plane q for ring column (lx, ly):
    // A = W interpolated along c, Q = component c interpolated along z
    __host__ __device__ static void zflux(T *lds, int q, int lx, int ly, T j1[3], T j3[3])
    {
        const T ac[3] = {
            i4(R(lds, 2, q, ly, lx - 1), R(lds, 2, q, ly, lx), R(lds, 2, q, ly, lx + 1), R(lds, 2, q, ly, lx + 2)),
            i4(R(lds, 2, q, ly - 1, lx), R(lds, 2, q, ly, lx), R(lds, 2, q, ly + 1, lx), R(lds, 2, q, ly + 2, lx)),
            i4(R(lds, 2, q - 1, ly, lx), R(lds, 2, q, ly, lx), R(lds, 2, q + 1, ly, lx), R(lds, 2, q + 2, ly, lx))};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            j1[c] = ac[c] * i2(R(lds, c, q, ly, lx), R(lds, c, q + 1, ly, lx));
            j3[c] = ac[c] * i2(R(lds, c, q - 1, ly, lx), R(lds, c, q + 2, ly, lx));
        }
    }

    // phase B: in-plane fluxes at plane z (reads planes z-1..z+2)
    __host__ __device__ static void phase_fluxes(T *lds, int z, int tid)
    {
        // x-fluxes: A = U interpolated along c, Q = component c along x
        for (int e = tid; e < FX * TY; e += NT) {
            const int fx = e % FX, ty = e / FX;
            const int lx = fx + 1, ly = ty + 3;
            const T ac[3] = {
                i4(R(lds, 0, z, ly, lx - 1), R(lds, 0, z, ly, lx), R(lds, 0, z, ly, lx + 1), R(lds, 0, z, ly, lx + 2)),
                i4(R(lds, 0, z, ly - 1, lx), R(lds, 0, z, ly, lx), R(lds, 0, z, ly + 1, lx), R(lds, 0, z, ly + 2, lx)),
                i4(R(lds, 0, z - 1, ly, lx), R(lds, 0, z, ly, lx), R(lds, 0, z + 1, ly, lx), R(lds, 0, z + 2, ly, lx))};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                JXr(lds, 2 * c, ty, fx) = ac[c] * i2(R(lds, c, z, ly, lx), R(lds, c, z, ly, lx + 1));
                JXr(lds, 2 * c + 1, ty, fx) = ac[c] * i2(R(lds, c, z, ly, lx - 1), R(lds, c, z, ly, lx + 2));
            }
        }
        // y-fluxes: A = V interpolated along c, Q = component c along y
        for (int e = tid; e < TX * FY; e += NT) {
            const int tx = e % TX, fy = e / TX;
            const int lx = tx + 3, ly = fy + 1;
            const T ac[3] = {
                i4(R(lds, 1, z, ly, lx - 1), R(lds, 1, z, ly, lx), R(lds, 1, z, ly, lx + 1), R(lds, 1, z, ly, lx + 2)),
                i4(R(lds, 1, z, ly - 1, lx), R(lds, 1, z, ly, lx), R(lds, 1, z, ly + 1, lx), R(lds, 1, z, ly + 2, lx)),
                i4(R(lds, 1, z - 1, ly, lx), R(lds, 1, z, ly, lx), R(lds, 1, z + 1, ly, lx), R(lds, 1, z + 2, ly, lx))};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                JYr(lds, 2 * c, fy, tx) = ac[c] * i2(R(lds, c, z, ly, lx), R(lds, c, z, ly + 1, lx));
                JYr(lds, 2 * c + 1, fy, tx) = ac[c] * i2(R(lds, c, z, ly - 1, lx), R(lds, c, z, ly + 2, lx));
            }
        }
    }

    // z-flux queues before the first output plane z0 (reads planes z0-3..z0+2)
    __host__ __device__ static void warmup(T *lds, int z0, int tx, int ty, Thread &th)
    {
        const int lx = tx + 3, ly = ty + 3;
#pragma unroll
        for (int s = 0; s < 3; s++) {  // positions z0-2, z0-1, z0
            T j1[3], j3[3];
            zflux(lds, z0 - 2 + s, lx, ly, j1, j3);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                th.j3[c][s] = j3[c];
                if (s >= 1) th.j1[c][s - 1] = j1[c];
            }
        }
    }

    // phase C: result at plane z for cell (X0+tx, Y0+ty) (reads planes z-2..z+3)
    template <bool FUSED>
    __host__ __device__ static void phase_output(const RhsArgs<T> &a, T *lds, int X0, int Y0, int z, int tx, int ty,
                                                 T dt, Thread &th)
    {
        const Geom &g = a.g;
        const int lx = tx + 3, ly = ty + 3;
        {
            T j1[3], j3[3];
            zflux(lds, z + 1, lx, ly, j1, j3);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                th.j1[c][2] = j1[c];
                th.j3[c][3] = j3[c];
            }
        }
        const int x = X0 + tx, y = Y0 + ty;
        if (x <= g.nx && y <= g.ny) {
            const long p = g.at(x, y, z);
            const long st[3] = {1, g.sx, g.sx * g.sy};
            const T P0 = a.P[p];
            const int fx = tx + 2, fy = ty + 2;  // flux slots of positions x and y
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const T dx2 = d2(R(lds, c, z, ly, lx - 2), R(lds, c, z, ly, lx - 1), R(lds, c, z, ly, lx),
                                 R(lds, c, z, ly, lx + 1), R(lds, c, z, ly, lx + 2));
                const T dy2 = d2(R(lds, c, z, ly - 2, lx), R(lds, c, z, ly - 1, lx), R(lds, c, z, ly, lx),
                                 R(lds, c, z, ly + 1, lx), R(lds, c, z, ly + 2, lx));
                const T dz2 = d2(R(lds, c, z - 2, ly, lx), R(lds, c, z - 1, ly, lx), R(lds, c, z, ly, lx),
                                 R(lds, c, z + 1, ly, lx), R(lds, c, z + 2, ly, lx));
                const T diff = ((a.di2[0] * dx2 + a.di2[1] * dy2) + a.di2[2] * dz2) * a.Rei;
                const T pres = a.di[c] * (-P0 + a.P[p + st[c]]);
                const T fxs = sf(JXr(lds, 2 * c, ty, fx - 1), JXr(lds, 2 * c, ty, fx),
                                 JXr(lds, 2 * c + 1, ty, fx - 2), JXr(lds, 2 * c + 1, ty, fx + 1));
                const T fys = sf(JYr(lds, 2 * c, fy - 1, tx), JYr(lds, 2 * c, fy, tx),
                                 JYr(lds, 2 * c + 1, fy - 2, tx), JYr(lds, 2 * c + 1, fy + 1, tx));
                const T fzs = sf(th.j1[c][0], th.j1[c][1], th.j3[c][0], th.j3[c][3]);
                const T advect = (a.di[0] * fxs + a.di[1] * fys) + a.di[2] * fzs;
                const T dq = (diff - pres) - advect;
                if (FUSED)
                    a.out[c][p] = a.base[c][p] + dq * dt;
                else
                    a.out[c][g.interior(x, y, z)] = dq;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            th.j1[c][0] = th.j1[c][1];
            th.j1[c][1] = th.j1[c][2];
            th.j3[c][0] = th.j3[c][1];
            th.j3[c][1] = th.j3[c][2];
            th.j3[c][2] = th.j3[c][3];
        }
    }

    // The whole march of one workgroup, for a host harness that runs the
    // threads of each phase one after the other; the kernel body follows the
    // same sequence with a barrier between the phases.
    template <bool FUSED>
    static void emulate_block(const RhsArgs<T> &a, T *lds, Thread *th, int bx, int by, int bz, T dt)
    {
        const int X0 = bx * TX + 1, Y0 = by * TY + 1, z0 = bz * a.zc + 1;
        const int z1 = z0 + a.zc - 1 < a.g.nz ? z0 + a.zc - 1 : a.g.nz;
        for (int q = z0 - 3; q <= z0 + 3; q++)
            for (int t = 0; t < NT; t++) {
                fetch(a, X0, Y0, q, t, th[t]);
                store(lds, q, t, th[t]);
            }
        for (int t = 0; t < NT; t++) warmup(lds, z0, t % TX, t / TX, th[t]);
        for (int z = z0; z <= z1; z++) {
            for (int t = 0; t < NT; t++) {
                if (z < z1) fetch(a, X0, Y0, z + 4, t, th[t]);
                phase_fluxes(lds, z, t);
            }
            for (int t = 0; t < NT; t++) {
                phase_output<FUSED>(a, lds, X0, Y0, z, t % TX, t / TX, dt, th[t]);
                if (z < z1) store(lds, z + 4, t, th[t]);
            }
        }
    }
};

}  // namespace lcs
