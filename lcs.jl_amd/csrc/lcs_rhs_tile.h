// Tiled RHS (+ RK update) of src/flows/rhs.jl:44-252 and src/flows/update.jl:37-51.
//
// A workgroup owns a TX x TY column of interior cells and a chunk of z-planes
// and marches up in z.  Each thread keeps its own column's U, V, W for planes
// z-2..z+3 in registers (every z-direction stencil reads only those); LDS holds
// the planes in-plane neighbours need, over the tile plus a 3-cell halo:
// U and V for z-1..z+2 (the in-plane fluxes interpolate them along z), W for
// z and z+1 (plus the slot being refilled).  Per output plane z:
//   fetch    global loads for the next plane: own column at z+4, the halo of
//            U, V at z+3 and of W at z+2 (latency hidden behind the phases);
//   phase B  every in-plane flux J1/J3 of the three components (x-fluxes at
//            x in [X0-2, X0+TX], y-fluxes at y in [Y0-2, Y0+TY]) is computed
//            once into LDS;
//   phase C  each thread forms its cell's z-fluxes at z+1 (register queues
//            hold J1 at z-1..z+1 and J3 at z-2..z+1), the viscous and
//            pressure terms and the flux divergence, writes the result, and
//            stores U, V of plane z+3 and W of plane z+2 into the ring slots
//            neither phase reads.
// Two barriers per plane: after B (J arrays complete) and after C (J arrays
// free, new planes visible).  The arithmetic of each term is that of
// src/flows/rhs.jl, operand for operand.
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
    const T *base[3];        // fused: velocity at the start of the step (null: the stage input itself)
    T di[3], di2[3], Rei;
    int zc;  // planes per chunk
};

template <typename T, int TX_, int TY_>
struct RhsTile {
    static constexpr int TX = TX_, TY = TY_, NT = TX * TY;
    static constexpr int LX = TX + 6, LY = TY + 6, PLANE = LY * LX;
    static constexpr int FX = TX + 3, FY = TY + 3;  // x-flux columns, y-flux rows
    static constexpr int OFF_V = 4 * PLANE, OFF_W = 8 * PLANE, RING = 11 * PLANE;
    static constexpr int JX = 6 * TY * FX;  // [2c + (0: J1, 1: J3)][ty][fx]
    static constexpr int JY = 6 * FY * TX;  // [2c + ...][fy][tx]
    static constexpr int LDS_ELEMS = RING + JX + JY;
    static constexpr int NH = (PLANE - NT + NT - 1) / NT;  // halo elements per thread

    struct Thread {
        int tid, X0, Y0, z0, z1;
        long o_off;      // in-plane offset of the own column, -1 outside the array
        bool o_in;       // own column is an interior cell of the box
        int h_lds[NH];   // LDS offsets of this thread's halo elements, -1 if none
        long h_off[NH];  // their in-plane offsets, -1 outside the array
        T own[3][6];     // own column, planes z-2..z+3
        T fo[3];         // own column, plane z+4 (fetched)
        T fh[3][NH];     // halo: U, V at z+3, W at z+2 (fetched)
        T j1[3][3];      // J1 z-flux of component c at z-1, z, z+1
        T j3[3][4];      // J3 z-flux at z-2 .. z+1
    };

    __host__ __device__ static int m4(int z) { return z & 3; }
    __host__ __device__ static int m3(int z) { return ((z % 3) + 3) % 3; }
    // field a (0 U, 1 V, 2 W) at plane z, ring coordinates (ly, lx)
    __host__ __device__ static T &R(T *lds, int a, int z, int ly, int lx)
    {
        const int s = a == 2 ? OFF_W + m3(z) * PLANE : a * OFF_V + m4(z) * PLANE;
        return lds[s + ly * LX + lx];
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

    __host__ __device__ static long plane_base(const Geom &g, int z) { return (long)(z - 1 + g.h) * g.sx * g.sy; }
    __host__ __device__ static bool plane_in(const Geom &g, int z) { return z >= 1 - g.h && z <= g.nz + g.h; }
    __host__ __device__ static long in_plane(const Geom &g, int x, int y)
    {
        const bool ok = x >= 1 - g.h && x <= g.nx + g.h && y >= 1 - g.h && y <= g.ny + g.h;
        return ok ? (long)(x - 1 + g.h) + g.sx * (long)(y - 1 + g.h) : -1;
    }
    __host__ __device__ static T ld(const T *A, const Geom &g, long off, int z)
    {
        return (off >= 0 && plane_in(g, z)) ? A[plane_base(g, z) + off] : T(0);
    }

    __host__ __device__ static void init(const RhsArgs<T> &a, int bx, int by, int bz, int tid, Thread &th)
    {
        const Geom &g = a.g;
        th.tid = tid;
        th.X0 = bx * TX + 1;
        th.Y0 = by * TY + 1;
        th.z0 = bz * a.zc + 1;
        th.z1 = th.z0 + a.zc - 1 < g.nz ? th.z0 + a.zc - 1 : g.nz;
        const int x = th.X0 + tid % TX, y = th.Y0 + tid / TX;
        th.o_in = x <= g.nx && y <= g.ny;
        th.o_off = in_plane(g, x, y);
        // halo elements: ring positions outside the owned TX x TY block, in
        // row-major order
        for (int s = 0; s < NH; s++) {
            th.h_lds[s] = -1;
            th.h_off[s] = -1;
            const int e = tid + s * NT;
            if (e >= PLANE - NT) continue;
            int lx, ly;
            if (e < 3 * LX) {  // rows 0..2
                ly = e / LX;
                lx = e % LX;
            } else if (e < 6 * LX) {  // rows TY+3..TY+5
                ly = TY + 3 + (e - 3 * LX) / LX;
                lx = (e - 3 * LX) % LX;
            } else {  // columns 0..2 and TX+3..TX+5 of rows 3..TY+2
                const int f = e - 6 * LX, c = f % 6;
                ly = 3 + f / 6;
                lx = c < 3 ? c : TX + c;
            }
            th.h_lds[s] = ly * LX + lx;
            th.h_off[s] = in_plane(g, th.X0 - 3 + lx, th.Y0 - 3 + ly);
        }
    }

    __host__ __device__ static int own_lds(const Thread &th) { return (th.tid / TX + 3) * LX + th.tid % TX + 3; }

    // prologue: own column z0-2..z0+3, LDS planes U, V z0-1..z0+2 and W
    // z0..z0+1, and the z-flux queues (from global memory)
    __host__ __device__ static void prologue_fetch(const RhsArgs<T> &a, Thread &th)
    {
        const Geom &g = a.g;
        const T *F[3] = {a.U, a.V, a.W};
        for (int c = 0; c < 3; c++)
            for (int s = 0; s < 6; s++) th.own[c][s] = ld(F[c], g, th.o_off, th.z0 - 2 + s);
    }
    __host__ __device__ static void prologue_store(const RhsArgs<T> &a, T *lds, const Thread &th)
    {
        const Geom &g = a.g;
        const T *F[3] = {a.U, a.V, a.W};
        for (int c = 0; c < 3; c++) {
            const int zlo = c == 2 ? th.z0 : th.z0 - 1, zhi = th.z0 + (c == 2 ? 1 : 2);
            for (int z = zlo; z <= zhi; z++) {
                R(lds, c, z, 0, own_lds(th)) = th.own[c][z - (th.z0 - 2)];
                for (int s = 0; s < NH; s++)
                    if (th.h_lds[s] >= 0) R(lds, c, z, 0, th.h_lds[s]) = ld(F[c], g, th.h_off[s], z);
            }
        }
    }
    __host__ __device__ static void prologue_queues(const RhsArgs<T> &a, Thread &th)
    {
        const Geom &g = a.g;
        const int x = th.X0 + th.tid % TX, y = th.Y0 + th.tid / TX;
        const T *F[3] = {a.U, a.V, a.W};
        for (int s = 0; s < 3; s++) {  // positions z0-2, z0-1, z0
            const int q = th.z0 - 2 + s;
            auto Wg = [&](int dx, int dy, int dz) { return ld(a.W, g, in_plane(g, x + dx, y + dy), q + dz); };
            const T ac[3] = {i4(Wg(-1, 0, 0), Wg(0, 0, 0), Wg(1, 0, 0), Wg(2, 0, 0)),
                             i4(Wg(0, -1, 0), Wg(0, 0, 0), Wg(0, 1, 0), Wg(0, 2, 0)),
                             i4(Wg(0, 0, -1), Wg(0, 0, 0), Wg(0, 0, 1), Wg(0, 0, 2))};
            for (int c = 0; c < 3; c++) {
                const T q1 = i2(ld(F[c], g, th.o_off, q), ld(F[c], g, th.o_off, q + 1));
                const T q3 = i2(ld(F[c], g, th.o_off, q - 1), ld(F[c], g, th.o_off, q + 2));
                th.j3[c][s] = ac[c] * q3;
                if (s >= 1) th.j1[c][s - 1] = ac[c] * q1;
            }
        }
    }

    // global loads for the next plane
    __host__ __device__ static void fetch(const RhsArgs<T> &a, int z, Thread &th)
    {
        const Geom &g = a.g;
        th.fo[0] = ld(a.U, g, th.o_off, z + 4);
        th.fo[1] = ld(a.V, g, th.o_off, z + 4);
        th.fo[2] = ld(a.W, g, th.o_off, z + 4);
        for (int s = 0; s < NH; s++) {
            th.fh[0][s] = ld(a.U, g, th.h_off[s], z + 3);
            th.fh[1][s] = ld(a.V, g, th.h_off[s], z + 3);
            th.fh[2][s] = ld(a.W, g, th.h_off[s], z + 2);
        }
    }

    // U, V of plane z+3 and W of plane z+2 into LDS; own queue advances
    __host__ __device__ static void store(T *lds, int z, Thread &th)
    {
        const int o = own_lds(th);
        R(lds, 0, z + 3, 0, o) = th.own[0][5];
        R(lds, 1, z + 3, 0, o) = th.own[1][5];
        R(lds, 2, z + 2, 0, o) = th.own[2][4];
        for (int s = 0; s < NH; s++) {
            const int h = th.h_lds[s];
            if (h < 0) continue;
            R(lds, 0, z + 3, 0, h) = th.fh[0][s];
            R(lds, 1, z + 3, 0, h) = th.fh[1][s];
            R(lds, 2, z + 2, 0, h) = th.fh[2][s];
        }
        for (int c = 0; c < 3; c++) {
            for (int s = 0; s < 5; s++) th.own[c][s] = th.own[c][s + 1];
            th.own[c][5] = th.fo[c];
        }
    }

    // phase B: in-plane fluxes at plane z (reads U, V at z-1..z+2, W at z).
    // Each item reads every ring value it needs before it writes a flux: the
    // ring and the flux arrays never overlap, and with the reads first the
    // compiler issues them together instead of waiting on LDS after each
    // write.  The arithmetic is unchanged, operand for operand.
    __host__ __device__ static void phase_fluxes(T *lds, int z, int tid)
    {
        // x-fluxes: A = U interpolated along c, Q = component c along x
        for (int e = tid; e < FX * TY; e += NT) {
            const int fx = e % FX, ty = e / FX;
            const int lx = fx + 1, ly = ty + 3;
            T q[3][4];  // component c at x = lx-1 .. lx+2
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int k = 0; k < 4; k++) q[c][k] = R(lds, c, z, ly, lx - 1 + k);
            const T uy[3] = {R(lds, 0, z, ly - 1, lx), R(lds, 0, z, ly + 1, lx), R(lds, 0, z, ly + 2, lx)};
            const T uz[3] = {R(lds, 0, z - 1, ly, lx), R(lds, 0, z + 1, ly, lx), R(lds, 0, z + 2, ly, lx)};
            const T ac[3] = {i4(q[0][0], q[0][1], q[0][2], q[0][3]), i4(uy[0], q[0][1], uy[1], uy[2]),
                             i4(uz[0], q[0][1], uz[1], uz[2])};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                JXr(lds, 2 * c, ty, fx) = ac[c] * i2(q[c][1], q[c][2]);
                JXr(lds, 2 * c + 1, ty, fx) = ac[c] * i2(q[c][0], q[c][3]);
            }
        }
        // y-fluxes: A = V interpolated along c, Q = component c along y
        for (int e = tid; e < TX * FY; e += NT) {
            const int tx = e % TX, fy = e / TX;
            const int lx = tx + 3, ly = fy + 1;
            T q[3][4];  // component c at y = ly-1 .. ly+2
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int k = 0; k < 4; k++) q[c][k] = R(lds, c, z, ly - 1 + k, lx);
            const T vx[3] = {R(lds, 1, z, ly, lx - 1), R(lds, 1, z, ly, lx + 1), R(lds, 1, z, ly, lx + 2)};
            const T vz[3] = {R(lds, 1, z - 1, ly, lx), R(lds, 1, z + 1, ly, lx), R(lds, 1, z + 2, ly, lx)};
            const T ac[3] = {i4(vx[0], q[1][1], vx[1], vx[2]), i4(q[1][0], q[1][1], q[1][2], q[1][3]),
                             i4(vz[0], q[1][1], vz[1], vz[2])};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                JYr(lds, 2 * c, fy, tx) = ac[c] * i2(q[c][1], q[c][2]);
                JYr(lds, 2 * c + 1, fy, tx) = ac[c] * i2(q[c][0], q[c][3]);
            }
        }
    }

    // phase C: result at plane z for the own cell (reads U, V, W at z and W
    // at z+1 from LDS, the own column from registers)
    template <bool FUSED>
    __host__ __device__ static void phase_output(const RhsArgs<T> &a, T *lds, int z, T dt, Thread &th)
    {
        const Geom &g = a.g;
        const int tx = th.tid % TX, ty = th.tid / TX, lx = tx + 3, ly = ty + 3;
        // z-fluxes at z+1: A = W along c (in-plane from LDS, along z from the
        // own column), Q = component c along z (own column)
        {
            const T ac[3] = {
                i4(R(lds, 2, z + 1, ly, lx - 1), th.own[2][3], R(lds, 2, z + 1, ly, lx + 1), R(lds, 2, z + 1, ly, lx + 2)),
                i4(R(lds, 2, z + 1, ly - 1, lx), th.own[2][3], R(lds, 2, z + 1, ly + 1, lx), R(lds, 2, z + 1, ly + 2, lx)),
                i4(th.own[2][2], th.own[2][3], th.own[2][4], th.own[2][5])};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                th.j1[c][2] = ac[c] * i2(th.own[c][3], th.own[c][4]);
                th.j3[c][3] = ac[c] * i2(th.own[c][2], th.own[c][5]);
            }
        }
        if (th.o_in) {
            const long p = plane_base(g, z) + th.o_off;
            const long st[3] = {1, g.sx, g.sx * g.sy};
            const T P0 = a.P[p];
            const int fx = tx + 2, fy = ty + 2;  // flux slots of positions x and y
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const T ctr = th.own[c][2];
                const T dx2 = d2(R(lds, c, z, ly, lx - 2), R(lds, c, z, ly, lx - 1), ctr, R(lds, c, z, ly, lx + 1),
                                 R(lds, c, z, ly, lx + 2));
                const T dy2 = d2(R(lds, c, z, ly - 2, lx), R(lds, c, z, ly - 1, lx), ctr, R(lds, c, z, ly + 1, lx),
                                 R(lds, c, z, ly + 2, lx));
                const T dz2 = d2(th.own[c][0], th.own[c][1], ctr, th.own[c][3], th.own[c][4]);
                const T diff = ((a.di2[0] * dx2 + a.di2[1] * dy2) + a.di2[2] * dz2) * a.Rei;
                const T pres = a.di[c] * (-P0 + a.P[p + st[c]]);
                const T fxs = sf(JXr(lds, 2 * c, ty, fx - 1), JXr(lds, 2 * c, ty, fx),
                                 JXr(lds, 2 * c + 1, ty, fx - 2), JXr(lds, 2 * c + 1, ty, fx + 1));
                const T fys = sf(JYr(lds, 2 * c, fy - 1, tx), JYr(lds, 2 * c, fy, tx),
                                 JYr(lds, 2 * c + 1, fy - 2, tx), JYr(lds, 2 * c + 1, fy + 1, tx));
                const T fzs = sf(th.j1[c][0], th.j1[c][1], th.j3[c][0], th.j3[c][3]);
                const T advect = (a.di[0] * fxs + a.di[1] * fys) + a.di[2] * fzs;
                const T dq = (diff - pres) - advect;
                if (FUSED)  // stage 1 updates its own input: the base is the register copy
                    a.out[c][p] = (a.base[c] ? a.base[c][p] : ctr) + dq * dt;
                else
                    a.out[c][g.interior(th.X0 + tx, th.Y0 + ty, z)] = dq;
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

    // The march of one workgroup, for a host harness that runs the threads of
    // each phase one after the other; the kernel body follows the same
    // sequence with a barrier between the phases.
    template <bool FUSED>
    static void emulate_block(const RhsArgs<T> &a, T *lds, Thread *th, int bx, int by, int bz, T dt)
    {
        for (int t = 0; t < NT; t++) {
            init(a, bx, by, bz, t, th[t]);
            prologue_fetch(a, th[t]);
            prologue_store(a, lds, th[t]);
            prologue_queues(a, th[t]);
        }
        const int z0 = th[0].z0, z1 = th[0].z1;
        for (int z = z0; z <= z1; z++) {
            for (int t = 0; t < NT; t++) {
                if (z < z1) fetch(a, z, th[t]);
                phase_fluxes(lds, z, t);
            }
            for (int t = 0; t < NT; t++) {
                phase_output<FUSED>(a, lds, z, dt, th[t]);
                if (z < z1) store(lds, z, th[t]);
            }
        }
    }
};

}  // namespace lcs
