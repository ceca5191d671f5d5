// One fused red+black HSMAC sweep (src/flows/couples/hsmac.jl:1-35 over a
// box R, colour 0 then colour 1 as lib/parallel/src/foraxes.jl:73-77 orders
// them), read from one buffer set and written to the other.
//
// Red corrections depend only on the sweep's input; black ones on the input
// plus the red corrections of their six neighbours.  So the final value of a
// face -- input, then the red correction, then the black one, in the order the
// two in-place passes apply them -- is formed by exactly one thread, and the
// whole sweep costs one read and one write of (U, V, W, P).
//
// A workgroup owns a 64 x TY column of cells and a chunk of planes [z0, z1]
// and marches up in z.  At step q:
//   phase A  red dp on plane q over x in [X0-1, X1+2], y in [Y0-1, Y1+2]
//            (the red neighbours of the black cells below);
//   phase B  black dp on plane q-1 over x in [X0, X1+1], y in [Y0, Y1+1];
//            final U, V, W, P of plane q-3 for the owned cells; the
//            prefetched plane q+1 goes into the ring slot of plane q-4.
// LDS holds U, V, W for planes q-3..q+1 (ring of 5) over x in [X0-2, X1+2],
// y in [Y0-2, Y1+2], red dp for q-3..q (ring of 4), black dp for q-3..q-1
// (ring of 3).  Two barriers per plane.  Every cell (x, y, z) owned by the
// workgroup writes P and its upper faces U(x), V(y), W(z); the grid covers
// the whole array, so elements outside R are carried over to the other set.
//
// All of it is __host__ __device__ so a host harness can run the phases
// thread by thread against the oracle.
#pragma once

#include "lcs_internal.h"

namespace lcs {

template <typename T>
struct HsCoef {
    T dxi, dyi, dzi, dl, omega;
};

struct Box3 {
    int lo[3], hi[3];
};

__host__ __device__ inline bool in_box(const Box3 &r, int x, int y, int z)
{
    return x >= r.lo[0] && x <= r.hi[0] && y >= r.lo[1] && y <= r.hi[1] && z >= r.lo[2] && z <= r.hi[2];
}

// red-black (truncating %) never visits cells with a negative odd index sum
// (lib/parallel/src/foraxes.jl:84-88)
template <int COLORING>
__host__ __device__ inline bool visited(const Box3 &r, int x, int y, int z)
{
    if (!in_box(r, x, y, z)) return false;
    if (COLORING == LCS_COLORING_RED_BLACK) {
        int s = x + y + z;
        if (s < 0 && (s & 1)) return false;
    }
    return true;
}

// cells whose divergence enters the RMS test: (1:n+1)^3 (hsmac.jl:97-100)
__host__ __device__ inline bool in_d3(const Geom &g, int x, int y, int z)
{
    return x >= 1 && x <= g.nx + 1 && y >= 1 && y <= g.ny + 1 && z >= 1 && z <= g.nz + 1;
}

template <typename T>
struct HsArgs {
    Geom g;
    const T *Us, *Vs, *Ws, *Ps;  // sweep input
    T *Ud, *Vd, *Wd, *Pd;        // sweep output
    Box3 R;
    HsCoef<T> c;
    T dti;
    int zc;
};

template <typename T, int COLORING, int TY_>
struct HsTile {
    static constexpr int TX = 64, TY = TY_, NT = TX * TY;
    static constexpr int LX = TX + 4, LY = TY + 4, PLANE = LX * LY;  // fields: x in [X0-2, X1+2]
    static constexpr int RXN = TX + 3, RYN = TY + 3, RPL = RXN * RYN;  // red dp: x in [X0-1, X1+2]
    static constexpr int BXN = TX + 1, BYN = TY + 1, BPL = BXN * BYN;  // black dp: x in [X0, X1+1]
    static constexpr int NF = 5, NRS = 4, NBS = 3;                     // ring depths
    static constexpr int OFF_R = 3 * NF * PLANE, OFF_B = OFF_R + NRS * RPL;
    static constexpr int LDS_ELEMS = OFF_B + NBS * BPL;
    static constexpr int NLD = (PLANE + NT - 1) / NT;
    static constexpr int RPR = (RXN + 1) / 2, BPR = (BXN + 1) / 2;  // cells of one colour per row
    static_assert(RPR * RYN <= NT && BPR * BYN <= NT, "one red and one black slot per thread");

    struct Thread {
        int tid, X0, Y0, z0, z1;
        int ld_lds[NLD];
        long ld_off[NLD];  // in-plane offset, -1 outside the array
        int r_row, r_k, b_row, b_k;
        long o_off;  // in-plane offset of the owned cell, -1 outside the array
        T pf[3][NLD];
        T pP;
        double acc;
    };

    __host__ __device__ static int m5(int q) { return ((q % 5) + 5) % 5; }
    __host__ __device__ static int m4(int q) { return q & 3; }
    __host__ __device__ static int m3(int q) { return ((q % 3) + 3) % 3; }
    __host__ __device__ static T &F(T *lds, int f, int q, int ly, int lx)
    {
        return lds[(f * NF + m5(q)) * PLANE + ly * LX + lx];
    }
    __host__ __device__ static T &Rd(T *lds, int q, int ry, int rx) { return lds[OFF_R + m4(q) * RPL + ry * RXN + rx]; }
    __host__ __device__ static T &Bd(T *lds, int q, int by, int bx) { return lds[OFF_B + m3(q) * BPL + by * BXN + bx]; }

    __host__ __device__ static long plane_base(const Geom &g, int q) { return (long)(q - 1 + g.h) * g.sx * g.sy; }
    __host__ __device__ static bool plane_in(const Geom &g, int q) { return q >= 1 - g.h && q <= g.nz + g.h; }

    __host__ __device__ static void init(const HsArgs<T> &a, int bx, int by, int bz, int tid, Thread &th)
    {
        const Geom &g = a.g;
        th.tid = tid;
        th.X0 = 1 - g.h + TX * bx;
        th.Y0 = 1 - g.h + TY * by;
        th.z0 = 1 - g.h + a.zc * bz;
        th.z1 = th.z0 + a.zc - 1 < g.nz + g.h ? th.z0 + a.zc - 1 : g.nz + g.h;
        for (int r = 0; r < NLD; r++) {
            const int e = tid + r * NT;
            th.ld_lds[r] = e < PLANE ? e : -1;
            const int x = th.X0 - 2 + e % LX, y = th.Y0 - 2 + e / LX;
            const bool ok = e < PLANE && x >= 1 - g.h && x <= g.nx + g.h && y >= 1 - g.h && y <= g.ny + g.h;
            th.ld_off[r] = ok ? (long)(x - 1 + g.h) + g.sx * (long)(y - 1 + g.h) : -1;
        }
        th.r_row = tid / RPR;
        th.r_k = tid % RPR;
        th.b_row = tid / BPR;
        th.b_k = tid % BPR;
        const int x = th.X0 + tid % TX, y = th.Y0 + tid / TX;
        th.o_off = (x <= g.nx + g.h && y <= g.ny + g.h) ? (long)(x - 1 + g.h) + g.sx * (long)(y - 1 + g.h) : -1;
        th.pP = T(0);
        th.acc = 0.0;
    }

    __host__ __device__ static void fetch(const HsArgs<T> &a, int q, Thread &th)
    {
        const bool zin = plane_in(a.g, q);
        const long base = plane_base(a.g, q);
        for (int r = 0; r < NLD; r++) {
            const bool ok = zin && th.ld_off[r] >= 0;
            const long p = ok ? base + th.ld_off[r] : 0;
            th.pf[0][r] = ok ? a.Us[p] : T(0);
            th.pf[1][r] = ok ? a.Vs[p] : T(0);
            th.pf[2][r] = ok ? a.Ws[p] : T(0);
        }
    }

    __host__ __device__ static void store(T *lds, int q, const Thread &th)
    {
        for (int r = 0; r < NLD; r++) {
            const int e = th.ld_lds[r];
            if (e < 0) continue;
            for (int f = 0; f < 3; f++) lds[(f * NF + m5(q)) * PLANE + e] = th.pf[f][r];
        }
    }

    __host__ __device__ static void fetch_p(const HsArgs<T> &a, int o, Thread &th)
    {
        if (th.o_off >= 0) th.pP = a.Ps[plane_base(a.g, o) + th.o_off];
    }

    __host__ __device__ static bool owned(const Thread &th, int x, int y, int z)
    {
        return x >= th.X0 && x < th.X0 + TX && y >= th.Y0 && y < th.Y0 + TY && z >= th.z0 && z <= th.z1;
    }

    // phase A: red dp on plane q
    __host__ __device__ static void red(const HsArgs<T> &a, T *lds, int q, Thread &th)
    {
        if (th.r_row >= RYN) return;
        const int xs = th.X0 - 1, y = th.Y0 - 1 + th.r_row;
        const int rx = ((xs + y + q) & 1) + 2 * th.r_k;
        if (rx >= RXN) return;
        const int x = xs + rx, ry = th.r_row, lx = rx + 1, ly = ry + 1;
        const HsCoef<T> &c = a.c;
        T dp = T(0);
        if (visited<COLORING>(a.R, x, y, q)) {
            const T dudx = c.dxi * (-F(lds, 0, q, ly, lx - 1) + F(lds, 0, q, ly, lx));
            const T dvdy = c.dyi * (-F(lds, 1, q, ly - 1, lx) + F(lds, 1, q, ly, lx));
            const T dwdz = c.dzi * (-F(lds, 2, q - 1, ly, lx) + F(lds, 2, q, ly, lx));
            const T div = (dudx + dvdy) + dwdz;
            dp = (-c.omega * div) * c.dl;
            if (owned(th, x, y, q) && in_d3(a.g, x, y, q)) th.acc += (double)(div * div);
        }
        Rd(lds, q, ry, rx) = dp;
    }

    // phase B, first part: black dp on plane p, on faces already corrected by
    // their red neighbours
    __host__ __device__ static void black(const HsArgs<T> &a, T *lds, int p, Thread &th)
    {
        if (th.b_row >= BYN) return;
        const int y = th.Y0 + th.b_row;
        const int bx = ((th.X0 + y + p + 1) & 1) + 2 * th.b_k;
        if (bx >= BXN) return;
        const int x = th.X0 + bx, by = th.b_row, rx = bx + 1, ry = by + 1, lx = bx + 2, ly = by + 2;
        const HsCoef<T> &c = a.c;
        T dp = T(0);
        if (visited<COLORING>(a.R, x, y, p)) {
            const T Ur = F(lds, 0, p, ly, lx) - Rd(lds, p, ry, rx + 1) * c.dxi;
            const T Ul = F(lds, 0, p, ly, lx - 1) + Rd(lds, p, ry, rx - 1) * c.dxi;
            const T Vr = F(lds, 1, p, ly, lx) - Rd(lds, p, ry + 1, rx) * c.dyi;
            const T Vl = F(lds, 1, p, ly - 1, lx) + Rd(lds, p, ry - 1, rx) * c.dyi;
            const T Wt = F(lds, 2, p, ly, lx) - Rd(lds, p + 1, ry, rx) * c.dzi;
            const T Wb = F(lds, 2, p - 1, ly, lx) + Rd(lds, p - 1, ry, rx) * c.dzi;
            const T dudx = c.dxi * (-Ul + Ur);
            const T dvdy = c.dyi * (-Vl + Vr);
            const T dwdz = c.dzi * (-Wb + Wt);
            const T div = (dudx + dvdy) + dwdz;
            dp = (-c.omega * div) * c.dl;
            if (owned(th, x, y, p) && in_d3(a.g, x, y, p)) th.acc += (double)(div * div);
        }
        Bd(lds, p, by, bx) = dp;
    }

    // phase B, second part: final values of the owned cell on plane o
    __host__ __device__ static void output(const HsArgs<T> &a, T *lds, int o, Thread &th)
    {
        if (th.o_off < 0) return;
        const int tx = th.tid % TX, ty = th.tid / TX;
        const int x = th.X0 + tx, y = th.Y0 + ty;
        const int rx = tx + 1, ry = ty + 1, lx = tx + 2, ly = ty + 2;
        const HsCoef<T> &c = a.c;
        const bool isred = ((x + y + o) & 1) == 0;
        // correction of this cell and of its +x, +y, +z neighbours (0 outside R)
        const T d0 = isred ? Rd(lds, o, ry, rx) : Bd(lds, o, ty, tx);
        const T dx1 = isred ? Bd(lds, o, ty, tx + 1) : Rd(lds, o, ry, rx + 1);
        const T dy1 = isred ? Bd(lds, o, ty + 1, tx) : Rd(lds, o, ry + 1, rx);
        const T dz1 = isred ? Bd(lds, o + 1, ty, tx) : Rd(lds, o + 1, ry, rx);
        T pv = th.pP;
        if (visited<COLORING>(a.R, x, y, o)) pv += d0 * a.dti;
        T u = F(lds, 0, o, ly, lx), v = F(lds, 1, o, ly, lx), w = F(lds, 2, o, ly, lx);
        if (isred) {
            u += d0 * c.dxi;
            u -= dx1 * c.dxi;
            v += d0 * c.dyi;
            v -= dy1 * c.dyi;
            w += d0 * c.dzi;
            w -= dz1 * c.dzi;
        } else {
            u -= dx1 * c.dxi;
            u += d0 * c.dxi;
            v -= dy1 * c.dyi;
            v += d0 * c.dyi;
            w -= dz1 * c.dzi;
            w += d0 * c.dzi;
        }
        const long gi = plane_base(a.g, o) + th.o_off;
        a.Pd[gi] = pv;
        a.Ud[gi] = u;
        a.Vd[gi] = v;
        a.Wd[gi] = w;
    }

    // The march of one workgroup for a host harness: the threads of each
    // phase run one after the other; the kernel has a barrier where a phase
    // ends.
    static double emulate_block(const HsArgs<T> &a, T *lds, Thread *th, int bx, int by, int bz)
    {
        for (int t = 0; t < NT; t++) init(a, bx, by, bz, t, th[t]);
        const int z0 = th[0].z0, z1 = th[0].z1;
        for (int t = 0; t < NT; t++) {
            fetch(a, z0 - 2, th[t]);
            store(lds, z0 - 2, th[t]);
            fetch(a, z0 - 1, th[t]);
            store(lds, z0 - 1, th[t]);
        }
        for (int q = z0 - 1; q <= z1 + 3; q++) {
            for (int t = 0; t < NT; t++) {
                if (q + 1 <= z1 + 2) fetch(a, q + 1, th[t]);
                if (q - 3 >= z0) fetch_p(a, q - 3, th[t]);
                if (q <= z1 + 2) red(a, lds, q, th[t]);
            }
            for (int t = 0; t < NT; t++) {
                if (q - 1 >= z0 && q - 1 <= z1 + 1) black(a, lds, q - 1, th[t]);
                if (q - 3 >= z0) output(a, lds, q - 3, th[t]);
                if (q + 1 <= z1 + 2) store(lds, q + 1, th[t]);
            }
        }
        double s = 0.0;
        for (int t = 0; t < NT; t++) s += th[t].acc;
        return s;
    }
};

}  // namespace lcs
