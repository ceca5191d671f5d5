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
// A workgroup owns a TX x TY column of cells and a chunk of planes [z0, z1]
// and marches up in z.  At step q:
//   phase A  issue the loads of plane q+3 (U, V, W over the column and its
//            ring) and of P at plane q (owned cells) into registers: three
//            planes stay in flight; red dp on plane q over x in
//            [X0-1, X1+2], y in [Y0-1, Y1+2] (the red neighbours of the black
//            cells below);
//   phase B  black dp on plane q-1 over x in [X0, X1+1], y in [Y0, Y1+1];
//            final U, V, W, P of plane q-3 for the owned cells; plane q+1
//            (loaded two steps earlier) goes into LDS.
// LDS holds U and V for planes q-1..q+1, W for q-2..q+1 over x in
// [X0-2, X1+2], y in [Y0-2, Y1+2], and one dp plane per z for q-3..q in
// which red cells hold red dp and black cells black dp.  A thread keeps its
// own cell's U, V, W for planes q-3..q in registers (the only reader of plane
// q-3).  Two barriers per plane.  (A march with one barrier per plane, whose
// larger rings let red dp of plane q, black dp of q-2 and the output of q-4
// share a phase, was bit-exact and 9 % slower at 1024^3.)  Every cell
// (x, y, z) owned by the workgroup writes P and its upper faces U(x), V(y),
// W(z); the grid covers the whole array, so elements outside R are carried
// over to the other set.
//
// In 1024-thread tiles every global load and store is a range-checked plane
// access (lcs_bufio.h): elements outside the array read as 0 and stores
// outside it are dropped, so the march has no branch around a memory
// operation and the plane loads issued three steps ahead stay in flight until
// the step that stores them.  Smaller tiles, which run several workgroups per
// CU over short chunks, use exec-masked global loads and stores and fetch two
// planes ahead: the plane resources' scalar work per step cost them more than
// the exact waits gained (96-box ensemble 4,530 against 4,160 box-steps/s),
// while the 1024^3 sweep gains 2 % from them.
//
// All of it is __host__ __device__ so a host harness can run the phases
// thread by thread against the oracle.
#pragma once

#include "lcs_bufio.h"
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

// cells of one colour pass over r: cld(nx, 2) per row (foraxes.jl:95-107)
__host__ __device__ inline int pass_count(const Box3 &r)
{
    if (r.hi[0] < r.lo[0] || r.hi[1] < r.lo[1] || r.hi[2] < r.lo[2]) return 0;
    return ((r.hi[0] - r.lo[0] + 2) / 2) * (r.hi[1] - r.lo[1] + 1) * (r.hi[2] - r.lo[2] + 1);
}

// HSMAC regions of a box: 0 core (2:n), 1 d1 (-1:n+3), 2 d2 (0:n+2), 3 d3 (1:n+1)
// (hsmac.jl:97-100, 146-152)
__host__ __device__ inline Box3 hs_region(const Geom &g, int kind)
{
    const int lo = kind == 0 ? 2 : (kind == 1 ? -1 : (kind == 2 ? 0 : 1));
    const int up = kind == 0 ? 0 : (kind == 1 ? 3 : (kind == 2 ? 2 : 1));
    Box3 r = {{lo, lo, lo}, {g.nx + up, g.ny + up, g.nz + up}};
    return r;
}

// the two face slabs of one axis (0: z, 1: y, 2: x) of the overlap order
// around the core (2:n)^3 within region `domain` (lib/topologies/src/halo.jl:349-366)
__host__ __device__ inline void hs_faces(const Geom &g, int domain, int axis, Box3 &f0, Box3 &f1)
{
    const Box3 d = hs_region(g, domain);
    if (axis == 0) {
        f0 = {{d.lo[0], d.lo[1], d.lo[2]}, {d.hi[0], d.hi[1], 1}};
        f1 = {{d.lo[0], d.lo[1], g.nz + 1}, {d.hi[0], d.hi[1], d.hi[2]}};
    } else if (axis == 1) {
        f0 = {{d.lo[0], d.lo[1], 2}, {d.hi[0], 1, g.nz}};
        f1 = {{d.lo[0], g.ny + 1, 2}, {d.hi[0], d.hi[1], g.nz}};
    } else {
        f0 = {{d.lo[0], 2, 2}, {1, g.ny, g.nz}};
        f1 = {{g.nx + 1, 2, 2}, {d.hi[0], g.ny, g.nz}};
    }
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
    // Linear forcing applied to U, V, W as they are loaded (src/flows/forcings/
    // lf.jl:88-100): x <- x + (A (x - xmean)) dt, and the cell power
    // u usrc + v vsrc + w wsrc summed over the interior cells the workgroup
    // owns.  xf = 0: none.
    int xf = 0;
    T xA = T(0), xm[3] = {T(0), T(0), T(0)}, xdt = T(0);
    // U, V, W of a ghost cell are loaded from its periodic image
    // A[mod1(i, n)]: the sweep reads what a FullHalo fill of width h of
    // U, V, W (lib/topologies/src/halo.jl:138-159) would have left, without
    // the fill.  P is loaded as it is.  wrap = 0: plain loads.
    int wrap = 0;
};

// FORCE: the sweep applies the linear forcing of HsArgs::xf on load (the
// first sweep of couple! after forcing!); a separate instantiation, so the
// other sweeps carry none of its registers.
//
// Tiles of more than 1024 cells (128 x 16) give each thread CX = cells / 1024
// cells of one row, WX = TX / CX columns apart, so each load and store of a
// wave still covers 64 consecutive elements of a row; the red and black
// slots then take several passes per thread.
template <typename T, int COLORING, int TY_, int TX_ = 64, bool FORCE = false>
struct HsTile {
    static constexpr int TX = TX_, TY = TY_, NC = TX * TY;
    static constexpr int CX = NC > 1024 ? NC / 1024 : 1;  // owned cells per thread
    static constexpr int NT = NC / CX, WX = TX / CX;        // threads; threads per tile row
    static_assert(CX * NT == NC && WX * CX == TX, "cells split evenly over threads");
    static constexpr int LX = TX + 4, LY = TY + 4, PLANE = LX * LY;  // fields: x in [X0-2, X1+2]
    static constexpr int DXN = TX + 3, DYN = TY + 3, DPL = DXN * DYN;  // dp: x in [X0-1, X1+2]
    static constexpr int OFF_V = 3 * PLANE, OFF_W = 6 * PLANE, OFF_D = 10 * PLANE;
    static constexpr int LDS_ELEMS = OFF_D + 4 * DPL;
    static constexpr int NHALO = 4 * LX + 4 * TY;                   // ring of width 2 around the owned cells
    static constexpr int NHS = (NHALO + NT - 1) / NT;              // halo elements per thread
    static constexpr int RPR = (DXN + 1) / 2, BPR = (TX + 2) / 2;  // cells of one colour per row
    static constexpr int BYN = TY + 1;
    static constexpr int RS = (RPR * DYN + NT - 1) / NT, BS = (BPR * BYN + NT - 1) / NT;  // slots per thread
    // plane accesses (lcs_bufio.h) for 1024-thread tiles
    static constexpr bool PLANES = NT >= 1024;
    // register slots of fetched planes: plane p lives in slot (p - z0) mod
    // NSLOT from the step that fetches it (p - NSLOT) to the step that stores
    // it (p - 1)
    static constexpr int NSLOT = PLANES ? 3 : 2;

    struct Thread {
        int tid, X0, Y0, z0, z1;
        int h_lds[NHS];   // LDS offsets of this thread's halo elements, -1 if none
        long h_off[NHS];  // their in-plane offsets, -1 outside the array
        long o_off[CX];   // in-plane offsets of the owned cells, -1 outside the array
        long u_off[CX];   // where their U, V, W are loaded from (o_off, or the image when wrapping)
        uint32_t o_vo[CX], u_vo[CX];  // byte offsets of o_off and u_off (kOff0 outside the array)
        uint32_t h_vo[NHS];           // byte offsets of h_off (kOff0: none)
        int r_row[RS], r_k[RS], b_row[BS], b_k[BS];
        T pf[NSLOT][3][CX + NHS];  // fetched planes in flight: owned cells, halo elements
        T own[CX][3][4];  // owned cells' U, V, W at planes q-3..q
        T pq[CX][4];      // P of the owned cells at planes q-3..q (the last just fetched)
        double acc;
        bool own_xy[CX];  // the owned column is interior in x and y
        double pacc;      // forcing power of the owned interior cells
    };
    // column of owned cell k of thread tid within the tile, and its row
    __host__ __device__ static int cell_x(int tid, int k) { return tid % WX + k * WX; }
    __host__ __device__ static int cell_y(int tid) { return tid / WX; }

    __host__ __device__ static int m3(int q) { return ((q % 3) + 3) % 3; }
    __host__ __device__ static int m4(int q) { return q & 3; }
    __host__ __device__ static T &Uf(T *lds, int q, int ly, int lx) { return lds[m3(q) * PLANE + ly * LX + lx]; }
    __host__ __device__ static T &Vf(T *lds, int q, int ly, int lx)
    {
        return lds[OFF_V + m3(q) * PLANE + ly * LX + lx];
    }
    __host__ __device__ static T &Wf(T *lds, int q, int ly, int lx)
    {
        return lds[OFF_W + m4(q) * PLANE + ly * LX + lx];
    }
    __host__ __device__ static T &D(T *lds, int q, int dy, int dx) { return lds[OFF_D + m4(q) * DPL + dy * DXN + dx]; }

    __host__ __device__ static long plane_base(const Geom &g, int q) { return (long)(q - 1 + g.h) * g.sx * g.sy; }
    __host__ __device__ static bool plane_in(const Geom &g, int q) { return q >= 1 - g.h && q <= g.nz + g.h; }
    __host__ __device__ static long in_plane(const Geom &g, int x, int y)
    {
        const bool ok = x >= 1 - g.h && x <= g.nx + g.h && y >= 1 - g.h && y <= g.ny + g.h;
        return ok ? (long)(x - 1 + g.h) + g.sx * (long)(y - 1 + g.h) : -1;
    }
    __host__ __device__ static int mod1(int i, int n) { return ((i - 1) % n + n) % n + 1; }
    // in_plane of the periodic image of (x, y); -1 outside the array
    __host__ __device__ static long in_plane_img(const Geom &g, int x, int y)
    {
        return in_plane(g, x, y) < 0 ? -1 : in_plane(g, mod1(x, g.nx), mod1(y, g.ny));
    }
    __host__ __device__ static long load_plane(const HsArgs<T> &a, int q)
    {
        return plane_base(a.g, a.wrap ? mod1(q, a.g.nz) : q);
    }
    __host__ __device__ static uint32_t plane_bytes(const Geom &g) { return (uint32_t)(g.sx * g.sy * sizeof(T)); }
    __host__ __device__ static uint32_t vo(long off) { return off >= 0 ? (uint32_t)(off * (long)sizeof(T)) : kOff0; }
    // plane q of A, or an empty plane (every access out of range) when !ok
    __host__ __device__ static Plane plane_of(const T *A, const Geom &g, long base, bool ok)
    {
        return make_plane(ok ? A + base : A, ok ? plane_bytes(g) : 0u);
    }

    __host__ __device__ static void init(const HsArgs<T> &a, int bx, int by, int bz, int tid, Thread &th)
    {
        const Geom &g = a.g;
        th.tid = tid;
        th.X0 = 1 - g.h + TX * bx;
        th.Y0 = 1 - g.h + TY * by;
        th.z0 = 1 - g.h + a.zc * bz;
        th.z1 = th.z0 + a.zc - 1 < g.nz + g.h ? th.z0 + a.zc - 1 : g.nz + g.h;
        for (int k = 0; k < CX; k++) {
            const int ox = th.X0 + cell_x(tid, k), oy = th.Y0 + cell_y(tid);
            th.o_off[k] = in_plane(g, ox, oy);
            th.u_off[k] = a.wrap ? in_plane_img(g, ox, oy) : th.o_off[k];
            th.own_xy[k] = ox >= 1 && ox <= g.nx && oy >= 1 && oy <= g.ny;
        }
        for (int s = 0; s < NHS; s++) {
            const int e = tid + s * NT;
            th.h_lds[s] = -1;
            th.h_off[s] = -1;
            if (e >= NHALO) continue;
            int lx, ly;
            if (e < 4 * LX) {  // rows 0, 1, TY+2, TY+3
                const int r = e / LX;
                ly = r < 2 ? r : TY + r;
                lx = e % LX;
            } else {  // columns 0, 1, TX+2, TX+3 of rows 2..TY+1
                const int f = e - 4 * LX, c = f % 4;
                ly = 2 + f / 4;
                lx = c < 2 ? c : TX + c;
            }
            th.h_lds[s] = ly * LX + lx;
            const int hx = th.X0 - 2 + lx, hy = th.Y0 - 2 + ly;
            th.h_off[s] = a.wrap ? in_plane_img(g, hx, hy) : in_plane(g, hx, hy);
        }
        for (int k = 0; k < CX; k++) {
            th.o_vo[k] = vo(th.o_off[k]);
            th.u_vo[k] = vo(th.u_off[k]);
        }
        for (int s = 0; s < NHS; s++) th.h_vo[s] = vo(th.h_off[s]);
        for (int r = 0; r < RS; r++) {
            th.r_row[r] = (tid + r * NT) / RPR;
            th.r_k[r] = (tid + r * NT) % RPR;
        }
        for (int r = 0; r < BS; r++) {
            th.b_row[r] = (tid + r * NT) / BPR;
            th.b_k[r] = (tid + r * NT) % BPR;
        }
        for (int k = 0; k < CX; k++) {
            for (int f = 0; f < 3; f++)
                for (int s = 0; s < 4; s++) th.own[k][f][s] = T(0);
            for (int s = 0; s < 4; s++) th.pq[k][s] = T(0);
        }
        th.acc = 0.0;
        th.pacc = 0.0;
    }

    // loads of plane q into register slot s (planes past z1 + 2 are not
    // needed: their loads read as 0 without touching memory)
    __host__ __device__ static void fetch(const HsArgs<T> &a, int q, Thread &th, int s)
    {
        const bool zin = plane_in(a.g, q) && q <= th.z1 + 2;
        const long base = load_plane(a, q);
        if constexpr (PLANES) {
            const Plane pu = plane_of(a.Us, a.g, base, zin), pv = plane_of(a.Vs, a.g, base, zin),
                        pw = plane_of(a.Ws, a.g, base, zin);
            for (int e = 0; e < CX + NHS; e++) {
                const uint32_t off = e < CX ? th.u_vo[e] : th.h_vo[e - CX];
                th.pf[s][0][e] = pload<T>(pu, off);
                th.pf[s][1][e] = pload<T>(pv, off);
                th.pf[s][2][e] = pload<T>(pw, off);
            }
        } else {
            for (int e = 0; e < CX + NHS; e++) {
                const long off = e < CX ? th.u_off[e] : th.h_off[e - CX];
                const bool ok = zin && off >= 0;
                const long p = ok ? base + off : 0;
                th.pf[s][0][e] = ok ? a.Us[p] : T(0);
                th.pf[s][1][e] = ok ? a.Vs[p] : T(0);
                th.pf[s][2][e] = ok ? a.Ws[p] : T(0);
            }
        }
    }

    // FORCE: the linear forcing on the elements of plane q in slot s (at the
    // store, not the fetch, so the loads stay in flight for a whole step)
    __host__ __device__ static void force(const HsArgs<T> &a, int q, Thread &th, int s)
    {
        if (!plane_in(a.g, q)) return;
        for (int e = 0; e < CX + NHS; e++) {
            if ((e < CX ? th.o_off[e] : th.h_off[e - CX]) < 0) continue;
            const T u = th.pf[s][0][e], v = th.pf[s][1][e], w = th.pf[s][2][e];
            const T usrc = a.xA * (u - a.xm[0]), vsrc = a.xA * (v - a.xm[1]), wsrc = a.xA * (w - a.xm[2]);
            if (e < CX && th.own_xy[e] && q >= th.z0 && q <= th.z1 && q >= 1 && q <= a.g.nz)
                th.pacc += (double)((u * usrc + v * vsrc) + w * wsrc);
            th.pf[s][0][e] = u + usrc * a.xdt;
            th.pf[s][1][e] = v + vsrc * a.xdt;
            th.pf[s][2][e] = w + wsrc * a.xdt;
        }
    }

    // plane q (register slot s) into LDS and into the own-cell queue
    __host__ __device__ static void store(const HsArgs<T> &a, T *lds, int q, Thread &th, int s)
    {
        if (FORCE && a.xf) force(a, q, th, s);
        for (int k = 0; k < CX; k++) {
            const int e = (cell_y(th.tid) + 2) * LX + cell_x(th.tid, k) + 2;
            Uf(lds, q, 0, e) = th.pf[s][0][k];
            Vf(lds, q, 0, e) = th.pf[s][1][k];
            Wf(lds, q, 0, e) = th.pf[s][2][k];
        }
        for (int k = 0; k < NHS; k++) {
            const int h = th.h_lds[k];
            if (h < 0) continue;
            Uf(lds, q, 0, h) = th.pf[s][0][k + CX];
            Vf(lds, q, 0, h) = th.pf[s][1][k + CX];
            Wf(lds, q, 0, h) = th.pf[s][2][k + CX];
        }
        shift(th, s);
    }

    // advance the own-cell queue by one plane (the new top comes from slot s)
    __host__ __device__ static void shift(Thread &th, int s)
    {
        for (int k = 0; k < CX; k++)
            for (int f = 0; f < 3; f++) {
                th.own[k][f][0] = th.own[k][f][1];
                th.own[k][f][1] = th.own[k][f][2];
                th.own[k][f][2] = th.own[k][f][3];
                th.own[k][f][3] = th.pf[s][f][k];
            }
    }

    // P of the owned cell at plane o into the top of the P queue (planes
    // outside [z0, z1] read as 0)
    __host__ __device__ static void fetch_p(const HsArgs<T> &a, int o, Thread &th)
    {
        if constexpr (PLANES) {
            const Plane pp = plane_of(a.Ps, a.g, plane_base(a.g, o), o >= th.z0 && o <= th.z1);
            for (int k = 0; k < CX; k++) th.pq[k][3] = pload<T>(pp, th.o_vo[k]);
        } else {
            for (int k = 0; k < CX; k++)
                if (th.o_off[k] >= 0) th.pq[k][3] = a.Ps[plane_base(a.g, o) + th.o_off[k]];
        }
    }

    __host__ __device__ static bool owned(const Thread &th, int x, int y, int z)
    {
        return x >= th.X0 && x < th.X0 + TX && y >= th.Y0 && y < th.Y0 + TY && z >= th.z0 && z <= th.z1;
    }

    // phase A: red dp on plane q
    __host__ __device__ static void red(const HsArgs<T> &a, T *lds, int q, Thread &th)
    {
        for (int r = 0; r < RS; r++) red_slot(a, lds, q, th, th.r_row[r], th.r_k[r]);
    }
    __host__ __device__ static void red_slot(const HsArgs<T> &a, T *lds, int q, Thread &th, int r_row, int r_k)
    {
        if (r_row >= DYN) return;
        const int xs = th.X0 - 1, y = th.Y0 - 1 + r_row;
        const int dx = ((xs + y + q) & 1) + 2 * r_k;
        if (dx >= DXN) return;
        const int x = xs + dx, dy = r_row, lx = dx + 1, ly = dy + 1;
        const HsCoef<T> &c = a.c;
        T dp = T(0);
        if (visited<COLORING>(a.R, x, y, q)) {
            const T dudx = c.dxi * (-Uf(lds, q, ly, lx - 1) + Uf(lds, q, ly, lx));
            const T dvdy = c.dyi * (-Vf(lds, q, ly - 1, lx) + Vf(lds, q, ly, lx));
            const T dwdz = c.dzi * (-Wf(lds, q - 1, ly, lx) + Wf(lds, q, ly, lx));
            const T div = (dudx + dvdy) + dwdz;
            dp = (-c.omega * div) * c.dl;
            if (owned(th, x, y, q) && in_d3(a.g, x, y, q)) th.acc += (double)(div * div);
        }
        D(lds, q, dy, dx) = dp;
    }

    // phase B, first part: black dp on plane p, on faces already corrected by
    // their red neighbours
    __host__ __device__ static void black(const HsArgs<T> &a, T *lds, int p, Thread &th)
    {
        for (int r = 0; r < BS; r++) black_slot(a, lds, p, th, th.b_row[r], th.b_k[r]);
    }
    __host__ __device__ static void black_slot(const HsArgs<T> &a, T *lds, int p, Thread &th, int b_row, int b_k)
    {
        if (b_row >= BYN) return;
        const int y = th.Y0 + b_row;
        const int bx = ((th.X0 + y + p + 1) & 1) + 2 * b_k;
        if (bx > TX) return;
        const int x = th.X0 + bx, dx = bx + 1, dy = b_row + 1, lx = bx + 2, ly = b_row + 2;
        const HsCoef<T> &c = a.c;
        T dp = T(0);
        if (visited<COLORING>(a.R, x, y, p)) {
            const T Ur = Uf(lds, p, ly, lx) - D(lds, p, dy, dx + 1) * c.dxi;
            const T Ul = Uf(lds, p, ly, lx - 1) + D(lds, p, dy, dx - 1) * c.dxi;
            const T Vr = Vf(lds, p, ly, lx) - D(lds, p, dy + 1, dx) * c.dyi;
            const T Vl = Vf(lds, p, ly - 1, lx) + D(lds, p, dy - 1, dx) * c.dyi;
            const T Wt = Wf(lds, p, ly, lx) - D(lds, p + 1, dy, dx) * c.dzi;
            const T Wb = Wf(lds, p - 1, ly, lx) + D(lds, p - 1, dy, dx) * c.dzi;
            const T dudx = c.dxi * (-Ul + Ur);
            const T dvdy = c.dyi * (-Vl + Vr);
            const T dwdz = c.dzi * (-Wb + Wt);
            const T div = (dudx + dvdy) + dwdz;
            dp = (-c.omega * div) * c.dl;
            if (owned(th, x, y, p) && in_d3(a.g, x, y, p)) th.acc += (double)(div * div);
        }
        D(lds, p, dy, dx) = dp;
    }

    // phase B, second part: final values of the owned cell on plane o
    // (its U, V, W are own[.][0], its P is pq[0]).  Planes below z0 and cells
    // outside the array compute values that are not stored.
    __host__ __device__ static void output(const HsArgs<T> &a, T *lds, int o, Thread &th)
    {
        for (int k = 0; k < CX; k++) output_cell(a, lds, o, th, k);
    }
    __host__ __device__ static void output_cell(const HsArgs<T> &a, T *lds, int o, Thread &th, int k)
    {
        if (!PLANES && th.o_off[k] < 0) return;
        const int tx = cell_x(th.tid, k), ty = cell_y(th.tid);
        const int x = th.X0 + tx, y = th.Y0 + ty, dx = tx + 1, dy = ty + 1;
        const HsCoef<T> &c = a.c;
        const bool isred = ((x + y + o) & 1) == 0;
        // correction of this cell and of its +x, +y, +z neighbours (0 outside R)
        const T d0 = D(lds, o, dy, dx), dx1 = D(lds, o, dy, dx + 1), dy1 = D(lds, o, dy + 1, dx);
        const T dz1 = D(lds, o + 1, dy, dx);
        T pv = th.pq[k][0];
        if (visited<COLORING>(a.R, x, y, o)) pv += d0 * a.dti;
        T u = th.own[k][0][0], v = th.own[k][1][0], w = th.own[k][2][0];
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
        const long base = plane_base(a.g, o);
        if constexpr (PLANES) {
            const bool zin = o >= th.z0;
            pstore<T>(plane_of(a.Pd, a.g, base, zin), th.o_vo[k], pv);
            pstore<T>(plane_of(a.Ud, a.g, base, zin), th.o_vo[k], u);
            pstore<T>(plane_of(a.Vd, a.g, base, zin), th.o_vo[k], v);
            pstore<T>(plane_of(a.Wd, a.g, base, zin), th.o_vo[k], w);
        } else {
            const long gi = base + th.o_off[k];
            a.Pd[gi] = pv;
            a.Ud[gi] = u;
            a.Vd[gi] = v;
            a.Wd[gi] = w;
        }
    }

    // One step of the march.  Step q = z0 - 1 + u fetches plane q + NSLOT into
    // slot sf = (u + NSLOT - 1) mod NSLOT and stores plane q + 1 from slot
    // ss = u mod NSLOT; the kernel unrolls the march by NSLOT so both are
    // compile-time constants.  P of plane q (plane accesses) or q - 2 enters
    // the top of the P queue.  SYNC is a workgroup barrier in the kernel; the
    // host harness runs every thread's part A, then every thread's part B.
    __host__ __device__ static void step_a(const HsArgs<T> &a, T *lds, int q, Thread &th, int sf)
    {
        if constexpr (PLANES) {
            fetch(a, q + 3, th, sf);
            fetch_p(a, q, th);
        } else {
            if (q + 2 <= th.z1 + 2) fetch(a, q + 2, th, sf);
            if (q - 2 >= th.z0 && q - 2 <= th.z1) fetch_p(a, q - 2, th);
        }
        if (q <= th.z1 + 2) red(a, lds, q, th);
    }
    __host__ __device__ static void step_b(const HsArgs<T> &a, T *lds, int q, Thread &th, int ss)
    {
        const int z0 = th.z0, z1 = th.z1;
        if (q - 1 >= z0 && q - 1 <= z1 + 1) black(a, lds, q - 1, th);
        if (PLANES || q - 3 >= z0) output(a, lds, q - 3, th);
        if (q + 1 <= z1 + 2)
            store(a, lds, q + 1, th, ss);
        else
            shift(th, ss);
        for (int k = 0; k < CX; k++) {
            if constexpr (PLANES) {
                th.pq[k][0] = th.pq[k][1];
                th.pq[k][1] = th.pq[k][2];
                th.pq[k][2] = th.pq[k][3];
            } else {
                th.pq[k][0] = th.pq[k][3];
            }
        }
    }
    // the first and last steps of a chunk's march
    __host__ __device__ static int first_step(const Thread &th) { return th.z0 - 1; }
    __host__ __device__ static int last_step(const Thread &th) { return th.z1 + 3; }
    // before the first step: planes z0-2 and z0-1 in LDS, planes z0 .. z0 +
    // NSLOT - 2 in flight (slots 0 ..)
    __host__ __device__ static void prologue(const HsArgs<T> &a, T *lds, Thread &th)
    {
        fetch(a, th.z0 - 2, th, NSLOT - 1);
        store(a, lds, th.z0 - 2, th, NSLOT - 1);
        fetch(a, th.z0 - 1, th, NSLOT - 1);
        store(a, lds, th.z0 - 1, th, NSLOT - 1);
        for (int k = 0; k < NSLOT - 1; k++) fetch(a, th.z0 + k, th, k);
    }

    // The march of one workgroup for a host harness: the threads of each
    // part run one after the other.
    static double emulate_block(const HsArgs<T> &a, T *lds, Thread *th, int bx, int by, int bz,
                                double *power = nullptr)
    {
        for (int t = 0; t < NT; t++) init(a, bx, by, bz, t, th[t]);
        for (int t = 0; t < NT; t++) prologue(a, lds, th[t]);
        for (int q = first_step(th[0]), u = 0; q <= last_step(th[0]); q++, u++) {
            for (int t = 0; t < NT; t++) step_a(a, lds, q, th[t], (u + NSLOT - 1) % NSLOT);
            for (int t = 0; t < NT; t++) step_b(a, lds, q, th[t], u % NSLOT);
        }
        double sum = 0.0;
        for (int t = 0; t < NT; t++) sum += th[t].acc;
        if (power)
            for (int t = 0; t < NT; t++) *power += th[t].pacc;
        return sum;
    }
};

}  // namespace lcs
