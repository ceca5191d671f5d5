// Two fused red+black HSMAC sweeps in one pass over HBM: sweep A over box RA,
// then sweep B over box RB on A's result -- the d2 and d3 sweeps of a
// time-blocked group (src/flows/couples/hsmac.jl:146-189), in the groups
// before itrmin where neither sweep can end the iteration.
//
// A workgroup owns a TX x TY column of cells and a chunk of planes [z0, z1].
// Sweep B is exactly the single sweep of lcs_hsmac_tile.h for the owned cells;
// its input (A's result) over the owned cells plus a 2-cell ring never leaves
// LDS.  Sweep A therefore produces its result over that extended region E,
// which needs A's red corrections over E + 1..2 cells and A's input over
// E + 2 cells (x in [X0-4, X1+4]).  A runs planes [z0-2, z1+2] (what B reads)
// and B follows five planes behind, so at step q (two barriers):
//   phase 1  A: loads of plane q+2 into registers, red dp on plane q;
//            B: red dp on plane q-5;
//   phase 2  A: black dp on plane q-1, its result on plane q-3 over E into
//            B's rings (and P of the owned cells into a 2-plane ring), plane
//            q+1 into A's rings;
//            B: black dp on plane q-6, final U, V, W, P of plane q-8 for the
//            owned cells, and its own-cell queues take plane q-4 of A's result.
// Ring slots never overlap between writer and readers: B's W ring has five
// planes (q-7..q-3), its U and V rings four (q-6..q-3), the P ring two.
// Cells of E outside the array, and planes outside it, hold 0 -- what the
// second of two separate sweeps would load there.
//
// Global loads and stores are range-checked plane accesses (lcs_bufio.h), with
// no branch around them, as in the single sweep.
//
// The div^2 sums of both sweeps are kept apart (one per check of the
// reference).  Every element of the output set is written by exactly one
// workgroup, so the result equals two launches of the single sweep bit for bit.
#pragma once

#include "lcs_hsmac_tile.h"

namespace lcs {

template <typename T>
struct Hs2Args {
    Geom g;
    const T *Us, *Vs, *Ws, *Ps;  // input of sweep A
    T *Ud, *Vd, *Wd, *Pd;        // result of sweep B
    Box3 RA, RB;
    HsCoef<T> c;
    T dti;
    int zc;
};

template <typename T, int COLORING, int TY_, int TX_>
struct HsTile2 {
    static constexpr int TX = TX_, TY = TY_, NT = TX * TY;
    static constexpr int AX = TX + 8, AY = TY + 8, APL = AX * AY;       // A input: x in [X0-4, X1+4]
    static constexpr int ADX = TX + 7, ADY = TY + 7, ADPL = ADX * ADY;  // A dp (red region): x in [X0-3, X1+4]
    static constexpr int EX = TX + 4, EY = TY + 4, EPL = EX * EY;       // A result = B input: x in [X0-2, X1+2]
    static constexpr int BDX = TX + 3, BDY = TY + 3, BDPL = BDX * BDY;  // B dp: x in [X0-1, X1+2]
    static constexpr int A_U = 0, A_V = 3 * APL, A_W = 6 * APL, A_D = 10 * APL;
    static constexpr int B_U = A_D + 4 * ADPL, B_V = B_U + 4 * EPL, B_W = B_V + 4 * EPL, B_D = B_W + 5 * EPL;
    static constexpr int P_R = B_D + 4 * BDPL;
    static constexpr int LDS_ELEMS = P_R + 2 * NT;
    static constexpr int NA = (APL + NT - 1) / NT;  // A-input elements per thread (cells of E first)
    static constexpr int NE = (EPL + NT - 1) / NT;  // cells of E per thread
    static constexpr int A_RPR = (ADX + 1) / 2;                                    // A red: x in [X0-3, X1+4]
    static constexpr int A_BXN = TX + 5, A_BPR = (A_BXN + 1) / 2, A_BYN = TY + 5;  // A black: [X0-2, X1+3]
    static constexpr int B_RPR = (BDX + 1) / 2;                                    // B red: [X0-1, X1+2]
    static constexpr int B_BPR = (TX + 2) / 2, B_BYN = TY + 1;                     // B black: [X0, X1+1]
    static_assert(A_RPR * ADY <= NT && A_BPR * A_BYN <= NT, "one A red and one A black slot per thread");
    static_assert(B_RPR * BDY <= NT && B_BPR * B_BYN <= NT, "one B red and one B black slot per thread");

    struct Thread {
        int tid, X0, Y0, z0, z1, za0, za1;
        int a_lds[NA];    // A-input ring index (ay * AX + ax) of element tid + k NT, -1 if none
        long a_off[NA];   // its in-plane offset, -1 outside the array
        uint32_t a_vo[NA];  // byte offsets of a_off (kOff0 outside the array)
        uint32_t p_vo[NE];  // byte offsets of the owned E cells' P (kOff0 for the others)
        uint32_t o_vo;      // byte offset of o_off (kOff0 outside the array)
        int e_own[NE];    // E slot k: index of the owned cell it is, -1 if it is not one
        long o_off;       // the thread's owned cell (sweep B), -1 outside the array
        T pf[2][3][NA];   // two A-input planes in flight
        T aown[NE][3][4];  // E cells' U, V, W at planes q-3..q (A's result reads plane q-3)
        T apP[NE], apPn[NE];  // A's P input of owned E cells, planes q-3 and q-2
        T bown[3][4];     // B's own cell U, V, W (A's result), planes qb-3..qb
        T bpq[4];         // A's P result at the own cell, planes qb-3..qb
        int r_row, r_k, b_row, b_k, rb_row, rb_k, bb_row, bb_k;
        double accA, accB;
    };

    __host__ __device__ static int m3(int q) { return ((q % 3) + 3) % 3; }
    __host__ __device__ static int m5(int q) { return ((q % 5) + 5) % 5; }
    __host__ __device__ static T &AU(T *l, int q, int y, int x) { return l[A_U + m3(q) * APL + y * AX + x]; }
    __host__ __device__ static T &AV(T *l, int q, int y, int x) { return l[A_V + m3(q) * APL + y * AX + x]; }
    __host__ __device__ static T &AW(T *l, int q, int y, int x) { return l[A_W + (q & 3) * APL + y * AX + x]; }
    __host__ __device__ static T &AD(T *l, int q, int y, int x) { return l[A_D + (q & 3) * ADPL + y * ADX + x]; }
    __host__ __device__ static T &BU(T *l, int q, int y, int x) { return l[B_U + (q & 3) * EPL + y * EX + x]; }
    __host__ __device__ static T &BV(T *l, int q, int y, int x) { return l[B_V + (q & 3) * EPL + y * EX + x]; }
    __host__ __device__ static T &BW(T *l, int q, int y, int x) { return l[B_W + m5(q) * EPL + y * EX + x]; }
    __host__ __device__ static T &BD(T *l, int q, int y, int x) { return l[B_D + (q & 3) * BDPL + y * BDX + x]; }
    __host__ __device__ static T &PR(T *l, int q, int i) { return l[P_R + (q & 1) * NT + i]; }

    using One = HsTile<T, COLORING, TY, TX>;  // shared helpers: plane_base, plane_in, in_plane

    __host__ __device__ static bool owned(const Thread &th, int x, int y, int z)
    {
        return x >= th.X0 && x < th.X0 + TX && y >= th.Y0 && y < th.Y0 + TY && z >= th.z0 && z <= th.z1;
    }

    __host__ __device__ static void init(const Hs2Args<T> &a, int bx, int by, int bz, int tid, Thread &th)
    {
        const Geom &g = a.g;
        th.tid = tid;
        th.X0 = 1 - g.h + TX * bx;
        th.Y0 = 1 - g.h + TY * by;
        th.z0 = 1 - g.h + a.zc * bz;
        th.z1 = th.z0 + a.zc - 1 < g.nz + g.h ? th.z0 + a.zc - 1 : g.nz + g.h;
        th.za0 = th.z0 - 2;
        th.za1 = th.z1 + 2;
        th.o_off = One::in_plane(g, th.X0 + tid % TX, th.Y0 + tid / TX);
        for (int k = 0; k < NA; k++) {
            const int e = tid + k * NT;
            th.a_lds[k] = -1;
            th.a_off[k] = -1;
            if (e >= APL) continue;
            int ax, ay;
            if (e < EPL) {  // the cells of E, row-major
                ax = e % EX + 2;
                ay = e / EX + 2;
            } else {
                const int f = e - EPL;
                if (f < 4 * AX) {  // rows 0, 1, AY-2, AY-1
                    const int r = f / AX;
                    ay = r < 2 ? r : AY - 4 + r;
                    ax = f % AX;
                } else {  // columns 0, 1, AX-2, AX-1 of rows 2..AY-3
                    const int f2 = f - 4 * AX, c = f2 % 4;
                    ay = 2 + f2 / 4;
                    ax = c < 2 ? c : AX - 4 + c;
                }
            }
            th.a_lds[k] = ay * AX + ax;
            th.a_off[k] = One::in_plane(g, th.X0 - 4 + ax, th.Y0 - 4 + ay);
        }
        for (int k = 0; k < NA; k++) th.a_vo[k] = One::vo(th.a_off[k]);
        th.o_vo = One::vo(th.o_off);
        for (int k = 0; k < NE; k++) {
            const int e = tid + k * NT;
            th.e_own[k] = -1;
            for (int f = 0; f < 3; f++)
                for (int s = 0; s < 4; s++) th.aown[k][f][s] = T(0);
            th.apP[k] = th.apPn[k] = T(0);
            if (e >= EPL) continue;
            const int ex = e % EX, ey = e / EX;
            if (ex >= 2 && ex < TX + 2 && ey >= 2 && ey < TY + 2) th.e_own[k] = (ey - 2) * TX + (ex - 2);
        }
        for (int k = 0; k < NE; k++) th.p_vo[k] = th.e_own[k] >= 0 && k < NA ? th.a_vo[k] : kOff0;
        th.r_row = tid / A_RPR;
        th.r_k = tid % A_RPR;
        th.b_row = tid / A_BPR;
        th.b_k = tid % A_BPR;
        th.rb_row = tid / B_RPR;
        th.rb_k = tid % B_RPR;
        th.bb_row = tid / B_BPR;
        th.bb_k = tid % B_BPR;
        for (int f = 0; f < 3; f++)
            for (int s = 0; s < 4; s++) th.bown[f][s] = T(0);
        for (int s = 0; s < 4; s++) th.bpq[s] = T(0);
        th.accA = th.accB = 0.0;
    }

    // ---- sweep A ---------------------------------------------------------------
    // A's input plane q into register slot s (planes past za1 + 2 read as 0
    // without touching memory)
    __host__ __device__ static void fetch(const Hs2Args<T> &a, int q, Thread &th, int s)
    {
        const bool zin = One::plane_in(a.g, q) && q <= th.za1 + 2;
        const long base = One::plane_base(a.g, q);
        const Plane pu = One::plane_of(a.Us, a.g, base, zin), pv = One::plane_of(a.Vs, a.g, base, zin),
                    pw = One::plane_of(a.Ws, a.g, base, zin);
        for (int k = 0; k < NA; k++) {
            th.pf[s][0][k] = pload<T>(pu, th.a_vo[k]);
            th.pf[s][1][k] = pload<T>(pv, th.a_vo[k]);
            th.pf[s][2][k] = pload<T>(pw, th.a_vo[k]);
        }
    }
    // the E cells' queues advance by one plane (the new top comes from slot s)
    __host__ __device__ static void shift(Thread &th, int s)
    {
        for (int k = 0; k < NE; k++)
            for (int f = 0; f < 3; f++) {
                th.aown[k][f][0] = th.aown[k][f][1];
                th.aown[k][f][1] = th.aown[k][f][2];
                th.aown[k][f][2] = th.aown[k][f][3];
                th.aown[k][f][3] = th.pf[s][f][k];
            }
    }
    __host__ __device__ static void store(T *lds, int q, Thread &th, int s)
    {
        for (int k = 0; k < NA; k++) {
            const int i = th.a_lds[k];
            if (i < 0) continue;
            AU(lds, q, 0, i) = th.pf[s][0][k];
            AV(lds, q, 0, i) = th.pf[s][1][k];
            AW(lds, q, 0, i) = th.pf[s][2][k];
        }
        shift(th, s);
    }
    // A's P input of the owned E cells at plane o (A's planes reach two beyond
    // the chunk, and so past the array: those read as 0 and are not used)
    __host__ __device__ static void fetch_p(const Hs2Args<T> &a, int o, Thread &th)
    {
        const bool zin = One::plane_in(a.g, o) && o >= th.za0 && o <= th.za1;
        const Plane pp = One::plane_of(a.Ps, a.g, One::plane_base(a.g, o), zin);
        for (int k = 0; k < NE; k++) th.apPn[k] = pload<T>(pp, th.p_vo[k]);
    }
    __host__ __device__ static void red_a(const Hs2Args<T> &a, T *lds, int q, Thread &th)
    {
        if (th.r_row >= ADY) return;
        const int xs = th.X0 - 3, y = th.Y0 - 3 + th.r_row;
        const int dx = ((xs + y + q) & 1) + 2 * th.r_k;
        if (dx >= ADX) return;
        const int x = xs + dx, dy = th.r_row, ax = dx + 1, ay = dy + 1;
        const HsCoef<T> &c = a.c;
        T dp = T(0);
        if (visited<COLORING>(a.RA, x, y, q)) {
            const T dudx = c.dxi * (-AU(lds, q, ay, ax - 1) + AU(lds, q, ay, ax));
            const T dvdy = c.dyi * (-AV(lds, q, ay - 1, ax) + AV(lds, q, ay, ax));
            const T dwdz = c.dzi * (-AW(lds, q - 1, ay, ax) + AW(lds, q, ay, ax));
            const T div = (dudx + dvdy) + dwdz;
            dp = (-c.omega * div) * c.dl;
            if (owned(th, x, y, q) && in_d3(a.g, x, y, q)) th.accA += (double)(div * div);
        }
        AD(lds, q, dy, dx) = dp;
    }
    __host__ __device__ static void black_a(const Hs2Args<T> &a, T *lds, int p, Thread &th)
    {
        if (th.b_row >= A_BYN) return;
        const int y = th.Y0 - 2 + th.b_row;
        const int bx = ((th.X0 - 2 + y + p + 1) & 1) + 2 * th.b_k;
        if (bx >= A_BXN) return;
        const int x = th.X0 - 2 + bx, dx = bx + 1, dy = th.b_row + 1, ax = bx + 2, ay = th.b_row + 2;
        const HsCoef<T> &c = a.c;
        T dp = T(0);
        if (visited<COLORING>(a.RA, x, y, p)) {
            const T Ur = AU(lds, p, ay, ax) - AD(lds, p, dy, dx + 1) * c.dxi;
            const T Ul = AU(lds, p, ay, ax - 1) + AD(lds, p, dy, dx - 1) * c.dxi;
            const T Vr = AV(lds, p, ay, ax) - AD(lds, p, dy + 1, dx) * c.dyi;
            const T Vl = AV(lds, p, ay - 1, ax) + AD(lds, p, dy - 1, dx) * c.dyi;
            const T Wt = AW(lds, p, ay, ax) - AD(lds, p + 1, dy, dx) * c.dzi;
            const T Wb = AW(lds, p - 1, ay, ax) + AD(lds, p - 1, dy, dx) * c.dzi;
            const T dudx = c.dxi * (-Ul + Ur);
            const T dvdy = c.dyi * (-Vl + Vr);
            const T dwdz = c.dzi * (-Wb + Wt);
            const T div = (dudx + dvdy) + dwdz;
            dp = (-c.omega * div) * c.dl;
            if (owned(th, x, y, p) && in_d3(a.g, x, y, p)) th.accA += (double)(div * div);
        }
        AD(lds, p, dy, dx) = dp;
    }
    // A's result on plane o over E into B's rings, P of owned cells into the P ring
    __host__ __device__ static void output_a(const Hs2Args<T> &a, T *lds, int o, Thread &th)
    {
        const HsCoef<T> &c = a.c;
        const bool zin = One::plane_in(a.g, o);
        for (int k = 0; k < NE; k++) {
            const int e = th.tid + k * NT;
            if (e >= EPL) continue;
            const int ex = e % EX, ey = e / EX;
            T u = T(0), v = T(0), w = T(0), pv = T(0);
            if (zin && th.a_off[k] >= 0) {
                const int x = th.X0 - 2 + ex, y = th.Y0 - 2 + ey, dx = ex + 1, dy = ey + 1;
                const bool isred = ((x + y + o) & 1) == 0;
                const T d0 = AD(lds, o, dy, dx), dx1 = AD(lds, o, dy, dx + 1), dy1 = AD(lds, o, dy + 1, dx);
                const T dz1 = AD(lds, o + 1, dy, dx);
                pv = th.apP[k];
                if (visited<COLORING>(a.RA, x, y, o)) pv += d0 * a.dti;
                u = th.aown[k][0][0];
                v = th.aown[k][1][0];
                w = th.aown[k][2][0];
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
            }
            BU(lds, o, ey, ex) = u;
            BV(lds, o, ey, ex) = v;
            BW(lds, o, ey, ex) = w;
            if (th.e_own[k] >= 0) PR(lds, o, th.e_own[k]) = pv;
        }
    }

    // ---- sweep B (the single sweep of lcs_hsmac_tile.h on LDS input) ------------
    __host__ __device__ static void red_b(const Hs2Args<T> &a, T *lds, int q, Thread &th)
    {
        if (th.rb_row >= BDY) return;
        const int xs = th.X0 - 1, y = th.Y0 - 1 + th.rb_row;
        const int dx = ((xs + y + q) & 1) + 2 * th.rb_k;
        if (dx >= BDX) return;
        const int x = xs + dx, dy = th.rb_row, lx = dx + 1, ly = dy + 1;
        const HsCoef<T> &c = a.c;
        T dp = T(0);
        if (visited<COLORING>(a.RB, x, y, q)) {
            const T dudx = c.dxi * (-BU(lds, q, ly, lx - 1) + BU(lds, q, ly, lx));
            const T dvdy = c.dyi * (-BV(lds, q, ly - 1, lx) + BV(lds, q, ly, lx));
            const T dwdz = c.dzi * (-BW(lds, q - 1, ly, lx) + BW(lds, q, ly, lx));
            const T div = (dudx + dvdy) + dwdz;
            dp = (-c.omega * div) * c.dl;
            if (owned(th, x, y, q) && in_d3(a.g, x, y, q)) th.accB += (double)(div * div);
        }
        BD(lds, q, dy, dx) = dp;
    }
    __host__ __device__ static void black_b(const Hs2Args<T> &a, T *lds, int p, Thread &th)
    {
        if (th.bb_row >= B_BYN) return;
        const int y = th.Y0 + th.bb_row;
        const int bx = ((th.X0 + y + p + 1) & 1) + 2 * th.bb_k;
        if (bx > TX) return;
        const int x = th.X0 + bx, dx = bx + 1, dy = th.bb_row + 1, lx = bx + 2, ly = th.bb_row + 2;
        const HsCoef<T> &c = a.c;
        T dp = T(0);
        if (visited<COLORING>(a.RB, x, y, p)) {
            const T Ur = BU(lds, p, ly, lx) - BD(lds, p, dy, dx + 1) * c.dxi;
            const T Ul = BU(lds, p, ly, lx - 1) + BD(lds, p, dy, dx - 1) * c.dxi;
            const T Vr = BV(lds, p, ly, lx) - BD(lds, p, dy + 1, dx) * c.dyi;
            const T Vl = BV(lds, p, ly - 1, lx) + BD(lds, p, dy - 1, dx) * c.dyi;
            const T Wt = BW(lds, p, ly, lx) - BD(lds, p + 1, dy, dx) * c.dzi;
            const T Wb = BW(lds, p - 1, ly, lx) + BD(lds, p - 1, dy, dx) * c.dzi;
            const T dudx = c.dxi * (-Ul + Ur);
            const T dvdy = c.dyi * (-Vl + Vr);
            const T dwdz = c.dzi * (-Wb + Wt);
            const T div = (dudx + dvdy) + dwdz;
            dp = (-c.omega * div) * c.dl;
            if (owned(th, x, y, p) && in_d3(a.g, x, y, p)) th.accB += (double)(div * div);
        }
        BD(lds, p, dy, dx) = dp;
    }
    // final values of the owned cell on plane o; stored only for o in [z0, z1]
    // and cells inside the array
    __host__ __device__ static void output_b(const Hs2Args<T> &a, T *lds, int o, Thread &th)
    {
        const int tx = th.tid % TX, ty = th.tid / TX;
        const int x = th.X0 + tx, y = th.Y0 + ty, dx = tx + 1, dy = ty + 1;
        const HsCoef<T> &c = a.c;
        const bool isred = ((x + y + o) & 1) == 0;
        const T d0 = BD(lds, o, dy, dx), dx1 = BD(lds, o, dy, dx + 1), dy1 = BD(lds, o, dy + 1, dx);
        const T dz1 = BD(lds, o + 1, dy, dx);
        T pv = th.bpq[0];
        if (visited<COLORING>(a.RB, x, y, o)) pv += d0 * a.dti;
        T u = th.bown[0][0], v = th.bown[1][0], w = th.bown[2][0];
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
        const long base = One::plane_base(a.g, o);
        const bool zin = o >= th.z0 && o <= th.z1;
        pstore<T>(One::plane_of(a.Pd, a.g, base, zin), th.o_vo, pv);
        pstore<T>(One::plane_of(a.Ud, a.g, base, zin), th.o_vo, u);
        pstore<T>(One::plane_of(a.Vd, a.g, base, zin), th.o_vo, v);
        pstore<T>(One::plane_of(a.Wd, a.g, base, zin), th.o_vo, w);
    }
    // B's own-cell queues take plane r of A's result
    __host__ __device__ static void take_b(T *lds, int r, Thread &th)
    {
        const int tx = th.tid % TX, ty = th.tid / TX, ex = tx + 2, ey = ty + 2;
        const T nu = BU(lds, r, ey, ex), nv = BV(lds, r, ey, ex), nw = BW(lds, r, ey, ex), np = PR(lds, r, th.tid);
        for (int s = 0; s < 3; s++) {
            th.bown[0][s] = th.bown[0][s + 1];
            th.bown[1][s] = th.bown[1][s + 1];
            th.bown[2][s] = th.bown[2][s + 1];
            th.bpq[s] = th.bpq[s + 1];
        }
        th.bown[0][3] = nu;
        th.bown[1][3] = nv;
        th.bown[2][3] = nw;
        th.bpq[3] = np;
    }

    // ---- the march -----------------------------------------------------------------
    __host__ __device__ static void step_1(const Hs2Args<T> &a, T *lds, int q, Thread &th, int s)
    {
        fetch(a, q + 2, th, s ^ 1);
        fetch_p(a, q - 2, th);
        if (q <= th.za1 + 2) red_a(a, lds, q, th);
        const int qb = q - 5;
        if (qb >= th.z0 - 1 && qb <= th.z1 + 2) red_b(a, lds, qb, th);
    }
    __host__ __device__ static void step_2(const Hs2Args<T> &a, T *lds, int q, Thread &th, int s)
    {
        if (q - 1 >= th.za0 && q - 1 <= th.za1 + 1) black_a(a, lds, q - 1, th);
        if (q - 3 >= th.za0 && q - 3 <= th.za1) output_a(a, lds, q - 3, th);
        if (q + 1 <= th.za1 + 2)
            store(lds, q + 1, th, s);
        else
            shift(th, s);
        for (int k = 0; k < NE; k++) th.apP[k] = th.apPn[k];
        const int qb = q - 5;
        if (qb - 1 >= th.z0 && qb - 1 <= th.z1 + 1) black_b(a, lds, qb - 1, th);
        output_b(a, lds, qb - 3, th);
        take_b(lds, q - 4, th);
    }
    // before the first step: A's planes za0-2 and za0-1 in LDS, za0 in flight (slot 0)
    __host__ __device__ static void prologue(const Hs2Args<T> &a, T *lds, Thread &th)
    {
        fetch(a, th.za0 - 2, th, 0);
        store(lds, th.za0 - 2, th, 0);
        fetch(a, th.za0 - 1, th, 0);
        store(lds, th.za0 - 1, th, 0);
        fetch(a, th.za0, th, 0);
    }
    __host__ __device__ static int first_step(const Thread &th) { return th.za0 - 1; }
    __host__ __device__ static int last_step(const Thread &th) { return th.z1 + 8; }

    // The march of one workgroup for a host harness.  Returns the div^2 sums of
    // sweep A and sweep B over this tile's owned cells.
    static void emulate_block(const Hs2Args<T> &a, T *lds, Thread *th, int bx, int by, int bz, double &sa,
                              double &sb)
    {
        for (int t = 0; t < NT; t++) init(a, bx, by, bz, t, th[t]);
        for (int t = 0; t < NT; t++) prologue(a, lds, th[t]);
        int s = 0;
        for (int q = first_step(th[0]); q <= last_step(th[0]); q++, s ^= 1) {
            for (int t = 0; t < NT; t++) step_1(a, lds, q, th[t], s);
            for (int t = 0; t < NT; t++) step_2(a, lds, q, th[t], s);
        }
        sa = sb = 0.0;
        for (int t = 0; t < NT; t++) {
            sa += th[t].accA;
            sb += th[t].accB;
        }
    }
};

}  // namespace lcs
