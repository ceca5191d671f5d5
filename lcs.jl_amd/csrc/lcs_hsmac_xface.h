// Both colour passes over the two x face slabs of the overlap order, as one
// launch over patches of slab rows plus a small commit launch, instead of one
// launch per colour.  The slabs are three cells (d1) or one cell (d3) thick,
// so a colour pass touches a few elements of every array row it visits and
// moves whole memory segments for them, and two passes move them twice.  Here
// a row's segments are read about once (patch windows overlap by two rows)
// and written once.
//
// Semantics (lib/topologies/src/halo.jl:349-366, src/flows/couples/hsmac.jl:
// 1-35, lib/parallel/src/foraxes.jl:73-107): over slab R, every visited red
// cell ((x + y + z) even) applies its correction dp to its six faces and to P,
// then every visited black cell does the same with the divergence of the
// faces the red cells left.  A face is touched by at most one red and one
// black cell, so its final value is ((f0 +- red dp*di) +- black dp*di).
//
// Launch 1 (compute): a workgroup owns a patch of PY x PZ slab rows (y, z) of
// one slab, loads the fields of the patch and two rows around it into LDS,
// forms red corrections one row beyond the patch and black corrections one
// row beyond it in +y and +z (the +y and +z faces of the patch rows take the
// next row's corrections), and writes the final U, V, W, P of its rows.  The
// two outer rows on each side of a patch are in a neighbour's window, read in
// the same launch, so their values go to a compact buffer; launch 2 (commit)
// copies them into the fields.  The inner rows are written in place.
#pragma once

#include "lcs_hsmac_tile.h"

namespace lcs {

template <typename T>
struct XfArgs {
    Geom g;
    T *U, *V, *W, *P;        // the current set (launch 1 reads it and writes its inner rows)
    Box3 R[2];               // the two slabs of the axis, lower one first
    HsCoef<T> c;
    T dti;
    T *buf;  // compact final values (commit_row)
};

// A pair of face slabs across axis TI (Box3 index: 0 x, 1 y, 2 z) as rows of
// 1 to 3 cells along TI; a row is one (a, b) position along the other two
// axes A1 < A2.  For the x slabs the rows are (y, z), for the y slabs (x, z),
// for the z slabs (x, y).
template <int TI>
struct SlabAxes {
    static constexpr int A1 = TI == 0 ? 1 : 0, A2 = TI == 2 ? 1 : 2;
    __host__ __device__ static long stride(const Geom &g, int axis)
    {
        return axis == 0 ? 1L : axis == 1 ? (long)g.sx : (long)g.sx * g.sy;
    }
    __host__ __device__ static int extent(const Geom &g, int axis) { return axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz; }
    // offset of the cell t along TI in row (a, b)
    __host__ __device__ static long at(const Geom &g, int t, int a, int b)
    {
        int c[3];
        c[TI] = t;
        c[A1] = a;
        c[A2] = b;
        return g.at(c[0], c[1], c[2]);
    }
    __host__ __device__ static void xyz(int t, int a, int b, int &x, int &y, int &z)
    {
        int c[3];
        c[TI] = t;
        c[A1] = a;
        c[A2] = b;
        x = c[0];
        y = c[1];
        z = c[2];
    }
    __host__ __device__ static int n1(const Box3 &R) { return R.hi[A1] - R.lo[A1] + 1; }
    __host__ __device__ static int n2(const Box3 &R) { return R.hi[A2] - R.lo[A2] + 1; }
    __host__ __device__ static long rows(const Box3 &R) { return 2L * n1(R) * n2(R); }
    // slab s, row (a, b)
    __host__ __device__ static long row(const Box3 &R, int s, int a, int b)
    {
        return (long)s * n1(R) * n2(R) + (long)(b - R.lo[A2]) * n1(R) + (a - R.lo[A1]);
    }
    // The slab pairs this form handles: 1 to 3 cells across TI, the same row
    // range in both slabs, faces and cells of every row within the array
    // (lo >= 2 - h, hi <= n + h), and no face shared between the two slabs.
    __host__ __device__ static bool fits(const Geom &g, const Box3 *R)
    {
        for (int s = 0; s < 2; s++) {
            const int w = R[s].hi[TI] - R[s].lo[TI] + 1;
            if (w < 1 || w > 3) return false;
            for (int d = 0; d < 3; d++)
                if (R[s].lo[d] < 2 - g.h || R[s].hi[d] > extent(g, d) + g.h || R[s].hi[d] < R[s].lo[d]) return false;
        }
        for (int d : {A1, A2})
            if (R[0].lo[d] != R[1].lo[d] || R[0].hi[d] != R[1].hi[d]) return false;
        return g.nx >= 2 && g.ny >= 2 && g.nz >= 2 && R[0].hi[TI] + 1 < R[1].lo[TI];
    }
};

// the x slabs (TI = 0), kept for callers that name them
__host__ __device__ inline bool xface_fits(const Geom &g, const Box3 *R) { return SlabAxes<0>::fits(g, R); }

// Layout of the compact buffer: for each slab row r (slab-major, then b, then
// a) and value q < kXfVals, element q * rows + r.  Values, named for the x
// slabs (TI faces U, A1 faces V, A2 faces W): U faces ta-1..tb (4), V +a faces
// (3), W +b faces (3), P (3), the -a faces of the first row in a (3) and the
// -b faces of the first row in b (3).
constexpr int kXfVals = 19;
constexpr int kXfU = 0, kXfV = 4, kXfW = 7, kXfP = 10, kXfVb = 13, kXfWb = 16;
constexpr int kXfMaxW = 3;  // slab cells per row

template <typename T, int COLORING, int PY = 16, int PZ = 16, int TI = 0>
struct XFace {
    using Ax = SlabAxes<TI>;
    static constexpr int A1 = Ax::A1, A2 = Ax::A2;
    static constexpr int NT = PY * PZ;
    static constexpr int WY = PY + 4, WZ = PZ + 4, WROWS = WY * WZ;  // window rows J0-2..J0+PY+1, K0-2..K0+PZ+1
    static constexpr int MAXW = kXfMaxW;
    // LDS, element-major so neighbouring threads (rows) hit neighbouring words;
    // "U" holds the TI faces, "V" the A1 faces, "W" the A2 faces
    static constexpr int OFF_U = 0;                           // faces ta-1..tb
    static constexpr int OFF_V = OFF_U + (MAXW + 1) * WROWS;  // +a faces of ta..tb
    static constexpr int OFF_W = OFF_V + MAXW * WROWS;        // +b faces
    static constexpr int OFF_DR = OFF_W + MAXW * WROWS;       // red corrections
    static constexpr int OFF_DB = OFF_DR + MAXW * WROWS;      // black corrections
    static constexpr int OFF_P = OFF_DB + MAXW * WROWS;       // P of the patch rows
    static constexpr int LDS_ELEMS = OFF_P + MAXW * WROWS;
    static constexpr int LOADS = (WROWS + NT - 1) / NT;        // window rows per thread

    // patches of one slab; workgroups of launch 1
    __host__ __device__ static int patches_a(const Box3 &R) { return (Ax::n1(R) + PY - 1) / PY; }
    __host__ __device__ static int patches_b(const Box3 &R) { return (Ax::n2(R) + PZ - 1) / PZ; }
    __host__ __device__ static int blocks(const Box3 &R) { return 2 * patches_a(R) * patches_b(R); }

    struct Patch {
        Box3 r;      // the slab
        int slab;    // 0: lower, 1: upper
        int xa, nw;  // first cell along TI and cells per row
        int J0, K0;  // first patch row (a, b)
    };

    // the three face arrays by role: TI, A1, A2
    __host__ __device__ static T *FT(const XfArgs<T> &a) { return TI == 0 ? a.U : TI == 1 ? a.V : a.W; }
    __host__ __device__ static T *F1(const XfArgs<T> &a) { return A1 == 0 ? a.U : a.V; }
    __host__ __device__ static T *F2(const XfArgs<T> &a) { return A2 == 1 ? a.V : a.W; }
    __host__ __device__ static T cT(const HsCoef<T> &c) { return TI == 0 ? c.dxi : TI == 1 ? c.dyi : c.dzi; }
    __host__ __device__ static T c1(const HsCoef<T> &c) { return A1 == 0 ? c.dxi : c.dyi; }
    __host__ __device__ static T c2(const HsCoef<T> &c) { return A2 == 1 ? c.dyi : c.dzi; }
    // the divergence with the three terms added in x, y, z order
    __host__ __device__ static T div_xyz(T dT, T d1, T d2)
    {
        if (TI == 0) return (dT + d1) + d2;
        if (TI == 1) return (d1 + dT) + d2;
        return (d1 + d2) + dT;
    }

    __host__ __device__ static void patch(const XfArgs<T> &a, int blk, Patch &p)
    {
        const int npy = patches_a(a.R[0]), npz = patches_b(a.R[0]);
        p.slab = blk / (npy * npz);
        const int rest = blk % (npy * npz);
        p.r = p.slab ? a.R[1] : a.R[0];  // constant indices: no private array
        p.xa = p.r.lo[TI];
        p.nw = p.r.hi[TI] - p.r.lo[TI] + 1;
        p.J0 = p.r.lo[A1] + PY * (rest % npy);
        p.K0 = p.r.lo[A2] + PZ * (rest / npy);
    }

    __host__ __device__ static int wrow(const Patch &p, int j, int k) { return (j - p.J0 + 2) + WY * (k - p.K0 + 2); }
    __host__ __device__ static bool row_in(const Patch &p, int j, int k)
    {
        return j >= p.r.lo[A1] && j <= p.r.hi[A1] && k >= p.r.lo[A2] && k <= p.r.hi[A2];
    }
    __host__ __device__ static bool row_in_array(const Geom &g, int j, int k)
    {
        return j >= 1 - g.h && j <= Ax::extent(g, A1) + g.h && k >= 1 - g.h && k <= Ax::extent(g, A2) + g.h;
    }
    // cell t of row (j, k) is in the slab, visited, and of colour col
    __host__ __device__ static bool cell(const Patch &p, int t, int j, int k, int col)
    {
        int x, y, z;
        Ax::xyz(t, j, k, x, y, z);
        return ((t + j + k) & 1) == col && visited<COLORING>(p.r, x, y, z);
    }
    __host__ __device__ static bool in_d3_row(const Geom &g, int t, int j, int k)
    {
        int x, y, z;
        Ax::xyz(t, j, k, x, y, z);
        return in_d3(g, x, y, z);
    }
    __host__ __device__ static bool owner(const Patch &p, int j, int k)
    {
        return j >= p.J0 && j < p.J0 + PY && k >= p.K0 && k < p.K0 + PZ;
    }
    // Rows two or more rows inside their patch on every side are in no other
    // patch's window: launch 1 writes them in place; the others go through
    // the compact buffer.  ly, lz: the row's position in its patch.
    __host__ __device__ static bool inner(int ly, int lz) { return ly >= 2 && ly < PY - 2 && lz >= 2 && lz < PZ - 2; }

    // Phase L: the window's fields into LDS (TI faces where corrections are
    // formed, A1 faces one row further down in a, A2 faces one row further
    // down in b, P on the patch rows); every load of the thread is issued
    // before the first store
    __host__ __device__ static void load(const XfArgs<T> &a, const Patch &p, T *lds, int tid)
    {
        const Geom &g = a.g;
        const long st = Ax::stride(g, TI);
        const T *Ft = FT(a), *Fa = F1(a), *Fb = F2(a);
        T u[LOADS][MAXW + 1], v[LOADS][MAXW], w[LOADS][MAXW], pp[LOADS][MAXW];
        for (int l = 0; l < LOADS; l++) {
            const int e = tid + l * NT;
            const int j = p.J0 - 2 + e % WY, k = p.K0 - 2 + e / WY;
            const bool in = e < WROWS && row_in_array(g, j, k);
            const bool needU = in && j >= p.J0 - 1 && k >= p.K0 - 1;
            const bool needV = in && k >= p.K0 - 1;
            const bool needW = in && j >= p.J0 - 1;
            const bool needP = in && owner(p, j, k) && row_in(p, j, k);
            const long base = in ? Ax::at(g, p.xa, j, k) : 0;
            for (int q = 0; q <= MAXW; q++) u[l][q] = needU && q <= p.nw ? Ft[base + (q - 1) * st] : T(0);
            for (int q = 0; q < MAXW; q++) {
                v[l][q] = needV && q < p.nw ? Fa[base + q * st] : T(0);
                w[l][q] = needW && q < p.nw ? Fb[base + q * st] : T(0);
                pp[l][q] = needP && q < p.nw ? a.P[base + q * st] : T(0);
            }
        }
        for (int l = 0; l < LOADS; l++) {
            const int e = tid + l * NT;
            if (e >= WROWS) break;
            for (int q = 0; q <= MAXW; q++) lds[OFF_U + q * WROWS + e] = u[l][q];
            for (int q = 0; q < MAXW; q++) {
                lds[OFF_V + q * WROWS + e] = v[l][q];
                lds[OFF_W + q * WROWS + e] = w[l][q];
                lds[OFF_P + q * WROWS + e] = pp[l][q];
                lds[OFF_DR + q * WROWS + e] = T(0);
                lds[OFF_DB + q * WROWS + e] = T(0);
            }
        }
    }

    // Phase R: red corrections of rows J0-1..J0+PY+1, K0-1..K0+PZ+1
    __host__ __device__ static void red(const XfArgs<T> &a, const Patch &p, T *lds, int tid, double &acc)
    {
        const HsCoef<T> &c = a.c;
        const T ct = cT(c), ca = c1(c), cb = c2(c);
        constexpr int RY = PY + 3, RROWS = RY * (PZ + 3);
        for (int t = tid; t < RROWS; t += NT) {
            const int j = p.J0 - 1 + t % RY, k = p.K0 - 1 + t / RY;
            if (!row_in(p, j, k)) continue;
            const int w = wrow(p, j, k);
            for (int q = 0; q < p.nw; q++) {
                const int x = p.xa + q;
                if (!cell(p, x, j, k, 0)) continue;
                const T dT = ct * (-lds[OFF_U + q * WROWS + w] + lds[OFF_U + (q + 1) * WROWS + w]);
                const T d1 = ca * (-lds[OFF_V + q * WROWS + w - 1] + lds[OFF_V + q * WROWS + w]);
                const T d2 = cb * (-lds[OFF_W + q * WROWS + w - WY] + lds[OFF_W + q * WROWS + w]);
                const T div = div_xyz(dT, d1, d2);
                lds[OFF_DR + q * WROWS + w] = (-c.omega * div) * c.dl;
                if (owner(p, j, k) && in_d3_row(a.g, x, j, k)) acc += (double)(div * div);
            }
        }
    }

    // Phase B: black corrections of rows J0..J0+PY, K0..K0+PZ on the faces the
    // red cells left (each face has at most one red neighbour)
    __host__ __device__ static void black(const XfArgs<T> &a, const Patch &p, T *lds, int tid, double &acc)
    {
        const HsCoef<T> &c = a.c;
        const T ct = cT(c), ca = c1(c), cb = c2(c);
        constexpr int BY = PY + 1, BROWS = BY * (PZ + 1);
        for (int t = tid; t < BROWS; t += NT) {
            const int j = p.J0 + t % BY, k = p.K0 + t / BY;
            if (!row_in(p, j, k)) continue;
            const int w = wrow(p, j, k);
            for (int q = 0; q < p.nw; q++) {
                const int x = p.xa + q;
                if (!cell(p, x, j, k, 1)) continue;
                T Ul = lds[OFF_U + q * WROWS + w], Ur = lds[OFF_U + (q + 1) * WROWS + w];
                if (cell(p, x - 1, j, k, 0)) Ul += lds[OFF_DR + (q - 1) * WROWS + w] * ct;
                if (cell(p, x + 1, j, k, 0)) Ur -= lds[OFF_DR + (q + 1) * WROWS + w] * ct;
                T Vl = lds[OFF_V + q * WROWS + w - 1], Vr = lds[OFF_V + q * WROWS + w];
                if (cell(p, x, j - 1, k, 0)) Vl += lds[OFF_DR + q * WROWS + w - 1] * ca;
                if (cell(p, x, j + 1, k, 0)) Vr -= lds[OFF_DR + q * WROWS + w + 1] * ca;
                T Wl = lds[OFF_W + q * WROWS + w - WY], Wr = lds[OFF_W + q * WROWS + w];
                if (cell(p, x, j, k - 1, 0)) Wl += lds[OFF_DR + q * WROWS + w - WY] * cb;
                if (cell(p, x, j, k + 1, 0)) Wr -= lds[OFF_DR + q * WROWS + w + WY] * cb;
                const T dT = ct * (-Ul + Ur);
                const T d1 = ca * (-Vl + Vr);
                const T d2 = cb * (-Wl + Wr);
                const T div = div_xyz(dT, d1, d2);
                lds[OFF_DB + q * WROWS + w] = (-c.omega * div) * c.dl;
                if (owner(p, j, k) && in_d3_row(a.g, x, j, k)) acc += (double)(div * div);
            }
        }
    }

    __host__ __device__ static T dpc(const T *lds, int col, int q, int w)
    {
        return lds[(col ? OFF_DB : OFF_DR) + q * WROWS + w];
    }

    // Phase F: final values of the thread's own row, in place or into the
    // compact buffer, the corrections of each face applied red first, then black
    __host__ __device__ static void finals(const XfArgs<T> &a, const Patch &p, const T *lds, int tid)
    {
        const HsCoef<T> &c = a.c;
        const T ct = cT(c), ca = c1(c), cb = c2(c);
        const Geom &g = a.g;
        const long st = Ax::stride(g, TI);
        T *Ft = FT(a), *Fa = F1(a), *Fb = F2(a);
        const int j = p.J0 + tid % PY, k = p.K0 + tid / PY;
        if (!row_in(p, j, k)) return;
        const bool in_place = inner(tid % PY, tid / PY);
        const int w = wrow(p, j, k);
        const long rows = Ax::rows(p.r), r = Ax::row(p.r, p.slab, j, k);
        const long base = Ax::at(g, p.xa, j, k);
        for (int q = 0; q <= p.nw; q++) {  // face between cells xa-1+q and xa+q
            T u = lds[OFF_U + q * WROWS + w];
            for (int col = 0; col < 2; col++) {
                if (cell(p, p.xa - 1 + q, j, k, col)) u += dpc(lds, col, q - 1, w) * ct;
                if (cell(p, p.xa + q, j, k, col)) u -= dpc(lds, col, q, w) * ct;
            }
            if (in_place)
                Ft[base + (q - 1) * st] = u;
            else
                a.buf[(kXfU + q) * rows + r] = u;
        }
        for (int q = 0; q < p.nw; q++) {
            const int x = p.xa + q;
            T pv = lds[OFF_P + q * WROWS + w];
            for (int col = 0; col < 2; col++)
                if (cell(p, x, j, k, col)) pv += dpc(lds, col, q, w) * a.dti;
            T v = lds[OFF_V + q * WROWS + w];  // +a face: this cell (+), the next row's (-)
            for (int col = 0; col < 2; col++) {
                if (cell(p, x, j, k, col)) v += dpc(lds, col, q, w) * ca;
                if (cell(p, x, j + 1, k, col)) v -= dpc(lds, col, q, w + 1) * ca;
            }
            T wv = lds[OFF_W + q * WROWS + w];
            for (int col = 0; col < 2; col++) {
                if (cell(p, x, j, k, col)) wv += dpc(lds, col, q, w) * cb;
                if (cell(p, x, j, k + 1, col)) wv -= dpc(lds, col, q, w + WY) * cb;
            }
            if (in_place) {
                a.P[base + q * st] = pv;
                Fa[base + q * st] = v;
                Fb[base + q * st] = wv;
                continue;
            }
            a.buf[(kXfP + q) * rows + r] = pv;
            a.buf[(kXfV + q) * rows + r] = v;
            a.buf[(kXfW + q) * rows + r] = wv;
            if (j == p.r.lo[A1]) {  // -a face of the first row: only this cell
                T vb = lds[OFF_V + q * WROWS + w - 1];
                for (int col = 0; col < 2; col++)
                    if (cell(p, x, j, k, col)) vb -= dpc(lds, col, q, w) * ca;
                a.buf[(kXfVb + q) * rows + r] = vb;
            }
            if (k == p.r.lo[A2]) {
                T wb = lds[OFF_W + q * WROWS + w - WY];
                for (int col = 0; col < 2; col++)
                    if (cell(p, x, j, k, col)) wb -= dpc(lds, col, q, w) * cb;
                a.buf[(kXfWb + q) * rows + r] = wb;
            }
        }
    }

    // Launch 2: row r of the compact buffer into the fields (rows launch 1
    // wrote in place are skipped)
    __host__ __device__ static void commit_row(const Geom &g, const Box3 &r0, const Box3 &r1, T *U, T *V, T *W, T *P,
                                               const T *buf, long r)
    {
        const long rows = Ax::rows(r0), per = rows / 2;
        const int slab = (int)(r / per);
        const long rr = r % per;
        const int n1 = Ax::n1(r0);
        const int j = r0.lo[A1] + (int)(rr % n1), k = r0.lo[A2] + (int)(rr / n1);
        if (inner((j - r0.lo[A1]) % PY, (k - r0.lo[A2]) % PZ)) return;
        const Box3 R = slab ? r1 : r0;
        const int xa = R.lo[TI], nw = R.hi[TI] - R.lo[TI] + 1;
        const long base = Ax::at(g, xa, j, k), st = Ax::stride(g, TI), sa = Ax::stride(g, A1), sb = Ax::stride(g, A2);
        T *Ft = TI == 0 ? U : TI == 1 ? V : W, *Fa = A1 == 0 ? U : V, *Fb = A2 == 1 ? V : W;
        for (int q = 0; q <= nw; q++) Ft[base + (q - 1) * st] = buf[(kXfU + q) * rows + r];
        for (int q = 0; q < nw; q++) {
            Fa[base + q * st] = buf[(kXfV + q) * rows + r];
            Fb[base + q * st] = buf[(kXfW + q) * rows + r];
            P[base + q * st] = buf[(kXfP + q) * rows + r];
            if (j == R.lo[A1]) Fa[base + q * st - sa] = buf[(kXfVb + q) * rows + r];
            if (k == R.lo[A2]) Fb[base + q * st - sb] = buf[(kXfWb + q) * rows + r];
        }
    }

    // One workgroup of launch 1 for a host harness: every thread's phase,
    // phase by phase.  Returns the div^2 sum of the patch's cells in d3.
    static double emulate_block(const XfArgs<T> &a, T *lds, int blk)
    {
        Patch p;
        patch(a, blk, p);
        double acc = 0.0;
        for (int t = 0; t < NT; t++) load(a, p, lds, t);
        for (int t = 0; t < NT; t++) red(a, p, lds, t, acc);
        for (int t = 0; t < NT; t++) black(a, p, lds, t, acc);
        for (int t = 0; t < NT; t++) finals(a, p, lds, t);
        return acc;
    }
};

// Patch edge by element type: 32 x 32 rows (1024 threads, 98 KB of LDS) for
// FP32, 16 x 16 (256 threads, 61 KB, two workgroups per CU) for FP64; LDS
// holds 19 values for each of the (N + 4)^2 window rows.  Bigger patches
// re-read fewer window rows and send fewer rows through the compact buffer.
template <typename T>
struct XfPatch {
    static constexpr int N = sizeof(T) == 4 ? 32 : 16;
};
template <typename T, int COLORING, int TI = 0>
using XFaceOf = XFace<T, COLORING, XfPatch<T>::N, XfPatch<T>::N, TI>;

}  // namespace lcs
