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
    Box3 R[2];               // the x- and x+ slabs (rows (y, z) in [2, ny] x [2, nz])
    HsCoef<T> c;
    T dti;
    T *buf;  // compact final values (xface_commit_row)
};

// The slabs this form handles: rows (y, z) = [2, ny] x [2, nz], 1 to 3 cells
// wide, x- below x+, and no face shared between them (nx >= 2).
__host__ __device__ inline bool xface_fits(const Geom &g, const Box3 *R)
{
    for (int s = 0; s < 2; s++) {
        const int w = R[s].hi[0] - R[s].lo[0] + 1;
        if (w < 1 || w > 3 || R[s].lo[1] != 2 || R[s].hi[1] != g.ny || R[s].lo[2] != 2 || R[s].hi[2] != g.nz)
            return false;
    }
    return g.nx >= 2 && g.ny >= 2 && g.nz >= 2 && R[0].hi[0] + 1 < R[1].lo[0];
}

// Layout of the compact buffer: for each slab row r (slab-major, then z, then
// y) and value q < kXfVals, element q * rows + r.  Values: U faces xa-1..xb
// (4), V +y faces (3), W +z faces (3), P (3), the -y faces of the first slab
// row in y (3) and the -z faces of the first row in z (3).
constexpr int kXfVals = 19;
constexpr int kXfU = 0, kXfV = 4, kXfW = 7, kXfP = 10, kXfVb = 13, kXfWb = 16;
constexpr int kXfMaxW = 3;  // slab cells per row

__host__ __device__ inline long xf_rows(const Geom &g) { return 2L * (g.ny - 1) * (g.nz - 1); }

// slab s, row (j, k) of the slab region (j in [2, ny], k in [2, nz])
__host__ __device__ inline long xf_row(const Geom &g, int s, int j, int k)
{
    return (long)s * (g.ny - 1) * (g.nz - 1) + (long)(k - 2) * (g.ny - 1) + (j - 2);
}

template <typename T, int COLORING, int PY = 16, int PZ = 16>
struct XFace {
    static constexpr int NT = PY * PZ;
    static constexpr int WY = PY + 4, WZ = PZ + 4, WROWS = WY * WZ;  // window rows J0-2..J0+PY+1, K0-2..K0+PZ+1
    static constexpr int MAXW = kXfMaxW;
    // LDS, element-major so neighbouring threads (rows) hit neighbouring words
    static constexpr int OFF_U = 0;                           // faces xa-1..xb
    static constexpr int OFF_V = OFF_U + (MAXW + 1) * WROWS;  // +y faces of xa..xb
    static constexpr int OFF_W = OFF_V + MAXW * WROWS;        // +z faces
    static constexpr int OFF_DR = OFF_W + MAXW * WROWS;       // red corrections
    static constexpr int OFF_DB = OFF_DR + MAXW * WROWS;      // black corrections
    static constexpr int OFF_P = OFF_DB + MAXW * WROWS;       // P of the patch rows
    static constexpr int LDS_ELEMS = OFF_P + MAXW * WROWS;
    static constexpr int LOADS = (WROWS + NT - 1) / NT;        // window rows per thread

    // patches of one slab; workgroups of launch 1
    __host__ __device__ static int patches_y(const Geom &g) { return (g.ny - 1 + PY - 1) / PY; }
    __host__ __device__ static int patches_z(const Geom &g) { return (g.nz - 1 + PZ - 1) / PZ; }
    __host__ __device__ static int blocks(const Geom &g) { return 2 * patches_y(g) * patches_z(g); }

    struct Patch {
        Box3 r;      // the slab
        int slab;    // 0: x-, 1: x+
        int xa, nw;  // first cell and cells per row
        int J0, K0;  // first patch row
    };

    __host__ __device__ static void patch(const XfArgs<T> &a, int blk, Patch &p)
    {
        const int npy = patches_y(a.g), npz = patches_z(a.g);
        p.slab = blk / (npy * npz);
        const int rest = blk % (npy * npz);
        p.r = p.slab ? a.R[1] : a.R[0];  // constant indices: no private array
        p.xa = p.r.lo[0];
        p.nw = p.r.hi[0] - p.r.lo[0] + 1;
        p.J0 = p.r.lo[1] + PY * (rest % npy);
        p.K0 = p.r.lo[2] + PZ * (rest / npy);
    }

    __host__ __device__ static int wrow(const Patch &p, int j, int k) { return (j - p.J0 + 2) + WY * (k - p.K0 + 2); }
    __host__ __device__ static bool row_in(const Patch &p, int j, int k)
    {
        return j >= p.r.lo[1] && j <= p.r.hi[1] && k >= p.r.lo[2] && k <= p.r.hi[2];
    }
    __host__ __device__ static bool row_in_array(const Geom &g, int j, int k)
    {
        return j >= 1 - g.h && j <= g.ny + g.h && k >= 1 - g.h && k <= g.nz + g.h;
    }
    // cell x of row (j, k) is in the slab, visited, and of colour col
    __host__ __device__ static bool cell(const Patch &p, int x, int j, int k, int col)
    {
        return ((x + j + k) & 1) == col && visited<COLORING>(p.r, x, j, k);
    }
    __host__ __device__ static bool owner(const Patch &p, int j, int k)
    {
        return j >= p.J0 && j < p.J0 + PY && k >= p.K0 && k < p.K0 + PZ;
    }
    // Rows two or more rows inside their patch on every side are in no other
    // patch's window: launch 1 writes them in place; the others go through
    // the compact buffer.  ly, lz: the row's position in its patch.
    __host__ __device__ static bool inner(int ly, int lz) { return ly >= 2 && ly < PY - 2 && lz >= 2 && lz < PZ - 2; }

    // Phase L: the window's fields into LDS (U where corrections are formed,
    // V one row further down in y, W one row further down in z, P on the
    // patch rows); every load of the thread is issued before the first store
    __host__ __device__ static void load(const XfArgs<T> &a, const Patch &p, T *lds, int tid)
    {
        const Geom &g = a.g;
        T u[LOADS][MAXW + 1], v[LOADS][MAXW], w[LOADS][MAXW], pp[LOADS][MAXW];
        for (int l = 0; l < LOADS; l++) {
            const int e = tid + l * NT;
            const int j = p.J0 - 2 + e % WY, k = p.K0 - 2 + e / WY;
            const bool in = e < WROWS && row_in_array(g, j, k);
            const bool needU = in && j >= p.J0 - 1 && k >= p.K0 - 1;
            const bool needV = in && k >= p.K0 - 1;
            const bool needW = in && j >= p.J0 - 1;
            const bool needP = in && owner(p, j, k) && row_in(p, j, k);
            const long base = in ? g.at(p.xa, j, k) : 0;
            for (int q = 0; q <= MAXW; q++) u[l][q] = needU && q <= p.nw ? a.U[base + q - 1] : T(0);
            for (int q = 0; q < MAXW; q++) {
                v[l][q] = needV && q < p.nw ? a.V[base + q] : T(0);
                w[l][q] = needW && q < p.nw ? a.W[base + q] : T(0);
                pp[l][q] = needP && q < p.nw ? a.P[base + q] : T(0);
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
        constexpr int RY = PY + 3, RROWS = RY * (PZ + 3);
        for (int t = tid; t < RROWS; t += NT) {
            const int j = p.J0 - 1 + t % RY, k = p.K0 - 1 + t / RY;
            if (!row_in(p, j, k)) continue;
            const int w = wrow(p, j, k);
            for (int q = 0; q < p.nw; q++) {
                const int x = p.xa + q;
                if (!cell(p, x, j, k, 0)) continue;
                const T dudx = c.dxi * (-lds[OFF_U + q * WROWS + w] + lds[OFF_U + (q + 1) * WROWS + w]);
                const T dvdy = c.dyi * (-lds[OFF_V + q * WROWS + w - 1] + lds[OFF_V + q * WROWS + w]);
                const T dwdz = c.dzi * (-lds[OFF_W + q * WROWS + w - WY] + lds[OFF_W + q * WROWS + w]);
                const T div = (dudx + dvdy) + dwdz;
                lds[OFF_DR + q * WROWS + w] = (-c.omega * div) * c.dl;
                if (owner(p, j, k) && in_d3(a.g, x, j, k)) acc += (double)(div * div);
            }
        }
    }

    // Phase B: black corrections of rows J0..J0+PY, K0..K0+PZ on the faces the
    // red cells left (each face has at most one red neighbour)
    __host__ __device__ static void black(const XfArgs<T> &a, const Patch &p, T *lds, int tid, double &acc)
    {
        const HsCoef<T> &c = a.c;
        constexpr int BY = PY + 1, BROWS = BY * (PZ + 1);
        for (int t = tid; t < BROWS; t += NT) {
            const int j = p.J0 + t % BY, k = p.K0 + t / BY;
            if (!row_in(p, j, k)) continue;
            const int w = wrow(p, j, k);
            for (int q = 0; q < p.nw; q++) {
                const int x = p.xa + q;
                if (!cell(p, x, j, k, 1)) continue;
                T Ul = lds[OFF_U + q * WROWS + w], Ur = lds[OFF_U + (q + 1) * WROWS + w];
                if (cell(p, x - 1, j, k, 0)) Ul += lds[OFF_DR + (q - 1) * WROWS + w] * c.dxi;
                if (cell(p, x + 1, j, k, 0)) Ur -= lds[OFF_DR + (q + 1) * WROWS + w] * c.dxi;
                T Vl = lds[OFF_V + q * WROWS + w - 1], Vr = lds[OFF_V + q * WROWS + w];
                if (cell(p, x, j - 1, k, 0)) Vl += lds[OFF_DR + q * WROWS + w - 1] * c.dyi;
                if (cell(p, x, j + 1, k, 0)) Vr -= lds[OFF_DR + q * WROWS + w + 1] * c.dyi;
                T Wl = lds[OFF_W + q * WROWS + w - WY], Wr = lds[OFF_W + q * WROWS + w];
                if (cell(p, x, j, k - 1, 0)) Wl += lds[OFF_DR + q * WROWS + w - WY] * c.dzi;
                if (cell(p, x, j, k + 1, 0)) Wr -= lds[OFF_DR + q * WROWS + w + WY] * c.dzi;
                const T dudx = c.dxi * (-Ul + Ur);
                const T dvdy = c.dyi * (-Vl + Vr);
                const T dwdz = c.dzi * (-Wl + Wr);
                const T div = (dudx + dvdy) + dwdz;
                lds[OFF_DB + q * WROWS + w] = (-c.omega * div) * c.dl;
                if (owner(p, j, k) && in_d3(a.g, x, j, k)) acc += (double)(div * div);
            }
        }
    }

    __host__ __device__ static T dpc(const T *lds, int col, int q, int w)
    {
        return lds[(col ? OFF_DB : OFF_DR) + q * WROWS + w];
    }

    // Phase F: final values of the thread's own row into the compact buffer,
    // the corrections of each face applied red first, then black
    __host__ __device__ static void finals(const XfArgs<T> &a, const Patch &p, const T *lds, int tid)
    {
        const HsCoef<T> &c = a.c;
        const Geom &g = a.g;
        const int j = p.J0 + tid % PY, k = p.K0 + tid / PY;
        if (!row_in(p, j, k)) return;
        const bool in_place = inner(tid % PY, tid / PY);
        const int w = wrow(p, j, k);
        const long rows = xf_rows(g), r = xf_row(g, p.slab, j, k);
        const long base = g.at(p.xa, j, k);
        for (int q = 0; q <= p.nw; q++) {  // face between cells xa-1+q and xa+q
            T u = lds[OFF_U + q * WROWS + w];
            for (int col = 0; col < 2; col++) {
                if (cell(p, p.xa - 1 + q, j, k, col)) u += dpc(lds, col, q - 1, w) * c.dxi;
                if (cell(p, p.xa + q, j, k, col)) u -= dpc(lds, col, q, w) * c.dxi;
            }
            if (in_place)
                a.U[base + q - 1] = u;
            else
                a.buf[(kXfU + q) * rows + r] = u;
        }
        for (int q = 0; q < p.nw; q++) {
            const int x = p.xa + q;
            T pv = lds[OFF_P + q * WROWS + w];
            for (int col = 0; col < 2; col++)
                if (cell(p, x, j, k, col)) pv += dpc(lds, col, q, w) * a.dti;
            T v = lds[OFF_V + q * WROWS + w];  // +y face: this cell (+), the cell above (-)
            for (int col = 0; col < 2; col++) {
                if (cell(p, x, j, k, col)) v += dpc(lds, col, q, w) * c.dyi;
                if (cell(p, x, j + 1, k, col)) v -= dpc(lds, col, q, w + 1) * c.dyi;
            }
            T wv = lds[OFF_W + q * WROWS + w];
            for (int col = 0; col < 2; col++) {
                if (cell(p, x, j, k, col)) wv += dpc(lds, col, q, w) * c.dzi;
                if (cell(p, x, j, k + 1, col)) wv -= dpc(lds, col, q, w + WY) * c.dzi;
            }
            if (in_place) {
                a.P[base + q] = pv;
                a.V[base + q] = v;
                a.W[base + q] = wv;
                continue;
            }
            a.buf[(kXfP + q) * rows + r] = pv;
            a.buf[(kXfV + q) * rows + r] = v;
            a.buf[(kXfW + q) * rows + r] = wv;
            if (j == p.r.lo[1]) {  // -y face of the first slab row: only this cell
                T vb = lds[OFF_V + q * WROWS + w - 1];
                for (int col = 0; col < 2; col++)
                    if (cell(p, x, j, k, col)) vb -= dpc(lds, col, q, w) * c.dyi;
                a.buf[(kXfVb + q) * rows + r] = vb;
            }
            if (k == p.r.lo[2]) {
                T wb = lds[OFF_W + q * WROWS + w - WY];
                for (int col = 0; col < 2; col++)
                    if (cell(p, x, j, k, col)) wb -= dpc(lds, col, q, w) * c.dzi;
                a.buf[(kXfWb + q) * rows + r] = wb;
            }
        }
    }

    // Launch 2: row r of the compact buffer into the fields (rows launch 1
    // wrote in place are skipped)
    __host__ __device__ static void commit_row(const Geom &g, const Box3 &r0, const Box3 &r1, T *U, T *V, T *W, T *P,
                                               const T *buf, long r)
    {
        const long rows = xf_rows(g), per = (long)(g.ny - 1) * (g.nz - 1);
        const int slab = (int)(r / per);
        const long rr = r % per;
        const int j = 2 + (int)(rr % (g.ny - 1)), k = 2 + (int)(rr / (g.ny - 1));
        if (inner((j - 2) % PY, (k - 2) % PZ)) return;
        const Box3 R = slab ? r1 : r0;
        const int xa = R.lo[0], nw = R.hi[0] - R.lo[0] + 1;
        const long base = g.at(xa, j, k), sy = g.sx, sz = g.sx * g.sy;
        for (int q = 0; q <= nw; q++) U[base + q - 1] = buf[(kXfU + q) * rows + r];
        for (int q = 0; q < nw; q++) {
            V[base + q] = buf[(kXfV + q) * rows + r];
            W[base + q] = buf[(kXfW + q) * rows + r];
            P[base + q] = buf[(kXfP + q) * rows + r];
            if (j == 2) V[base + q - sy] = buf[(kXfVb + q) * rows + r];
            if (k == 2) W[base + q - sz] = buf[(kXfWb + q) * rows + r];
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
// FP32, 16 x 16 (256 threads, 51 KB) for FP64.  Bigger patches re-read fewer
// window rows and send fewer rows through the compact buffer.
template <typename T>
struct XfPatch {
    static constexpr int N = sizeof(T) == 4 ? 32 : 16;
};
template <typename T, int COLORING>
using XFaceOf = XFace<T, COLORING, XfPatch<T>::N, XfPatch<T>::N>;

}  // namespace lcs
