/*
 * lcs_oracle.c -- CPU restatement of the LCS.jl per-time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  It is never part of the product path.
 *
 * It restates the reference algorithm phase by phase, keeping the reference's
 * unfused phase structure, loop domains, pass order and floating-point
 * expression order (Julia evaluates a+b+c and a*b*c left to right and does not
 * contract to FMA; this file is built with -ffp-contract=off).  Each function
 * cites the reference file:line it follows.
 *
 * It is compiled twice: REAL=double (the reference's float type, FP64) and
 * REAL=float (the FP32 extension).  FP32 semantics: every scalar coefficient is
 * computed in double and rounded once to float; per-cell arithmetic is in
 * float; reductions accumulate in double.
 *
 * Layout (reference src/flows/buffer.jl:52-55, lib/offsets/src/Offsets.jl:23-43):
 * column-major, x fastest, halo h on every side.  Domain index d in
 * [1-h, n+h] lives at 0-based array index d-1+h.
 *
 * Threads: loops run over the last axis with OpenMP, like
 * lib/parallel/src/foraxes.jl:16-23.  Reductions are deterministic: per-plane
 * partial sums are combined in plane order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Thread count of the OpenMP loops (the CPU baseline times 1 thread and all). */
int o_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

#ifndef REAL
#define REAL double
#endif
typedef REAL real;
#define R(x) ((real)(x))

/* src/LCS.jl:16-17 */
static const double DOMAIN_ORIGIN = 0.0;
static const double DOMAIN_LENGTH = 6.283185307179586; /* 2π */

typedef struct {
    int nx, ny, nz, h;
    long sx, sy;
} grid_t;

static grid_t mkgrid(int nx, int ny, int nz, int h)
{
    grid_t g = {nx, ny, nz, h, nx + 2 * h, ny + 2 * h};
    return g;
}

/* domain index (1-based, may be <= 0) -> linear array index */
#define AT(g, i, j, k)                                                                         \
    ((size_t)((i) + (g).h - 1) +                                                               \
     (size_t)(g).sx * ((size_t)((j) + (g).h - 1) + (size_t)(g).sy * (size_t)((k) + (g).h - 1)))

static inline int mod1(int i, int n)
{
    int r = (i - 1) % n;
    if (r < 0) r += n;
    return r + 1;
}

/* ------------------------------------------------------------------------- */
/* Stencils: src/diff.jl:1-31                                                */
/* ------------------------------------------------------------------------- */
static inline real interp_o2(real a, real b) { return R(0.5) * (a + b); }
static inline real interp_o4(real a, real b, real c, real d)
{
    return R(1.0 / 16.0) * (((-a + R(9) * b) + R(9) * c) - d);
}
static inline real diff1_o2(real a, real b) { return -a + b; }
static inline real diff2_o4(real a, real b, real c, real d, real e)
{
    return R(1.0 / 12.0) * ((((-a + R(16) * b) - R(30) * c) + R(16) * d) - e);
}
static inline real sum_flux_o4(real j1a, real j1b, real j3a, real j3b)
{
    return R(9.0 / 8.0) * (-j1a + j1b) - R((1.0 / 8.0) * (1.0 / 3.0)) * (-j3a + j3b);
}

static double spacing(int n) { return DOMAIN_LENGTH / (double)n; } /* src/config.jl:89 */

/* ------------------------------------------------------------------------- */
/* Periodic ghost fill, single process: lib/topologies/src/halo.jl:138-159,   */
/* index maps :298-347, directions lib/topologies/src/topology.jl:36-66.      */
/* Result: A[i] = A[mod1(i, n)] on the ghost set of width r.  FaceHalo (full=0)*/
/* fills cells with exactly one axis outside 1..n; FullHalo (full=1) fills    */
/* all 26 directions.                                                          */
/* ------------------------------------------------------------------------- */
void o_ghost_fill(real *A, int nx, int ny, int nz, int h, int r, int full)
{
    grid_t g = mkgrid(nx, ny, nz, h);
#pragma omp parallel for schedule(static)
    for (int k = 1 - r; k <= nz + r; k++) {
        int ok = (k < 1 || k > nz);
        for (int j = 1 - r; j <= ny + r; j++) {
            int oj = (j < 1 || j > ny);
            for (int i = 1 - r; i <= nx + r; i++) {
                int oi = (i < 1 || i > nx);
                int out = oi + oj + ok;
                if (out == 0 || (!full && out > 1)) continue;
                A[AT(g, i, j, k)] = A[AT(g, mod1(i, nx), mod1(j, ny), mod1(k, nz))];
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Flow RHS: src/flows/rhs.jl:44-252.                                          */
/* Component c (0=x,1=y,2=z) of velocity Q.  First pass over domain1 =        */
/* (-1:n+1)^3 writes dQdt (domain2 only) and the six flux scratches; second   */
/* pass over domain2 = (1:n)^3 subtracts the flux divergence.                 */
/* ------------------------------------------------------------------------- */
static void rhs_component(int c, real *dQ, const real *U, const real *V, const real *W,
                          const real *P, real *J[6], grid_t g, double Re)
{
    const real *Vel[3] = {U, V, W};
    const real *Q = Vel[c];
    const long st[3] = {1, g.sx, g.sx * g.sy};
    const real di[3] = {R(1.0 / spacing(g.nx)), R(1.0 / spacing(g.ny)), R(1.0 / spacing(g.nz))};
    const real Rei = R(1.0 / Re);
    const long sc = st[c];

#pragma omp parallel for schedule(static)
    for (int k = -1; k <= g.nz + 1; k++)
        for (int j = -1; j <= g.ny + 1; j++)
            for (int i = -1; i <= g.nx + 1; i++) {
                size_t p = AT(g, i, j, k);
                /* advecting velocities interpolated along axis c: rhs.jl:106-108 */
                real ac[3];
                for (int a = 0; a < 3; a++) {
                    const real *A = Vel[a];
                    ac[a] = interp_o4(A[p - sc], A[p], A[p + sc], A[p + 2 * sc]);
                }
                /* transported component interpolated along each axis: rhs.jl:110-116 */
                real q1[3], q3[3];
                for (int d = 0; d < 3; d++) {
                    long s = st[d];
                    q1[d] = interp_o2(Q[p], Q[p + s]);
                    q3[d] = interp_o2(Q[p - s], Q[p + 2 * s]);
                }
                if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) {
                    real d2[3];
                    for (int d = 0; d < 3; d++) {
                        long s = st[d];
                        d2[d] = diff2_o4(Q[p - 2 * s], Q[p - s], Q[p], Q[p + s], Q[p + 2 * s]);
                    }
                    real diff = ((di[0] * di[0] * d2[0] + di[1] * di[1] * d2[1]) +
                                 di[2] * di[2] * d2[2]) *
                                Rei;
                    real pres = di[c] * diff1_o2(P[p], P[p + sc]);
                    size_t o = (size_t)(i - 1) + (size_t)g.nx * ((size_t)(j - 1) + (size_t)g.ny * (size_t)(k - 1));
                    dQ[o] = diff - pres;
                }
                for (int d = 0; d < 3; d++) {
                    J[d][p] = ac[d] * q1[d];
                    J[3 + d][p] = ac[d] * q3[d];
                }
            }

#pragma omp parallel for schedule(static)
    for (int k = 1; k <= g.nz; k++)
        for (int j = 1; j <= g.ny; j++)
            for (int i = 1; i <= g.nx; i++) {
                size_t p = AT(g, i, j, k);
                real f[3];
                for (int d = 0; d < 3; d++) {
                    long s = st[d];
                    f[d] = sum_flux_o4(J[d][p - s], J[d][p], J[3 + d][p - 2 * s], J[3 + d][p + s]);
                }
                real advect = (di[0] * f[0] + di[1] * f[1]) + di[2] * f[2];
                size_t o = (size_t)(i - 1) + (size_t)g.nx * ((size_t)(j - 1) + (size_t)g.ny * (size_t)(k - 1));
                dQ[o] -= advect;
            }
}

/* rhs!: src/flows/rhs.jl:59-81 -- dU, dV, dW are interior-only (nx*ny*nz). */
void o_rhs(int nx, int ny, int nz, int h, double Re, const real *U, const real *V, const real *W,
           const real *P, real *dU, real *dV, real *dW)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    size_t len = (size_t)g.sx * g.sy * (size_t)(nz + 2 * h);
    real *J[6];
    for (int a = 0; a < 6; a++) J[a] = (real *)calloc(len, sizeof(real));
    rhs_component(0, dU, U, V, W, P, J, g, Re);
    rhs_component(1, dV, U, V, W, P, J, g, Re);
    rhs_component(2, dW, U, V, W, P, J, g, Re);
    for (int a = 0; a < 6; a++) free(J[a]);
}

/* update!: src/flows/update.jl:23-54.  U = U0 + dUdt*dt on (1:n)^3. */
void o_update(int nx, int ny, int nz, int h, double dt, real *U, real *V, real *W, const real *U0,
              const real *V0, const real *W0, const real *dU, const real *dV, const real *dW)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    real rdt = R(dt);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t o = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
                size_t p = AT(g, i, j, k);
                real du = dU[o] * rdt, dv = dV[o] * rdt, dw = dW[o] * rdt;
                real u = U0[p], v = V0[p], w = W0[p];
                U[p] = u + du;
                V[p] = v + dv;
                W[p] = w + dw;
            }
}

/* ------------------------------------------------------------------------- */
/* HSMAC: src/flows/couples/hsmac.jl:1-201                                     */
/* ------------------------------------------------------------------------- */
typedef struct {
    real dti, dxi, dyi, dzi, dl, omega;
} hsmac_coef;

/* hsmac_cellwise!: hsmac.jl:1-35 */
static inline void hsmac_cell(grid_t g, int i, int j, int k, real *U, real *V, real *W, real *P,
                              real *Div2, hsmac_coef c)
{
    size_t p = AT(g, i, j, k);
    const long sy = g.sx, sz = g.sx * g.sy;
    real dudx = c.dxi * diff1_o2(U[p - 1], U[p]);
    real dvdy = c.dyi * diff1_o2(V[p - sy], V[p]);
    real dwdz = c.dzi * diff1_o2(W[p - sz], W[p]);
    real div = (dudx + dvdy) + dwdz;
    real dp = (-c.omega * div) * c.dl;
    Div2[p] = div * div;
    P[p] += dp * c.dti;
    U[p - 1] -= dp * c.dxi;
    U[p] += dp * c.dxi;
    V[p - sy] -= dp * c.dyi;
    V[p] += dp * c.dyi;
    W[p - sz] -= dp * c.dzi;
    W[p] += dp * c.dzi;
}

enum { COLORING_RED_BLACK_FAST = 0, COLORING_RED_BLACK = 1 };

/* One colour pass over the box [lo, hi] (domain coords).
 * red-black-fast (lib/parallel/src/foraxes.jl:95-107): cells with
 *   (i+j+k) & 1 == color (two's complement parity, valid for negatives).
 * red-black (foraxes.jl:84-88): cells with (i+j+k) % 2 == color using the
 *   truncating remainder, so cells with a negative odd index sum are never
 *   visited -- the reference's own behaviour, reproduced. */
static inline int colour_hit(int i, int j, int k, int color, int coloring)
{
    int s = i + j + k;
    return (coloring == COLORING_RED_BLACK) ? (s % 2 == color) : ((s & 1) == color);
}

static void colour_pass(grid_t g, const int lo[3], const int hi[3], int color, int coloring,
                        real *U, real *V, real *W, real *P, real *Div2, hsmac_coef c)
{
#pragma omp parallel for schedule(static)
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++)
                if (colour_hit(i, j, k, color, coloring)) hsmac_cell(g, i, j, k, U, V, W, P, Div2, c);
}

/* Cells one colour pass visits (lib/parallel/test/runtests.jl:73-113 KAT):
 * mask is a (hi-lo+1)^3 column-major int array set to 1 where visited. */
void o_colour_mask(const int *lo, const int *hi, int color, int coloring, int *mask)
{
    int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++)
                mask[(i - lo[0]) + (size_t)nx * ((j - lo[1]) + (size_t)ny * (k - lo[2]))] =
                    colour_hit(i, j, k, color, coloring);
}

/* One colour pass of hsmac_cellwise! over a box (domain coords), exported so
 * tests can replay the reference's pass order step by step.  Div2 receives
 * div^2 of visited cells. */
void o_hsmac_pass(int nx, int ny, int nz, int h, const int *lo, const int *hi, int color, int coloring,
                  double dt, double omega, real *U, real *V, real *W, real *P, real *Div2)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    real dxi = R(1.0 / spacing(nx)), dyi = R(1.0 / spacing(ny)), dzi = R(1.0 / spacing(nz));
    hsmac_coef c;
    c.dti = R(1.0 / dt);
    c.dxi = dxi;
    c.dyi = dyi;
    c.dzi = dzi;
    c.dl = R(1) / (R(2) * ((dxi * dxi + dyi * dyi) + dzi * dzi));
    c.omega = R(omega);
    colour_pass(g, lo, hi, color, coloring, U, V, W, P, Div2, c);
}

/* foraxes_colored: colour 0 then colour 1 (foraxes.jl:73-77) */
static void sweep_box(grid_t g, const int lo[3], const int hi[3], int coloring, real *U, real *V,
                      real *W, real *P, real *Div2, hsmac_coef c)
{
    colour_pass(g, lo, hi, 0, coloring, U, V, W, P, Div2, c);
    colour_pass(g, lo, hi, 1, coloring, U, V, W, P, Div2, c);
}

/* Topologies.faces(domain, core): lib/topologies/src/halo.jl:349-366.
 * Order z-, z+, y-, y+, x-, x+. */
static void sweep_core_then_faces(grid_t g, const int dlo[3], const int dhi[3], int coloring,
                                  real *U, real *V, real *W, real *P, real *Div2, hsmac_coef c)
{
    const int clo[3] = {2, 2, 2};
    const int chi[3] = {g.nx, g.ny, g.nz};
    sweep_box(g, clo, chi, coloring, U, V, W, P, Div2, c);
    int flo[6][3] = {
        {dlo[0], dlo[1], dlo[2]},      {dlo[0], dlo[1], chi[2] + 1}, {dlo[0], dlo[1], clo[2]},
        {dlo[0], chi[1] + 1, clo[2]}, {dlo[0], clo[1], clo[2]},      {chi[0] + 1, clo[1], clo[2]},
    };
    int fhi[6][3] = {
        {dhi[0], dhi[1], clo[2] - 1}, {dhi[0], dhi[1], dhi[2]}, {dhi[0], clo[1] - 1, chi[2]},
        {dhi[0], dhi[1], chi[2]},      {clo[0] - 1, chi[1], chi[2]}, {dhi[0], chi[1], chi[2]},
    };
    for (int f = 0; f < 6; f++) sweep_box(g, flo[f], fhi[f], coloring, U, V, W, P, Div2, c);
}

/* sqrt(allmean(Div2[(1:n+1)^3])): hsmac.jl:120,177,191; operators.jl:44-46 */
static double rms_div(grid_t g, const real *Div2)
{
    int n3 = g.nz + 1;
    double *part = (double *)calloc((size_t)n3, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= g.nz + 1; k++) {
        double s = 0.0;
        for (int j = 1; j <= g.ny + 1; j++)
            for (int i = 1; i <= g.nx + 1; i++) s += (double)Div2[AT(g, i, j, k)];
        part[k - 1] = s;
    }
    double s = 0.0;
    for (int k = 0; k < n3; k++) s += part[k];
    free(part);
    double len = (double)(g.nx + 1) * (double)(g.ny + 1) * (double)(g.nz + 1);
    return sqrt(s / len);
}

/* couple!(::HSMAC): hsmac.jl:52-75; _couple_naive! :77-129;
 * _couple_time_blocking! :131-201.  Returns (itr, div) like the reference. */
void o_hsmac(int nx, int ny, int nz, int h, double dt, int time_blocking, int overlap, int coloring,
             double omega, double epsp, int itrmin, int itrmax, real *U, real *V, real *W, real *P,
             int *itr_out, double *div_out)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    size_t len = (size_t)g.sx * g.sy * (size_t)(nz + 2 * h);
    real *Div2 = (real *)calloc(len, sizeof(real));
    double dx = spacing(nx), dy = spacing(ny), dz = spacing(nz);
    double dxmin = fmin(fmin(dx, dy), dz);
    real dxi = R(1.0 / dx), dyi = R(1.0 / dy), dzi = R(1.0 / dz);
    hsmac_coef c;
    c.dti = R(1.0 / dt);
    c.dxi = dxi;
    c.dyi = dyi;
    c.dzi = dzi;
    c.dl = R(1) / (R(2) * ((dxi * dxi + dyi * dyi) + dzi * dzi));
    c.omega = R(omega);

    int itr = 0;
    double div = 0.0;
    const int d3lo[3] = {1, 1, 1}, d3hi[3] = {nx + 1, ny + 1, nz + 1};
    if (!time_blocking) {
        while (itr < itrmax) {
            if (overlap) {
                o_ghost_fill(U, nx, ny, nz, h, 1, 1);
                o_ghost_fill(V, nx, ny, nz, h, 1, 1);
                o_ghost_fill(W, nx, ny, nz, h, 1, 1);
                sweep_core_then_faces(g, d3lo, d3hi, coloring, U, V, W, P, Div2, c);
            } else {
                o_ghost_fill(U, nx, ny, nz, h, 1, 1);
                o_ghost_fill(V, nx, ny, nz, h, 1, 1);
                o_ghost_fill(W, nx, ny, nz, h, 1, 1);
                o_ghost_fill(P, nx, ny, nz, h, 1, 1);
                sweep_box(g, d3lo, d3hi, coloring, U, V, W, P, Div2, c);
            }
            div = rms_div(g, Div2);
            if (div * dxmin < epsp && itrmin <= itr) break;
            itr += 1;
        }
    } else {
        const int d1lo[3] = {-1, -1, -1}, d1hi[3] = {nx + 3, ny + 3, nz + 3};
        const int d2lo[3] = {0, 0, 0}, d2hi[3] = {nx + 2, ny + 2, nz + 2};
        while (itr < itrmax) {
            int done = 0;
            for (int s = 0; s < 3 && !done; s++) {
                if (s == 0) {
                    o_ghost_fill(U, nx, ny, nz, h, h, 1);
                    o_ghost_fill(V, nx, ny, nz, h, h, 1);
                    o_ghost_fill(W, nx, ny, nz, h, h, 1);
                    if (overlap)
                        sweep_core_then_faces(g, d1lo, d1hi, coloring, U, V, W, P, Div2, c);
                    else
                        sweep_box(g, d1lo, d1hi, coloring, U, V, W, P, Div2, c);
                } else if (s == 1) {
                    sweep_box(g, d2lo, d2hi, coloring, U, V, W, P, Div2, c);
                } else {
                    sweep_box(g, d3lo, d3hi, coloring, U, V, W, P, Div2, c);
                }
                div = rms_div(g, Div2);
                if (div * dxmin < epsp && itrmin <= itr) {
                    done = 1;
                    break;
                }
                itr += 1;
            }
            if (done) break;
        }
    }
    free(Div2);
    *itr_out = itr;
    *div_out = div;
}

/* ------------------------------------------------------------------------- */
/* Flow time step: src/flows/timestep.jl:18-65.  Fills the face ghosts of     */
/* U, V, W (FaceHalo(h)) in place, then returns dt and the three maxima.      */
/* ------------------------------------------------------------------------- */
double o_flow_timestep(int nx, int ny, int nz, int h, double cfl, double dtmax, double Re, real *U,
                       real *V, real *W, double *umax3)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    o_ghost_fill(U, nx, ny, nz, h, h, 0);
    o_ghost_fill(V, nx, ny, nz, h, h, 0);
    o_ghost_fill(W, nx, ny, nz, h, h, 0);
    const long sy = g.sx, sz = g.sx * g.sy;
    double *part = (double *)calloc((size_t)nz * 3, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++) {
        double mu = 0, mv = 0, mw = 0;
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t p = AT(g, i, j, k);
                real uc = interp_o4(U[p - 2], U[p - 1], U[p], U[p + 1]);
                real vc = interp_o4(V[p - 2 * sy], V[p - sy], V[p], V[p + sy]);
                real wc = interp_o4(W[p - 2 * sz], W[p - sz], W[p], W[p + sz]);
                mu = fmax(mu, fabs((double)uc));
                mv = fmax(mv, fabs((double)vc));
                mw = fmax(mw, fabs((double)wc));
            }
        part[3 * (k - 1) + 0] = mu;
        part[3 * (k - 1) + 1] = mv;
        part[3 * (k - 1) + 2] = mw;
    }
    double m[3] = {0, 0, 0};
    for (int k = 0; k < nz; k++)
        for (int a = 0; a < 3; a++) m[a] = fmax(m[a], part[3 * k + a]);
    free(part);
    double umax = m[0] + 1e-10, vmax = m[1] + 1e-10, wmax = m[2] + 1e-10;
    if (umax3) {
        umax3[0] = umax;
        umax3[1] = vmax;
        umax3[2] = wmax;
    }
    double dx = spacing(nx), dy = spacing(ny), dz = spacing(nz);
    double cands[5] = {dtmax, cfl * dx / umax, cfl * dy / vmax, cfl * dz / wmax,
                       0.5 * ((dx / 2) * (dx / 2)) / (1 / Re)};
    double dt = cands[0];
    for (int a = 1; a < 5; a++) dt = fmin(dt, cands[a]);
    return dt;
}

/* ------------------------------------------------------------------------- */
/* Constant-power linear forcing: src/flows/forcings/lf.jl:27-106.           */
/* FullHalo(h) on U,V,W in place; returns mean power p.                       */
/* ------------------------------------------------------------------------- */
double o_force_linear(int nx, int ny, int nz, int h, double power, double dt, real *U, real *V,
                      real *W, double *means5)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    o_ghost_fill(U, nx, ny, nz, h, h, 1);
    o_ghost_fill(V, nx, ny, nz, h, h, 1);
    o_ghost_fill(W, nx, ny, nz, h, h, 1);
    const long sy = g.sx, sz = g.sx * g.sy;
    double *part = (double *)calloc((size_t)nz * 4, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++) {
        double s2 = 0, su = 0, sv = 0, sw = 0;
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t p = AT(g, i, j, k);
                real u = interp_o2(U[p - 1], U[p]);
                real v = interp_o2(V[p - sy], V[p]);
                real w = interp_o2(W[p - sz], W[p]);
                real u2 = (u * u + v * v) + w * w;
                s2 += (double)u2;
                su += (double)U[p];
                sv += (double)V[p];
                sw += (double)W[p];
            }
        part[4 * (k - 1) + 0] = s2;
        part[4 * (k - 1) + 1] = su;
        part[4 * (k - 1) + 2] = sv;
        part[4 * (k - 1) + 3] = sw;
    }
    double s[4] = {0, 0, 0, 0};
    for (int k = 0; k < nz; k++)
        for (int a = 0; a < 4; a++) s[a] += part[4 * k + a];
    double len = (double)nx * (double)ny * (double)nz;
    double u2m = s[0] / len, um = s[1] / len, vm = s[2] / len, wm = s[3] / len;
    double e = 0.5 * (((u2m - um * um) - vm * vm) - wm * wm);
    double A = power / (2 * e);
    if (means5) {
        means5[0] = u2m;
        means5[1] = um;
        means5[2] = vm;
        means5[3] = wm;
        means5[4] = A;
    }
    /* _linear_forcing!: lf.jl:72-106 */
    real rA = R(A), rum = R(um), rvm = R(vm), rwm = R(wm), rdt = R(dt);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++) {
        double sp = 0;
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t p = AT(g, i, j, k);
                real usrc = rA * (U[p] - rum);
                real vsrc = rA * (V[p] - rvm);
                real wsrc = rA * (W[p] - rwm);
                real u = U[p], v = V[p], w = W[p];
                real pw = (u * usrc + v * vsrc) + w * wsrc;
                U[p] = u + usrc * rdt;
                V[p] = v + vsrc * rdt;
                W[p] = w + wsrc * rdt;
                sp += (double)pw;
            }
        part[k - 1] = sp;
    }
    double sp = 0;
    for (int k = 0; k < nz; k++) sp += part[k];
    free(part);
    return sp / len;
}

/* ------------------------------------------------------------------------- */
/* Flow statistics: Flows.stat, src/flows/stat.jl:67-155.                     */
/* Reads U,V,W with their ghosts as they are (the reference fills FullHalo    */
/* first, stat.jl:79; the caller does that here).  out[15]:                   */
/*   0-8   mean, max, min of U, V, W over the interior (LCS.Summary,          */
/*         src/stat.jl:20-22);                                                */
/*   9-13  interior means of u2 = u^2+v^2+w^2 (raw face values), s2 =          */
/*         2 S_ij S_ij, and dudx^2, dudx^3, dudx^4 (stat.jl:84-125);           */
/*   14    the integral length (stat.jl:157-182) if integral_length, else 0.   */
/* ------------------------------------------------------------------------- */
static inline real diff1_o4(real a, real b, real c, real d) /* src/diff.jl:13-15 */
{
    return R(1.0 / 24.0) * (((a - R(27) * b) + R(27) * c) - d);
}

void o_flow_stat(int nx, int ny, int nz, int h, const real *U, const real *V, const real *W,
                 int integral_length, double *out)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    const long sy = g.sx, sz = g.sx * g.sy;
    const real dxi = R(1.0 / spacing(nx)), dyi = R(1.0 / spacing(ny)), dzi = R(1.0 / spacing(nz));
    double *part = (double *)calloc((size_t)nz * 14, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++) {
        double s[14];
        for (int a = 0; a < 3; a++) {
            s[3 * a] = 0.0;
            s[3 * a + 1] = -INFINITY;
            s[3 * a + 2] = INFINITY;
        }
        for (int a = 9; a < 14; a++) s[a] = 0.0;
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t p = AT(g, i, j, k);
                real u = U[p], v = V[p], w = W[p];
                real u2 = (u * u + v * v) + w * w;
                real dudx = dxi * diff1_o4(U[p - 2], U[p - 1], U[p], U[p + 1]);
                real dvdy = dyi * diff1_o4(V[p - 2 * sy], V[p - sy], V[p], V[p + sy]);
                real dwdz = dzi * diff1_o4(W[p - 2 * sz], W[p - sz], W[p], W[p + sz]);
                real dudy = dyi * diff1_o4(U[p - 2 * sy], U[p - sy], U[p], U[p + sy]);
                /* V[(i-1):(i+2), j-1, k] */
                real dvdx = dxi * diff1_o4(V[p - sy - 1], V[p - sy], V[p - sy + 1], V[p - sy + 2]);
                /* V[i, j-1, (k-1):(k+2)] */
                real dvdz = dzi * diff1_o4(V[p - sy - sz], V[p - sy], V[p - sy + sz], V[p - sy + 2 * sz]);
                real dwdy = dyi * diff1_o4(W[p - 2 * sy], W[p - sy], W[p], W[p + sy]);
                /* W[(i-1):(i+2), j, k-1] */
                real dwdx = dxi * diff1_o4(W[p - sz - 1], W[p - sz], W[p - sz + 1], W[p - sz + 2]);
                real dudz = dzi * diff1_o4(U[p - 2 * sz], U[p - sz], U[p], U[p + sz]);
                real s11 = R(0.5) * (dudx + dudx), s22 = R(0.5) * (dvdy + dvdy), s33 = R(0.5) * (dwdz + dwdz);
                real s12 = R(0.5) * (dudy + dvdx), s13 = R(0.5) * (dudz + dwdx), s23 = R(0.5) * (dvdz + dwdy);
                real s2 = R(2) * (((s11 * s11 + s22 * s22) + s33 * s33) +
                                  R(2) * ((s12 * s12 + s13 * s13) + s23 * s23));
                real d2 = dudx * dudx;
                real vals[3] = {u, v, w};
                for (int a = 0; a < 3; a++) {
                    s[3 * a] += (double)vals[a];
                    s[3 * a + 1] = fmax(s[3 * a + 1], (double)vals[a]);
                    s[3 * a + 2] = fmin(s[3 * a + 2], (double)vals[a]);
                }
                s[9] += (double)u2;
                s[10] += (double)s2;
                s[11] += (double)d2;
                s[12] += (double)(d2 * dudx);
                s[13] += (double)(d2 * d2);
            }
        for (int a = 0; a < 14; a++) part[14 * (k - 1) + a] = s[a];
    }
    double len = (double)nx * (double)ny * (double)nz;
    for (int a = 0; a < 14; a++) {
        double r = part[a];
        for (int k = 1; k < nz; k++) {
            double x = part[14 * k + a];
            if (a < 9 && a % 3 == 1)
                r = fmax(r, x);
            else if (a < 9 && a % 3 == 2)
                r = fmin(r, x);
            else
                r += x;
        }
        out[a] = (a < 9 && a % 3 != 0) ? r : r / len;
    }
    free(part);
    out[14] = 0.0;
    if (!integral_length) return;
    /* integral_length: stat.jl:157-182.  R = irfft(|rfft(U)|^2) along x is the
     * circular autocorrelation R[r] = sum_i U[i] U[i+r]; lags 0..n/2 are
     * summed up to the first negative one, Ls = acc dx / R[0], L = mean(Ls). */
    const int mid = nx / 2 + 1;
    const double dx = spacing(nx);
    double *ls = (double *)calloc((size_t)nz, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++) {
        double acc_k = 0.0;
        for (int j = 1; j <= ny; j++) {
            const real *row = U + AT(g, 1, j, k);
            double acc = 0.0, r0 = 0.0;
            for (int r = 0; r < mid; r++) {
                double c = 0.0;
                for (int i = 0; i < nx; i++) c += (double)row[i] * (double)row[(i + r) % nx];
                if (r == 0) r0 = c;
                if (c < 0) break;
                acc += c;
            }
            acc_k += acc * dx / r0;
        }
        ls[k - 1] = acc_k;
    }
    double L = 0.0;
    for (int k = 0; k < nz; k++) L += ls[k];
    free(ls);
    out[14] = L / ((double)ny * (double)nz);
}

/* ------------------------------------------------------------------------- */
/* Particle statistics on the cell index: src/particles/rdf.jl:1-103,          */
/* src/particles/stat.jl:171-231 (SPH density), src/particles/spectrum.jl.     */
/* Positions, ids, hashes and 1-based starts/stops are those makeindex!        */
/* leaves (hash = 1 + ci1 + (nc1+2) ci2 + (nc1+2)(nc2+2) ci3, axes 0:nc+1).     */
/* ------------------------------------------------------------------------- */
static void decode_hash(int64_t hsh, const int nc[3], int ci[3])
{
    int64_t r = hsh - 1;
    ci[0] = (int)(r % (nc[0] + 2));
    r /= nc[0] + 2;
    ci[1] = (int)(r % (nc[1] + 2));
    ci[2] = (int)(r / (nc[1] + 2));
}

/* Utils.norm (src/utils/fast.jl:42-46) of a difference, in the particle type */
static inline real pnorm(real dx, real dy, real dz) { return (real)sqrt((double)((dx * dx + dy * dy) + dz * dz)); }

/* rdf_assume_ci: pair counts per log-spaced bin (edges[k] < d <= edges[k+1])
 * and contact pairs |d/diam - 1| < tol, over pairs id1 < id2 found within
 * +-ncell coarse cells of a particle whose cell is inside the domain. */
void o_rdf(int nx, int ny, int nz, int grid_ratio, long n, const real *xs, const real *ys, const real *zs,
           const int64_t *ids, const int64_t *hashes, const int64_t *starts, const int64_t *stops, double diam,
           double tol, int nbins, const double *edges, int ncell, int64_t *npairs, int64_t *ncontact)
{
    const int nc[3] = {nx / grid_ratio, ny / grid_ratio, nz / grid_ratio};
    for (int k = 0; k < nbins; k++) npairs[k] = 0;
    *ncontact = 0;
    for (long i1 = 0; i1 < n; i1++) {
        int ci[3];
        decode_hash(hashes[i1], nc, ci);
        if (ci[0] < 1 || ci[0] > nc[0] || ci[1] < 1 || ci[1] > nc[1] || ci[2] < 1 || ci[2] > nc[2]) continue;
        for (int di = -ncell; di <= ncell; di++)
            for (int dj = -ncell; dj <= ncell; dj++)
                for (int dk = -ncell; dk <= ncell; dk++) {
                    const int c2[3] = {ci[0] + di, ci[1] + dj, ci[2] + dk};
                    if (c2[0] < 0 || c2[0] > nc[0] + 1 || c2[1] < 0 || c2[1] > nc[1] + 1 || c2[2] < 0 ||
                        c2[2] > nc[2] + 1)
                        continue;
                    const int64_t h2 = 1 + c2[0] + (int64_t)(nc[0] + 2) * (c2[1] + (int64_t)(nc[1] + 2) * c2[2]);
                    for (int64_t i2 = starts[h2 - 1]; i2 <= stops[h2 - 1]; i2++) {
                        const long j = (long)i2 - 1;
                        if (!(ids[i1] < ids[j])) continue;
                        const double d = (double)pnorm(xs[i1] - xs[j], ys[i1] - ys[j], zs[i1] - zs[j]);
                        /* searchsortedfirst(edges, d) - 1: first edge >= d */
                        int lo = 0, hi = nbins + 1;
                        while (lo < hi) {
                            const int m = (lo + hi) / 2;
                            if (edges[m] < d)
                                lo = m + 1;
                            else
                                hi = m;
                        }
                        const int k = lo; /* 1-based bin = lo (0-based index of the first edge >= d) */
                        if (k >= 1 && k <= nbins) npairs[k - 1]++;
                        if (fabs(d / diam - 1) < tol) (*ncontact)++;
                    }
                }
    }
}

/* cubic_spline_kernel: src/particles/stat.jl:220-231 */
static double cubic_spline(double r, double h)
{
    const double q = r / h, factor = 1 / (M_PI * (h * h * h));
    if (q < 1) return factor * ((1 - 1.5 * (q * q)) + 0.75 * ((q * q) * q));
    if (q < 2) return (factor * 0.25) * (((2 - q) * (2 - q)) * (2 - q));
    return 0.0;
}

/* density: src/particles/stat.jl:171-217 (h = smallest fine spacing) */
void o_sph_density(int nx, int ny, int nz, int grid_ratio, long n, const real *xs, const real *ys, const real *zs,
                   const int64_t *hashes, const int64_t *starts, const int64_t *stops, double *dens)
{
    const int nc[3] = {nx / grid_ratio, ny / grid_ratio, nz / grid_ratio};
    const double h = fmin(fmin(spacing(nx), spacing(ny)), spacing(nz));
    for (long i1 = 0; i1 < n; i1++) {
        int ci[3];
        decode_hash(hashes[i1], nc, ci);
        double rho = 0.0;
        if (ci[0] >= 1 && ci[0] <= nc[0] && ci[1] >= 1 && ci[1] <= nc[1] && ci[2] >= 1 && ci[2] <= nc[2]) {
            for (int di = -1; di <= 1; di++)
                for (int dj = -1; dj <= 1; dj++)
                    for (int dk = -1; dk <= 1; dk++) {
                        const int c2[3] = {ci[0] + di, ci[1] + dj, ci[2] + dk};
                        const int64_t h2 =
                            1 + c2[0] + (int64_t)(nc[0] + 2) * (c2[1] + (int64_t)(nc[1] + 2) * c2[2]);
                        for (int64_t i2 = starts[h2 - 1]; i2 <= stops[h2 - 1]; i2++) {
                            const long j = (long)i2 - 1;
                            const double d = (double)pnorm(xs[i1] - xs[j], ys[i1] - ys[j], zs[i1] - zs[j]);
                            if (d < 2 * h) rho += cubic_spline(d, h);
                        }
                    }
        }
        dens[i1] = rho;
    }
}

/* _fibonacci_direction: src/particles/spectrum.jl:3-9 (i 0-based) */
void o_fib_dir(int i, int nsamples, double out[3])
{
    const double phi = M_PI * (3 - sqrt(5.0));
    const double y = 1 - ((double)i / (double)(nsamples - 1)) * 2;
    const double radius = sqrt(1 - y * y);
    const double theta = phi * i;
    out[0] = cos(theta) * radius;
    out[1] = y;
    out[2] = sin(theta) * radius;
}

/* spectrum: src/particles/spectrum.jl:11-78.  wavenumbers[nshells]; cos/sin
 * sums per (shell, direction) accumulated per particle in double; Enp out. */
void o_pspectrum(long n, const real *xs, const real *ys, const real *zs, int nshells, const int *wavenumbers,
                 int nsamples, double nvalid_g, double *Enp)
{
    double *cs = (double *)calloc((size_t)nshells * nsamples * 2, sizeof(double));
    for (int b = 0; b < nshells; b++)
        for (int i = 0; i < nsamples; i++) {
            double dir[3];
            o_fib_dir(i, nsamples, dir);
            const double kk = wavenumbers[b], k[3] = {kk * dir[0], kk * dir[1], kk * dir[2]};
            double c = 0, sn = 0;
            for (long p = 0; p < n; p++) {
                const double kx = (k[0] * (double)xs[p] + k[1] * (double)ys[p]) + k[2] * (double)zs[p];
                c += cos(kx);
                sn += sin(kx);
            }
            cs[2 * (b * nsamples + i)] = c;
            cs[2 * (b * nsamples + i) + 1] = sn;
        }
    for (int b = 0; b < nshells; b++) {
        double phi = 0.0;
        for (int i = 0; i < nsamples; i++) {
            const double c = cs[2 * (b * nsamples + i)], sn = cs[2 * (b * nsamples + i) + 1];
            phi += (c * c + sn * sn) / (nvalid_g * nvalid_g) - 1 / nvalid_g;
        }
        const double kk = wavenumbers[b];
        Enp[b] = phi / nsamples * 4 * M_PI * (kk * kk);
    }
    free(cs);
}

/* ------------------------------------------------------------------------- */
/* rc-energy forcing: src/flows/forcings/rcf.jl:21-191,                      */
/* lib/fouriertools/src/FourierTools.jl:9-208.                                */
/* ------------------------------------------------------------------------- */

/* boxmean_l! (AMD loop order and reciprocal weight): ext/AMDGPUExt.jl:18-41.
 * coarse is column-major cx*cy*cz. */
void o_boxmean(int nx, int ny, int nz, int h, int cx, int cy, int cz, const real *U, real *coarse)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    int kx = nx / cx, ky = ny / cy, kz = nz / cz;
    real w = R(1) / R((double)kx * ky * kz);
#pragma omp parallel for schedule(static)
    for (int kc = 1; kc <= cz; kc++)
        for (int jc = 1; jc <= cy; jc++)
            for (int ic = 1; ic <= cx; ic++) {
                real s = R(0);
                for (int dk = 1; dk <= kz; dk++)
                    for (int dj = 1; dj <= ky; dj++)
                        for (int di = 1; di <= kx; di++) {
                            int i = (ic - 1) * kx + di, j = (jc - 1) * ky + dj, k = (kc - 1) * kz + dk;
                            s += U[AT(g, i, j, k)];
                        }
                coarse[(ic - 1) + (size_t)cx * ((jc - 1) + (size_t)cy * (kc - 1))] = s * w;
            }
}

/* kcount(k0): FourierTools.jl:87-96 */
int o_kcount(int k0)
{
    int c = 0;
    double lo = (k0 - 0.5) * (k0 - 0.5), hi = (k0 + 0.5) * (k0 + 0.5);
    for (int kz = -k0; kz <= k0; kz++)
        for (int ky = -k0; ky <= k0; ky++)
            for (int kx = -k0; kx <= k0; kx++) {
                int r2 = kx * kx + ky * ky + kz * kz;
                if (lo <= r2 && r2 < hi) c++;
            }
    return c;
}

/* canon + kset(k0): FourierTools.jl:59-80.  Writes up to cap triples into
 * out (kx,ky,kz interleaved) in the reference's insertion order; returns count. */
int o_kset(int k0, int *out, int cap)
{
    int n = 0;
    double lo = (k0 - 0.5) * (k0 - 0.5), hi = (k0 + 0.5) * (k0 + 0.5);
    for (int kz = -k0; kz <= k0; kz++)
        for (int ky = -k0; ky <= k0; ky++)
            for (int kx = 0; kx <= k0; kx++) {
                int r2 = kx * kx + ky * ky + kz * kz;
                if (!(lo <= r2 && r2 < hi)) continue;
                int c[3] = {kx, ky, kz};
                int sgn = 0;
                for (int a = 0; a < 3 && sgn == 0; a++) sgn = (c[a] > 0) ? 1 : (c[a] < 0 ? -1 : 0);
                if (sgn < 0)
                    for (int a = 0; a < 3; a++) c[a] = -c[a];
                int dup = 0;
                for (int m = 0; m < n && !dup; m++)
                    dup = (out[3 * m] == c[0] && out[3 * m + 1] == c[1] && out[3 * m + 2] == c[2]);
                if (dup) continue;
                if (n < cap) {
                    out[3 * n] = c[0];
                    out[3 * n + 1] = c[1];
                    out[3 * n + 2] = c[2];
                }
                n++;
            }
    return n;
}

static inline int fftindex(int k, int n) { return k >= 0 ? k : n + k; } /* 0-based */

/* Forward real DFT coefficient (FFTW sign convention, unnormalised) of a
 * column-major c^3 grid at wavevector (kx,ky,kz).  Direct sum over all points;
 * stands in for FFTW's rfft (src/utils/fftw.jl:4, rcf.jl:45-47). */
static void dft_coeff(int cx, int cy, int cz, const real *A, int kx, int ky, int kz, double *re,
                      double *im)
{
    double sr = 0, si = 0;
    for (int z = 0; z < cz; z++)
        for (int y = 0; y < cy; y++)
            for (int x = 0; x < cx; x++) {
                /* phase = -2π (kx x/cx + ky y/cy + kz z/cz), reduced exactly */
                long num = ((long)kx * x * cy * cz + (long)ky * y * cx * cz + (long)kz * z * cx * cy);
                long den = (long)cx * cy * cz;
                long m = num % den;
                if (m < 0) m += den;
                double ang = -2.0 * M_PI * (double)m / (double)den;
                double a = (double)A[x + (size_t)cx * (y + (size_t)cy * z)];
                sr += a * cos(ang);
                si += a * sin(ang);
            }
    *re = sr;
    *im = si;
}

/* enespe0: rcf.jl:187-191 */
double o_enespe0(int k0, double e10)
{
    double c1 = (double)o_kcount(1), ck = (double)o_kcount(k0);
    double a = k0 + 0.5, b = k0 - 0.5;
    return e10 * pow((double)k0, -5.0 / 3.0) * (1.5 * 1.5 * 1.5 - 0.5 * 0.5 * 0.5) /
           ((a * a * a) - (b * b * b)) * ck / c1;
}

/* Spectral part of forcing!(::EnergyPreserveRCF): rcf.jl:43-67.
 * Input: the three coarse grids (c^3 each, column-major).  The real DFT of
 * each grid is held in the reference's rfft layout (c/2+1) x c x c (FFTW sign
 * convention, unnormalised), so wavevectors that alias onto one entry when
 * c < 2 kmax + 1 are scaled and read exactly as the reference does.
 * Output: the three cropped, normalised spectra in (kmax+1) x (2kmax+1) x
 * (2kmax+1) complex layout (re,im interleaved, component-major), i.e. the array
 * sirdft reads.  en_out (if not NULL) receives en per shell k0 = kmin..kmax
 * before scaling.  Returns 0, or -1 if kmax > c/2 (the reference would index
 * out of bounds). */
int o_rcf_spectrum(int c, int kmin, int kmax, double e10, const real *Uc, const real *Vc,
                   const real *Wc, double *crop3, double *en_out)
{
    if (kmax > c / 2 || kmin < 1 || kmax < kmin) return -1;
    int hx = c / 2 + 1, nk = 2 * kmax + 1, ncrop = (kmax + 1) * nk * nk;
    size_t nh = (size_t)hx * c * c;
    double *hat = (double *)calloc(nh * 6, sizeof(double));
    const real *A[3] = {Uc, Vc, Wc};
    for (int z = 0; z < c; z++)
        for (int y = 0; y < c; y++)
            for (int x = 0; x < hx; x++) {
                size_t q = (size_t)x + (size_t)hx * (y + (size_t)c * z);
                for (int a = 0; a < 3; a++)
                    dft_coeff(c, c, c, A[a], x, y, z, &hat[6 * q + 2 * a], &hat[6 * q + 2 * a + 1]);
            }
#define HQ(kx, ky, kz) ((size_t)(kx) + (size_t)hx * (fftindex(ky, c) + (size_t)c * fftindex(kz, c)))
    double norm = (double)(hx * 2 - 2) * c * c; /* FourierTools.jl:36 */
    for (int k0 = kmin; k0 <= kmax; k0++) {
        /* enespe: FourierTools.jl:103-120 (hasnyquist=false) */
        double e = 0;
        double lo = (k0 - 0.5) * (k0 - 0.5), hi = (k0 + 0.5) * (k0 + 0.5);
        for (int kz = -k0; kz <= k0; kz++)
            for (int ky = -k0; ky <= k0; ky++)
                for (int kx = 0; kx <= k0; kx++) {
                    int r2 = kx * kx + ky * ky + kz * kz;
                    if (!(lo <= r2 && r2 < hi)) continue;
                    size_t q = HQ(kx, ky, kz);
                    double w = (kx == 0) ? 1 : 2;
                    double s = 0;
                    for (int a = 0; a < 3; a++) {
                        double re = hat[6 * q + 2 * a] / norm, im = hat[6 * q + 2 * a + 1] / norm;
                        double ab = re * re + im * im;
                        s = (a == 0) ? ab : s + ab;
                    }
                    e += w * s;
                }
        double en = e / 2;
        if (en_out) en_out[k0 - kmin] = en;
        double en0 = o_enespe0(k0, e10);
        double r = sqrt(en0 / fmax(en, 1e-8)) - 1.0;
        /* shell scaling: rcf.jl:53-62 */
        for (int kz = -kmax; kz <= kmax; kz++)
            for (int ky = -kmax; ky <= kmax; ky++)
                for (int kx = 0; kx <= kmax; kx++) {
                    int r2 = kx * kx + ky * ky + kz * kz;
                    if (!(lo <= r2 && r2 < hi)) continue;
                    size_t q = HQ(kx, ky, kz);
                    for (int a = 0; a < 6; a++) hat[6 * q + a] *= r;
                }
    }
    /* crop with normalisation: FourierTools.jl:142-157 */
    for (int kz = -kmax; kz <= kmax; kz++)
        for (int ky = -kmax; ky <= kmax; ky++)
            for (int kx = 0; kx <= kmax; kx++) {
                size_t q = HQ(kx, ky, kz);
                int o = kx + (kmax + 1) * (fftindex(ky, nk) + nk * fftindex(kz, nk));
                for (int a = 0; a < 3; a++) {
                    crop3[2 * ((size_t)a * ncrop + o)] = hat[6 * q + 2 * a] / norm;
                    crop3[2 * ((size_t)a * ncrop + o) + 1] = hat[6 * q + 2 * a + 1] / norm;
                }
            }
#undef HQ
    free(hat);
    return 0;
}

typedef struct {
    double re, im;
} cplx;
static inline cplx cmul(cplx a, cplx b)
{
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cplx cconj(cplx a)
{
    cplx r = {a.re, -a.im};
    return r;
}
/* powi: FourierTools.jl:159-163 (x^n as repeated products, conj for n<0) */
static inline cplx cpowi(cplx x, int n)
{
    cplx one = {1.0, 0.0};
    if (n == 0) return one;
    if (n < 0) return cconj(cpowi(x, -n));
    cplx r = x;
    for (int m = 1; m < n; m++) r = cmul(x, r);
    return r;
}
/* cispi(x) = exp(iπx) with exact reduction of x modulo 2 */
static inline cplx ccispi(double x)
{
    double r = fmod(x, 2.0);
    cplx c = {cos(M_PI * r), sin(M_PI * r)};
    return c;
}

/* Apply part of forcing!(::EnergyPreserveRCF): rcf.jl:69-99.  crop3 as
 * produced by o_rcf_spectrum.  Returns mean power. */
double o_rcf_apply(int nx, int ny, int nz, int h, int c, int kmin, int kmax, double dt,
                   const double *crop3, real *U, real *V, real *W)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    int nk = 2 * kmax + 1, ncrop = (kmax + 1) * nk * nk;
    double phi = (double)(nx / c) / 2.0; /* phase_shift: src/flows/config.jl:147-149 */
    double nxi = 1.0 / nx, nyi = 1.0 / ny, nzi = 1.0 / nz;
    double dti = 1.0 / dt;
    int ks[512];
    int kof[16], kn[16];
    int nm = 0;
    for (int K = kmin; K <= kmax; K++) {
        kof[K - kmin] = nm;
        kn[K - kmin] = o_kset(K, ks + 3 * nm, (512 / 3) - nm);
        nm += kn[K - kmin];
    }
    double *part = (double *)calloc((size_t)nz, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++) {
        double sp = 0;
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                cplx e1k = ccispi(2 * ((double)k - phi) * nzi);
                cplx e1j = ccispi(2 * ((double)j - phi) * nyi);
                cplx e1i = ccispi(2 * ((double)i - phi) * nxi);
                double src[3];
                for (int a = 0; a < 3; a++) {
                    double tot = 0;
                    for (int K = kmin; K <= kmax; K++) {
                        double acc = 0;
                        for (int m = kof[K - kmin]; m < kof[K - kmin] + kn[K - kmin]; m++) {
                            int kx = ks[3 * m], ky = ks[3 * m + 1], kz = ks[3 * m + 2];
                            cplx xyz = cmul(cmul(cpowi(e1i, kx), cpowi(e1j, ky)), cpowi(e1k, kz));
                            int q = kx + (kmax + 1) * (fftindex(ky, nk) + nk * fftindex(kz, nk));
                            cplx A = {crop3[2 * (a * ncrop + q)], crop3[2 * (a * ncrop + q) + 1]};
                            acc += 2 * (A.re * xyz.re - A.im * xyz.im);
                        }
                        tot += acc;
                    }
                    src[a] = tot;
                }
                size_t p = AT(g, i, j, k);
                real usrc = R(src[0]), vsrc = R(src[1]), wsrc = R(src[2]);
                real u = U[p], v = V[p], w = W[p];
                real pw = ((u * usrc + v * vsrc) + w * wsrc) * R(dti);
                U[p] += usrc;
                V[p] += vsrc;
                W[p] += wsrc;
                sp += (double)pw;
            }
        part[k - 1] = sp;
    }
    double sp = 0;
    for (int k = 0; k < nz; k++) sp += part[k];
    free(part);
    return sp / ((double)nx * ny * nz);
}

/* ------------------------------------------------------------------------- */
/* Particles                                                                  */
/* ------------------------------------------------------------------------- */

/* apply_periodic_boundary!: src/particles/boundary.jl:81-120 */
void o_p_boundary(long n, real *xs, real *ys, real *zs, real *xs2, real *ys2, real *zs2)
{
    const real L = R(DOMAIN_LENGTH), O = R(DOMAIN_ORIGIN), E = R(DOMAIN_ORIGIN + DOMAIN_LENGTH);
    real *X[3] = {xs, ys, zs}, *X2[3] = {xs2, ys2, zs2};
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++)
        for (int a = 0; a < 3; a++) {
            real x = X[a][p], x2 = X2[a][p];
            if (x < O) {
                X[a][p] = x + L;
                X2[a][p] = x2 + L;
            }
            if (x >= E) {
                X[a][p] = x - L;
                X2[a][p] = x2 - L;
            }
        }
}

/* interp + _interp_trilinear: src/particles/interp.jl:62-90 */
static inline real trilinear(grid_t g, const real *A, int i, int j, int k, real r1, real r2, real r3)
{
    size_t p = AT(g, i, j, k);
    const long sy = g.sx, sz = g.sx * g.sy;
    real c0 = A[p], c1 = A[p + 1], c2 = A[p + 1 + sy], c3 = A[p + sy];
    real c4 = A[p + sz], c5 = A[p + 1 + sz], c6 = A[p + 1 + sy + sz], c7 = A[p + sy + sz];
    real o = R(1);
    /* eight terms summed left to right, each product left to right */
    real s = (((o - r1) * (o - r2)) * (o - r3)) * c0;
    s = s + ((r1 * (o - r2)) * (o - r3)) * c1;
    s = s + ((r1 * r2) * (o - r3)) * c2;
    s = s + (((o - r1) * r2) * (o - r3)) * c3;
    s = s + (((o - r1) * (o - r2)) * r3) * c4;
    s = s + ((r1 * (o - r2)) * r3) * c5;
    s = s + ((r1 * r2) * r3) * c6;
    s = s + (((o - r1) * r2) * r3) * c7;
    return s;
}

static inline int ufloor(real x) { return (int)floor((double)x); } /* unsafe_floor: src/utils/fast.jl:49-52 */

/* x/y/zmapping + interp: interp.jl:92-126.  axis 0,1,2 selects U,V,W staggering. */
static inline real interp_at(grid_t g, const real *A, int axis, real x, real y, real z, real dx,
                             real dy, real dz)
{
    real h5 = R(0.5);
    real qx = x / dx, qy = y / dy, qz = z / dz;
    int i, j, k;
    real r1, r2, r3;
    if (axis == 0) {
        i = ufloor(qx);
        j = ufloor(qy + h5);
        k = ufloor(qz + h5);
        r1 = qx - (real)i;
        r2 = (qy - (real)j) + h5;
        r3 = (qz - (real)k) + h5;
    } else if (axis == 1) {
        j = ufloor(qy);
        k = ufloor(qz + h5);
        i = ufloor(qx + h5);
        r2 = qy - (real)j;
        r3 = (qz - (real)k) + h5;
        r1 = (qx - (real)i) + h5;
    } else {
        k = ufloor(qz);
        i = ufloor(qx + h5);
        j = ufloor(qy + h5);
        r3 = qz - (real)k;
        r2 = (qy - (real)j) + h5;
        r1 = (qx - (real)i) + h5;
    }
    return trilinear(g, A, i, j, k, r1, r2, r3);
}

/* Fluid velocity at particle positions (used by the zero-slip initialiser and tests). */
void o_p_fluid_velocity(int nx, int ny, int nz, int h, const real *U, const real *V, const real *W,
                        long n, const real *xs, const real *ys, const real *zs, real *uf, real *vf,
                        real *wf)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    real dx = R(spacing(nx)), dy = R(spacing(ny)), dz = R(spacing(nz));
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++) {
        real x = xs[p] - R(DOMAIN_ORIGIN), y = ys[p] - R(DOMAIN_ORIGIN), z = zs[p] - R(DOMAIN_ORIGIN);
        uf[p] = interp_at(g, U, 0, x, y, z, dx, dy, dz);
        vf[p] = interp_at(g, V, 1, x, y, z, dx, dy, dz);
        wf[p] = interp_at(g, W, 2, x, y, z, dx, dy, dz);
    }
}

/* motion!: src/particles/motion.jl:38-84; coeff: src/particles/config.jl:128,139;
 * taup: config.jl:41.  drag 0 = linear, 1 = nonlinear (Schiller-Naumann). */
void o_p_motion(int nx, int ny, int nz, int h, double Re, double gravity, double density_ratio,
                int drag, const real *U, const real *V, const real *W, long n, const real *xs,
                const real *ys, const real *zs, const real *us, const real *vs, const real *ws,
                const real *diams, real *dudts, real *dvdts, real *dwdts)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    real dx = R(spacing(nx)), dy = R(spacing(ny)), dz = R(spacing(nz));
    real rRe = R(Re), rg = R(gravity);
    real kap = R(((1.0 / 18.0) * density_ratio) * Re); /* taup = kap * d^2, config.jl:41 */
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++) {
        real x = xs[p] - R(DOMAIN_ORIGIN), y = ys[p] - R(DOMAIN_ORIGIN), z = zs[p] - R(DOMAIN_ORIGIN);
        real up = us[p], vp = vs[p], wp = ws[p];
        real diam = diams[p];
        real uf = interp_at(g, U, 0, x, y, z, dx, dy, dz);
        real vf = interp_at(g, V, 1, x, y, z, dx, dy, dz);
        real wf = interp_at(g, W, 2, x, y, z, dx, dy, dz);
        real a = uf - up, b = vf - vp, c = wf - wp;
        real vrel = (real)sqrt((double)((a * a + b * b) + c * c));
        real Rep = (diam * vrel) * rRe;
        real f = drag ? (R(1.0) + R(0.15) * (real)pow((double)Rep, 0.687)) : R(1.0);
        real taup = kap * (diam * diam);
        real taui = R(1) / taup;
        dudts[p] = ((f * (uf - up)) * taui) - rg;
        dvdts[p] = (f * (vf - vp)) * taui;
        dwdts[p] = (f * (wf - wp)) * taui;
    }
}

/* update!: src/particles/update.jl:32-65 */
void o_p_update(long n, double dt, real *xs, real *ys, real *zs, const real *xs0, const real *ys0,
                const real *zs0, real *us, real *vs, real *ws, const real *us0, const real *vs0,
                const real *ws0, const real *dudts, const real *dvdts, const real *dwdts)
{
    real rdt = R(dt);
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++) {
        real x0 = xs0[p], y0 = ys0[p], z0 = zs0[p];
        real u0 = us0[p], v0 = vs0[p], w0 = ws0[p];
        real a = dudts[p], b = dvdts[p], c = dwdts[p];
        xs[p] = x0 + rdt * u0;
        ys[p] = y0 + rdt * v0;
        zs[p] = z0 + rdt * w0;
        us[p] = u0 + rdt * a;
        vs[p] = v0 + rdt * b;
        ws[p] = w0 + rdt * c;
    }
}

/* Particles.timestep: src/particles/timestep.jl:15-25 (max over all ntotal slots) */
double o_p_timestep(long ntotal, const real *diams, double Re, double density_ratio, double dtmax)
{
    double m = -INFINITY;
    for (long p = 0; p < ntotal; p++) m = fmax(m, (double)diams[p]);
    double taup = (((1.0 / 18.0) * density_ratio) * Re) * (m * m);
    return fmin(0.5 * taup, dtmax);
}

/* makeindex!: src/particles/makeindex.jl:29-67; encode/sub2ind
 * src/utils/hasher.jl:15-38; searchsortedfirst/last src/utils/fast.jl:8-40.
 * props: array of 16 real property pointers (diams, xss, xss2, uss, uss2,
 * dudtss) permuted together with ids (int64) and hashes.  The sort is stable
 * (ties keep input order).  starts/stops are 1-based, length (nc+2)^3. */
static long ssfirst(const int64_t *v, long n, int64_t x)
{
    long lo = 1, len = n;
    while (len != 0) {
        long half = len >> 1;
        long m = lo + half;
        int isless = v[m - 1] < x;
        lo = isless ? (m + 1) : lo;
        len = isless ? (len - (half + 1)) : half;
    }
    return lo;
}
static long sslast(const int64_t *v, long n, int64_t x)
{
    long lo = 0, hi = n + 1;
    while (lo < hi - 1) {
        long m = lo + ((hi - lo) >> 1);
        int isless = x < v[m - 1];
        hi = isless ? m : hi;
        lo = isless ? lo : m;
    }
    return lo;
}

void o_hash(int nx, int ny, int nz, int grid_ratio, long n, const real *xs, const real *ys,
            const real *zs, int64_t *hashes)
{
    int nc[3] = {nx / grid_ratio, ny / grid_ratio, nz / grid_ratio};
    real dxc[3] = {R(spacing(nx) * grid_ratio), R(spacing(ny) * grid_ratio), R(spacing(nz) * grid_ratio)};
    const real *X[3] = {xs, ys, zs};
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++) {
        int64_t idx = 1, stride = 1;
        for (int a = 0; a < 3; a++) {
            int32_t c = (int32_t)floor((double)((X[a][p] - R(DOMAIN_ORIGIN)) / dxc[a]));
            int64_t ci = (int64_t)c + 1;
            idx += ci * stride;
            stride *= (nc[a] + 2);
        }
        hashes[p] = idx;
    }
}

void o_makeindex(int nx, int ny, int nz, int grid_ratio, long n, real *xs, real *ys, real *zs,
                 int64_t *ids, real **props, int nprops, int64_t *hashes, int64_t *perm,
                 int64_t *starts, int64_t *stops)
{
    o_hash(nx, ny, nz, grid_ratio, n, xs, ys, zs, hashes);
    int nc[3] = {nx / grid_ratio, ny / grid_ratio, nz / grid_ratio};
    int64_t ncells = (int64_t)(nc[0] + 2) * (nc[1] + 2) * (nc[2] + 2);
    /* stable counting sort by hash (ties in input order) */
    int64_t *count = (int64_t *)calloc((size_t)ncells + 2, sizeof(int64_t));
    for (long p = 0; p < n; p++) count[hashes[p]]++;
    int64_t acc = 0;
    for (int64_t c = 0; c <= ncells + 1; c++) {
        int64_t t = count[c];
        count[c] = acc;
        acc += t;
    }
    for (long p = 0; p < n; p++) perm[count[hashes[p]]++] = p + 1; /* 1-based like sortperm */
    free(count);
    /* sortbyperm!: src/utils/sort.jl:7-22 */
    void *tmp = malloc((size_t)n * 8);
    int64_t *ti = (int64_t *)tmp;
    for (long p = 0; p < n; p++) ti[p] = hashes[perm[p] - 1];
    memcpy(hashes, ti, (size_t)n * 8);
    for (long p = 0; p < n; p++) ti[p] = ids[perm[p] - 1];
    memcpy(ids, ti, (size_t)n * 8);
    real *tr = (real *)tmp;
    for (int q = 0; q < nprops; q++) {
        for (long p = 0; p < n; p++) tr[p] = props[q][perm[p] - 1];
        memcpy(props[q], tr, (size_t)n * sizeof(real));
    }
    free(tmp);
#pragma omp parallel for schedule(static)
    for (int64_t x = 1; x <= ncells; x++) {
        starts[x - 1] = ssfirst(hashes, n, x);
        stops[x - 1] = sslast(hashes, n, x);
    }
}

/* Exported so tests can check the binary searches against numpy/bisect. */
long o_searchsortedfirst(const int64_t *v, long n, int64_t x) { return ssfirst(v, n, x); }
long o_searchsortedlast(const int64_t *v, long n, int64_t x) { return sslast(v, n, x); }

/* ------------------------------------------------------------------------- */
/* Builder-defined initial conditions (not in the reference; the reference's */
/* random-flow/random-position use Julia's RNG streams, which cannot be      */
/* reproduced here).  Counter-based RNG shared with the HIP library.         */
/* ------------------------------------------------------------------------- */
static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* uniform double in [0,1) from (seed, stream, index) */
double o_rand_u01(uint64_t seed, uint64_t stream, uint64_t idx)
{
    uint64_t z = mix64(seed ^ mix64(stream ^ mix64(idx)));
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* Spectrum cutoff used by the initialiser */
int o_spectrum_kcut(int nmin, double k0)
{
    int K = (int)ceil(3.0 * k0);
    if (K > nmin / 2 - 1) K = nmin / 2 - 1;
    return K;
}

/* Mode table of the solenoidal spectrum field E(k) ∝ k^4 exp(-2 (k/k0)^2).
 * Canonical half-space modes with 0 < |k| <= K.  Writes (kx,ky,kz) into kout and
 * complex amplitudes (ux,uy,uz re/im interleaved, 6 doubles) into aout.
 * Returns the number of modes. */
int o_spectrum_modes(uint64_t seed, double k0, int K, int *kout, double *aout, int cap)
{
    int n = 0;
    for (int kz = -K; kz <= K; kz++)
        for (int ky = -K; ky <= K; ky++)
            for (int kx = 0; kx <= K; kx++) {
                int canonical = (kx > 0) || (kx == 0 && ky > 0) || (kx == 0 && ky == 0 && kz > 0);
                int r2 = kx * kx + ky * ky + kz * kz;
                if (!canonical || r2 == 0 || r2 > K * K) continue;
                uint64_t idx = (uint64_t)(kx + (K + 1) * ((ky + K) + (2 * K + 1) * (kz + K)));
                double a[6];
                for (int c = 0; c < 6; c++) a[c] = 2.0 * o_rand_u01(seed, (uint64_t)c, idx) - 1.0;
                double kk = sqrt((double)r2);
                double amp = kk * exp(-(kk / k0) * (kk / k0));
                double kv[3] = {(double)kx, (double)ky, (double)kz};
                for (int part = 0; part < 2; part++) {
                    double d = 0;
                    for (int c = 0; c < 3; c++) d += kv[c] * a[2 * c + part];
                    for (int c = 0; c < 3; c++) a[2 * c + part] = amp * (a[2 * c + part] - kv[c] * d / r2);
                }
                if (n < cap) {
                    kout[3 * n] = kx;
                    kout[3 * n + 1] = ky;
                    kout[3 * n + 2] = kz;
                    for (int c = 0; c < 6; c++) aout[6 * n + c] = a[c];
                }
                n++;
            }
    return n;
}

/* Evaluate the mode table on the staggered grid and scale to u' = uprime.
 * U at (i dx, (j-1/2) dy, (k-1/2) dz), V at ((i-1/2) dx, j dy, (k-1/2) dz),
 * W at ((i-1/2) dx, (j-1/2) dy, k dz).  P = 0 everywhere. */
void o_init_spectrum(int nx, int ny, int nz, int h, uint64_t seed, double k0, double uprime, real *U,
                     real *V, real *W, real *P)
{
    grid_t g = mkgrid(nx, ny, nz, h);
    int nmin = nx < ny ? (nx < nz ? nx : nz) : (ny < nz ? ny : nz);
    int K = o_spectrum_kcut(nmin, k0);
    int cap = (K + 1) * (2 * K + 1) * (2 * K + 1);
    int *kt = (int *)malloc(sizeof(int) * 3 * (size_t)cap);
    double *at = (double *)malloc(sizeof(double) * 6 * (size_t)cap);
    int nm = o_spectrum_modes(seed, k0, K, kt, at, cap);
    size_t len = (size_t)g.sx * g.sy * (size_t)(nz + 2 * h);
    memset(U, 0, len * sizeof(real));
    memset(V, 0, len * sizeof(real));
    memset(W, 0, len * sizeof(real));
    memset(P, 0, len * sizeof(real));
    double dx = spacing(nx), dy = spacing(ny), dz = spacing(nz);
    real *F[3] = {U, V, W};
    double *part = (double *)calloc((size_t)nz, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++) {
        double ss = 0;
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                for (int c = 0; c < 3; c++) {
                    double x = (c == 0) ? i * dx : (i - 0.5) * dx;
                    double y = (c == 1) ? j * dy : (j - 0.5) * dy;
                    double z = (c == 2) ? k * dz : (k - 0.5) * dz;
                    double s = 0;
                    for (int m = 0; m < nm; m++) {
                        double ph = kt[3 * m] * x + kt[3 * m + 1] * y + kt[3 * m + 2] * z;
                        s += 2.0 * (at[6 * m + 2 * c] * cos(ph) - at[6 * m + 2 * c + 1] * sin(ph));
                    }
                    F[c][AT(g, i, j, k)] = R(s);
                    ss += s * s;
                }
            }
        part[k - 1] = ss;
    }
    double ss = 0;
    for (int k = 0; k < nz; k++) ss += part[k];
    free(part);
    double up = sqrt(ss / ((double)nx * ny * nz) / 3.0);
    double scale = uprime / up;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                for (int c = 0; c < 3; c++) {
                    size_t p = AT(g, i, j, k);
                    F[c][p] = R((double)F[c][p] * scale);
                }
    free(kt);
    free(at);
}

/* Particle initialiser: uniform positions in [0, 2π)^3, Stokes numbers
 * log-uniform in [st_min, st_max] (constant when equal), diameter from
 * St = τp/τη with τη = sqrt(1/(eps Re)) and τp from config.jl:41.  ids 1..n. */
void o_init_particles(uint64_t seed, long n, double st_min, double st_max, double eps, double Re,
                      double density_ratio, real *xs, real *ys, real *zs, real *diams, int64_t *ids)
{
    double taueta = sqrt(1.0 / (eps * Re));
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; p++) {
        xs[p] = R(DOMAIN_LENGTH * o_rand_u01(seed, 0, (uint64_t)p));
        ys[p] = R(DOMAIN_LENGTH * o_rand_u01(seed, 1, (uint64_t)p));
        zs[p] = R(DOMAIN_LENGTH * o_rand_u01(seed, 2, (uint64_t)p));
        double st = st_min;
        if (st_max != st_min) st = st_min * pow(st_max / st_min, o_rand_u01(seed, 3, (uint64_t)p));
        diams[p] = R(sqrt(18.0 * st * taueta / (density_ratio * Re)));
        ids[p] = (int64_t)p + 1;
    }
}
