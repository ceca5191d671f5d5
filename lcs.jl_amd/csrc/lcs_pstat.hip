// Particle statistics on the cell index (Particles.stat, src/particles/stat.jl:80-155):
// summaries of positions, velocities and diameters, the radial distribution
// function over the cell index (src/particles/rdf.jl:1-103), the number-density
// spectrum (src/particles/spectrum.jl:11-78) and the SPH density estimate
// (src/particles/stat.jl:171-231).  Both entry points re-bin the profile first
// (makeindex!), as the reference does.
//
// Distances are Utils.norm (src/utils/fast.jl:42-46) in the particle type.
// This file is compiled without contraction of multiplies into adds (the
// pragma below) and with correctly rounded square roots (sqrtf/sqrt; HIP's
// default), so every operation rounds as the reference's does and pair counts
// and densities are the reference's arithmetic, bit for bit.
#include <algorithm>
#include <cmath>
#include <vector>

#include "lcs_internal.h"

#pragma clang fp contract(off)

namespace lcs {

__device__ inline float psqrt(float a) { return sqrtf(a); }
__device__ inline double psqrt(double a) { return sqrt(a); }

template <typename T>
__device__ inline double pnorm(T dx, T dy, T dz)
{
    return (double)psqrt((dx * dx + dy * dy) + dz * dz);
}

__device__ inline void decode_hash(int64_t hsh, const int nc[3], int ci[3])
{
    int64_t r = hsh - 1;
    ci[0] = (int)(r % (nc[0] + 2));
    r /= nc[0] + 2;
    ci[1] = (int)(r % (nc[1] + 2));
    ci[2] = (int)(r / (nc[1] + 2));
}

__device__ inline double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// ---- summaries: sum, max, min of x, y, z, u, v, w, diam over 1:nvalid -------------
struct PSumArgs {
    const void *a[7];
};

template <typename T>
__global__ void __launch_bounds__(256) k_ps_summary(PSumArgs in, long n, double *part)
{
    __shared__ double lds[3 * 7 * 4];
    double s[7], mx[7], mn[7];
    for (int q = 0; q < 7; q++) {
        s[q] = 0.0;
        mx[q] = -INFINITY;
        mn[q] = INFINITY;
    }
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x)
        for (int q = 0; q < 7; q++) {
            const double v = (double)((const T *)in.a[q])[p];
            s[q] += v;
            mx[q] = fmax(mx[q], v);
            mn[q] = fmin(mn[q], v);
        }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int q = 0; q < 7; q++) {
        double a = wsum(s[q]), b = mx[q], c = mn[q];
        for (int o = 32; o > 0; o >>= 1) {
            b = fmax(b, __shfl_down(b, o, 64));
            c = fmin(c, __shfl_down(c, o, 64));
        }
        if (l == 0) {
            lds[(3 * q) * 4 + w] = a;
            lds[(3 * q + 1) * 4 + w] = b;
            lds[(3 * q + 2) * 4 + w] = c;
        }
    }
    __syncthreads();
    if (threadIdx.x < 21) {
        const int k = threadIdx.x, kind = k % 3;
        double r = lds[k * 4];
        for (int q = 1; q < 4; q++)
            r = kind == 0 ? r + lds[k * 4 + q] : kind == 1 ? fmax(r, lds[k * 4 + q]) : fmin(r, lds[k * 4 + q]);
        part[(long)blockIdx.x * 21 + k] = r;
    }
}

// ---- RDF ---------------------------------------------------------------------------
constexpr int kRdfMaxBins = 1024;

template <typename T>
__global__ void __launch_bounds__(256) k_ps_rdf(long n, const T *xs, const T *ys, const T *zs, const int64_t *ids,
                                               const int64_t *hashes, const int64_t *starts, const int64_t *stops,
                                               int nc0, int nc1, int nc2, int ncell, const double *edges, int nbins,
                                               double diam, double tol, unsigned long long *part)
{
    __shared__ double e_lds[kRdfMaxBins + 1];
    __shared__ unsigned int h_lds[kRdfMaxBins + 1];  // [nbins] = contact pairs
    for (int k = threadIdx.x; k <= nbins; k += blockDim.x) {
        e_lds[k] = edges[k];
        h_lds[k] = 0;
    }
    __syncthreads();
    const int nc[3] = {nc0, nc1, nc2};
    for (long i1 = (long)blockIdx.x * blockDim.x + threadIdx.x; i1 < n; i1 += (long)gridDim.x * blockDim.x) {
        int ci[3];
        decode_hash(hashes[i1], nc, ci);
        if (ci[0] < 1 || ci[0] > nc[0] || ci[1] < 1 || ci[1] > nc[1] || ci[2] < 1 || ci[2] > nc[2]) continue;
        const T x1 = xs[i1], y1 = ys[i1], z1 = zs[i1];
        const int64_t id1 = ids[i1];
        for (int di = -ncell; di <= ncell; di++)
            for (int dj = -ncell; dj <= ncell; dj++)
                for (int dk = -ncell; dk <= ncell; dk++) {
                    const int c0 = ci[0] + di, c1 = ci[1] + dj, c2 = ci[2] + dk;
                    if (c0 < 0 || c0 > nc[0] + 1 || c1 < 0 || c1 > nc[1] + 1 || c2 < 0 || c2 > nc[2] + 1) continue;
                    const int64_t h2 = 1 + c0 + (int64_t)(nc[0] + 2) * (c1 + (int64_t)(nc[1] + 2) * c2);
                    const int64_t j1 = starts[h2 - 1], j2 = stops[h2 - 1];
                    for (int64_t j = j1 - 1; j <= j2 - 1; j++) {
                        if (!(id1 < ids[j])) continue;
                        const double d = pnorm<T>(x1 - xs[j], y1 - ys[j], z1 - zs[j]);
                        int lo = 0, hi = nbins + 1;  // first edge >= d (searchsortedfirst)
                        while (lo < hi) {
                            const int m = (lo + hi) >> 1;
                            if (e_lds[m] < d)
                                lo = m + 1;
                            else
                                hi = m;
                        }
                        if (lo >= 1 && lo <= nbins) atomicAdd(&h_lds[lo - 1], 1u);
                        if (fabs(d / diam - 1) < tol) atomicAdd(&h_lds[nbins], 1u);
                    }
                }
    }
    __syncthreads();
    for (int k = threadIdx.x; k <= nbins; k += blockDim.x) part[(long)blockIdx.x * (nbins + 1) + k] = h_lds[k];
}

// ---- SPH density -------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_ps_density(long n, const T *xs, const T *ys, const T *zs,
                                                   const int64_t *hashes, const int64_t *starts,
                                                   const int64_t *stops, int nc0, int nc1, int nc2, double h,
                                                   double factor, double *dens)
{
    const long i1 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i1 >= n) return;
    const int nc[3] = {nc0, nc1, nc2};
    int ci[3];
    decode_hash(hashes[i1], nc, ci);
    double rho = 0.0;
    if (ci[0] >= 1 && ci[0] <= nc[0] && ci[1] >= 1 && ci[1] <= nc[1] && ci[2] >= 1 && ci[2] <= nc[2]) {
        const T x1 = xs[i1], y1 = ys[i1], z1 = zs[i1];
        const double h2x = 2 * h;
        for (int di = -1; di <= 1; di++)
            for (int dj = -1; dj <= 1; dj++)
                for (int dk = -1; dk <= 1; dk++) {
                    const int64_t hh =
                        1 + (ci[0] + di) + (int64_t)(nc[0] + 2) * ((ci[1] + dj) + (int64_t)(nc[1] + 2) * (ci[2] + dk));
                    const int64_t j1 = starts[hh - 1], j2 = stops[hh - 1];
                    for (int64_t j = j1 - 1; j <= j2 - 1; j++) {
                        const double d = pnorm<T>(x1 - xs[j], y1 - ys[j], z1 - zs[j]);
                        if (!(d < h2x)) continue;
                        // cubic_spline_kernel (stat.jl:220-231)
                        const double q = d / h;
                        double w;
                        if (q < 1) {
                            w = factor * ((1 - 1.5 * (q * q)) + 0.75 * ((q * q) * q));
                        } else {
                            const double t = 2 - q;
                            w = (factor * 0.25) * ((t * t) * t);
                        }
                        rho += w;
                    }
                }
    }
    dens[i1] = rho;
}

// ---- number-density spectrum: sum over particles of cos, sin (k . x) ---------------
template <typename T>
__global__ void __launch_bounds__(256) k_ps_spectrum(long n, const T *xs, const T *ys, const T *zs,
                                                    const double *kvec, double *part)
{
    __shared__ double lds[8];
    const int m = blockIdx.y;
    const double k0 = kvec[3 * m], k1 = kvec[3 * m + 1], k2 = kvec[3 * m + 2];
    double c = 0.0, s = 0.0;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        const double kx = (k0 * (double)xs[p] + k1 * (double)ys[p]) + k2 * (double)zs[p];
        double sv, cv;
        sincos(kx, &sv, &cv);
        c += cv;
        s += sv;
    }
    c = wsum(c);
    s = wsum(s);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        lds[w] = c;
        lds[4 + w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double cc = 0, ss = 0;
        for (int q = 0; q < 4; q++) {
            cc += lds[q];
            ss += lds[4 + q];
        }
        part[2 * ((long)m * gridDim.x + blockIdx.x)] = cc;
        part[2 * ((long)m * gridDim.x + blockIdx.x) + 1] = ss;
    }
}

}  // namespace lcs

using namespace lcs;

// ---- host side --------------------------------------------------------------------
namespace {

const int kWavenumbers[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768};

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

int nshells_of(int max_wavenumber)
{
    int n = 0;
    for (int k : kWavenumbers)
        if (k <= max_wavenumber) n++;
    return n;
}

template <typename T>
int summaries(Box &b, ParticleProfile &pp, double out[21])
{
    const long n = pp.desc.nvalid;
    const int props[7] = {LCS_PROP_X, LCS_PROP_Y, LCS_PROP_Z, LCS_PROP_U, LCS_PROP_V, LCS_PROP_W, LCS_PROP_DIAM};
    PSumArgs in;
    for (int q = 0; q < 7; q++) in.a[q] = pp.props[props[q]];
    const int blocks = (int)std::min<long>(std::max<long>(1, (n + 255) / 256), 1024);
    DevBuf part;
    LCS_HIP(hipMalloc(&part.p, sizeof(double) * 21 * blocks));
    hipLaunchKernelGGL(k_ps_summary<T>, dim3(blocks), dim3(256), 0, b.stream, in, n, (double *)part.p);
    LCS_HIP(hipGetLastError());
    std::vector<double> h((size_t)21 * blocks);
    LCS_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, b.stream));
    LCS_HIP(hipStreamSynchronize(b.stream));
    for (int k = 0; k < 21; k++) {
        const int kind = k % 3;
        double r = kind == 0 ? 0.0 : kind == 1 ? -INFINITY : INFINITY;
        for (int q = 0; q < blocks; q++) {
            const double v = h[(size_t)q * 21 + k];
            r = kind == 0 ? r + v : kind == 1 ? std::max(r, v) : std::min(r, v);
        }
        out[k] = kind == 0 ? r / (double)n : r;
    }
    return LCS_OK;
}

template <typename T>
int rdf(Box &b, ParticleProfile &pp, const lcs_particle_stat_params &p, double diam, const std::vector<double> &edges,
        int64_t *npairs, int64_t *ncontact)
{
    const long n = pp.desc.nvalid;
    const int nb = p.bin_size;
    const int blocks = (int)std::min<long>(std::max<long>(1, (n + 255) / 256), 2048);
    DevBuf part, e;
    LCS_HIP(hipMalloc(&part.p, sizeof(unsigned long long) * (size_t)blocks * (nb + 1)));
    LCS_HIP(hipMalloc(&e.p, sizeof(double) * (nb + 1)));
    LCS_HIP(hipMemcpyAsync(e.p, edges.data(), sizeof(double) * (nb + 1), hipMemcpyHostToDevice, b.stream));
    hipLaunchKernelGGL(k_ps_rdf<T>, dim3(blocks), dim3(256), 0, b.stream, n, (const T *)pp.props[LCS_PROP_X],
                       (const T *)pp.props[LCS_PROP_Y], (const T *)pp.props[LCS_PROP_Z],
                       (const int64_t *)pp.props[LCS_PROP_ID], (const int64_t *)pp.hashes,
                       (const int64_t *)pp.starts, (const int64_t *)pp.stops, pp.nc[0], pp.nc[1], pp.nc[2],
                       p.search_cell_size, (const double *)e.p, nb, diam, p.contact_tolerance,
                       (unsigned long long *)part.p);
    LCS_HIP(hipGetLastError());
    std::vector<unsigned long long> h((size_t)blocks * (nb + 1));
    LCS_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, b.stream));
    LCS_HIP(hipStreamSynchronize(b.stream));
    for (int k = 0; k <= nb; k++) {
        int64_t s = 0;
        for (int q = 0; q < blocks; q++) s += (int64_t)h[(size_t)q * (nb + 1) + k];
        if (k < nb)
            npairs[k] = s;
        else
            *ncontact = s;
    }
    return LCS_OK;
}

template <typename T>
int pspectrum(Box &b, ParticleProfile &pp, int nshells, int nsamples, double *Enp)
{
    const long n = pp.desc.nvalid;
    const int m = nshells * nsamples;
    // wave vectors: WAVENUMBERS[b] times the Fibonacci direction i (spectrum.jl:3-9)
    std::vector<double> kv((size_t)3 * m);
    const double phi = M_PI * (3 - std::sqrt(5.0));
    for (int s = 0; s < nshells; s++)
        for (int i = 0; i < nsamples; i++) {
            const double y = 1 - ((double)i / (double)(nsamples - 1)) * 2;
            const double radius = std::sqrt(1 - y * y), theta = phi * i, kk = kWavenumbers[s];
            double *k = &kv[(size_t)3 * (s * nsamples + i)];
            k[0] = kk * (std::cos(theta) * radius);
            k[1] = kk * y;
            k[2] = kk * (std::sin(theta) * radius);
        }
    const int bx = (int)std::min<long>(std::max<long>(1, (n + 255) / 256), 64);
    DevBuf part, kd;
    LCS_HIP(hipMalloc(&part.p, sizeof(double) * 2 * (size_t)bx * m));
    LCS_HIP(hipMalloc(&kd.p, sizeof(double) * kv.size()));
    LCS_HIP(hipMemcpyAsync(kd.p, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice, b.stream));
    hipLaunchKernelGGL(k_ps_spectrum<T>, dim3(bx, m), dim3(256), 0, b.stream, n, (const T *)pp.props[LCS_PROP_X],
                       (const T *)pp.props[LCS_PROP_Y], (const T *)pp.props[LCS_PROP_Z], (const double *)kd.p,
                       (double *)part.p);
    LCS_HIP(hipGetLastError());
    std::vector<double> h((size_t)2 * bx * m);
    LCS_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, b.stream));
    LCS_HIP(hipStreamSynchronize(b.stream));
    const double N = (double)n;
    for (int s = 0; s < nshells; s++) {
        double ph = 0.0;
        for (int i = 0; i < nsamples; i++) {
            double c = 0, sn = 0;
            for (int q = 0; q < bx; q++) {
                c += h[2 * ((size_t)(s * nsamples + i) * bx + q)];
                sn += h[2 * ((size_t)(s * nsamples + i) * bx + q) + 1];
            }
            ph += (c * c + sn * sn) / (N * N) - 1 / N;
        }
        const double kk = kWavenumbers[s];
        Enp[s] = ph / nsamples * 4 * M_PI * (kk * kk);
    }
    return LCS_OK;
}

template <typename T>
int density(Box &b, ParticleProfile &pp, double *out)
{
    const long n = pp.desc.nvalid;
    if (n == 0) return LCS_OK;
    const double h = std::min(std::min(spacing(b.g.nx), spacing(b.g.ny)), spacing(b.g.nz));
    const double factor = 1 / (M_PI * (h * h * h));
    DevBuf d;
    LCS_HIP(hipMalloc(&d.p, sizeof(double) * n));
    hipLaunchKernelGGL(k_ps_density<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b.stream, n,
                       (const T *)pp.props[LCS_PROP_X], (const T *)pp.props[LCS_PROP_Y],
                       (const T *)pp.props[LCS_PROP_Z], (const int64_t *)pp.hashes, (const int64_t *)pp.starts,
                       (const int64_t *)pp.stops, pp.nc[0], pp.nc[1], pp.nc[2], h, factor, (double *)d.p);
    LCS_HIP(hipGetLastError());
    LCS_HIP(hipMemcpyAsync(out, d.p, sizeof(double) * n, hipMemcpyDeviceToHost, b.stream));
    LCS_HIP(hipStreamSynchronize(b.stream));
    return LCS_OK;
}

}  // namespace

namespace lcs {

int particle_stat(Box &b, ParticleProfile &pp, const lcs_particle_stat_params &p, double tau_eta,
                  lcs_particle_stats &o, double *edges, int64_t *npairs, double *gr, double *wavenumbers, double *Enp)
{
    if (p.bin_size < 1 || p.bin_size > kRdfMaxBins) return fail(LCS_ERR_INVALID, "rdf bin_size must be in 1..1024");
    if (!(p.min_radius > 0) || p.search_cell_size < 1 || !(p.contact_tolerance > 0))
        return fail(LCS_ERR_INVALID, "rdf needs min_radius > 0, search_cell_size >= 1, contact_tolerance > 0");
    if (p.max_wavenumber > 0 && p.nsamples < 2) return fail(LCS_ERR_INVALID, "spectrum needs nsamples >= 2");
    if (pp.desc.nvalid < 2) return fail(LCS_ERR_INVALID, "particle statistics need at least two particles");
    int rc = launch_p_makeindex(b, pp, 1);  // stat() re-bins first (stat.jl:146)
    if (rc) return rc;
    double s[21];
    rc = b.esize == 8 ? summaries<double>(b, pp, s) : summaries<float>(b, pp, s);
    if (rc) return rc;
    double *f[21] = {&o.x_mean, &o.x_max, &o.x_min, &o.y_mean, &o.y_max, &o.y_min, &o.z_mean,
                     &o.z_max,  &o.z_min, &o.u_mean, &o.u_max, &o.u_min, &o.v_mean, &o.v_max,
                     &o.v_min,  &o.w_mean, &o.w_max, &o.w_min, &o.diam_mean, &o.diam_max, &o.diam_min};
    for (int k = 0; k < 21; k++) *f[k] = s[k];
    // taup of the mean diameter (src/particles/config.jl:41), St = taup / tau_eta
    o.taup = 1.0 / 18 * pp.desc.density_ratio * b.Re() * (o.diam_mean * o.diam_mean);
    o.St = o.taup / tau_eta;
    // RDF: log-spaced edges from min_radius to search_cell_size fine spacings
    const int nb = p.bin_size;
    const double rmax = p.search_cell_size * std::min(std::min(spacing(b.g.nx), spacing(b.g.ny)), spacing(b.g.nz));
    std::vector<double> e(nb + 1);
    const double l0 = std::log(p.min_radius), l1 = std::log(rmax);
    for (int k = 0; k <= nb; k++) e[k] = std::exp(l0 + (l1 - l0) * ((double)k / nb));
    e[0] = p.min_radius;
    e[nb] = rmax;
    int64_t ncontact = 0;
    std::vector<int64_t> np(nb);
    rc = b.esize == 8 ? rdf<double>(b, pp, p, o.diam_mean, e, np.data(), &ncontact)
                      : rdf<float>(b, pp, p, o.diam_mean, e, np.data(), &ncontact);
    if (rc) return rc;
    const double N = (double)pp.desc.nvalid, vol = kDomainLength * kDomainLength * kDomainLength;
    const double rho0 = (N * (N - 1) / 2) / vol;  // binomial(nvalid, 2) / volume
    for (int k = 0; k < nb; k++) {
        if (edges) edges[k] = e[k];
        if (npairs) npairs[k] = np[k];
        const double shell = (4.0 / 3.0) * M_PI * (e[k + 1] * e[k + 1] * e[k + 1] - e[k] * e[k] * e[k]);
        if (gr) gr[k] = ((double)np[k] / shell) / rho0;
    }
    if (edges) edges[nb] = e[nb];
    const double contact_shell = (4.0 / 3.0) * M_PI * (o.diam_mean * o.diam_mean * o.diam_mean) * 6 * p.contact_tolerance;
    o.npairs_contact = ncontact;
    o.gr_contact = ((double)ncontact / contact_shell) / rho0;
    o.nbins = nb;
    // number-density spectrum
    o.nshells = p.max_wavenumber > 0 ? nshells_of(p.max_wavenumber) : 0;
    if (o.nshells > 0) {
        std::vector<double> en(o.nshells);
        rc = b.esize == 8 ? pspectrum<double>(b, pp, o.nshells, p.nsamples, en.data())
                          : pspectrum<float>(b, pp, o.nshells, p.nsamples, en.data());
        if (rc) return rc;
        for (int s2 = 0; s2 < o.nshells; s2++) {
            if (wavenumbers) wavenumbers[s2] = kWavenumbers[s2];
            if (Enp) Enp[s2] = en[s2];
        }
    }
    return LCS_OK;
}

int particle_density(Box &b, ParticleProfile &pp, double *out)
{
    int rc = launch_p_makeindex(b, pp, 1);  // density() re-bins first (stat.jl:214)
    if (rc) return rc;
    return b.esize == 8 ? density<double>(b, pp, out) : density<float>(b, pp, out);
}

}  // namespace lcs
