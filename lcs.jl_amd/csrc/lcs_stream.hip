// Achievable HBM bandwidth on this device: hand-written streaming kernels
// (read, write, copy, triad) over arrays of several GiB.  They carry no
// physics; bench.py reports their rates beside the stencil kernels' rooflines
// so a kernel's fraction of the 8 TB/s peak can be read against what a pure
// stream reaches on the same card.  Built as its own library
// (liblcs_stream.so) outside the C ABI of the solver.
//
// Every kernel moves 16-byte vectors, a workgroup of 256 threads owning
// 256*U consecutive vectors: it issues all U loads before the first store,
// so each wave keeps U loads in flight.  Variants differ in U and in
// nontemporal (streaming) loads and stores; the probe reports the best.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace {

constexpr int kThreads = 256;

template <int U, bool NT>
__device__ inline float4 ld(const float4 *p)
{
    if constexpr (NT) {
        float4 v;
        v.x = __builtin_nontemporal_load(&p->x);
        v.y = __builtin_nontemporal_load(&p->y);
        v.z = __builtin_nontemporal_load(&p->z);
        v.w = __builtin_nontemporal_load(&p->w);
        return v;
    } else {
        return *p;
    }
}

template <int U, bool NT>
__device__ inline void st(float4 *p, float4 v)
{
    if constexpr (NT) {
        __builtin_nontemporal_store(v.x, &p->x);
        __builtin_nontemporal_store(v.y, &p->y);
        __builtin_nontemporal_store(v.z, &p->z);
        __builtin_nontemporal_store(v.w, &p->w);
    } else {
        *p = v;
    }
}

// b = a
template <int U, bool NT>
__global__ void __launch_bounds__(kThreads) k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, long n)
{
    const long i0 = (long)blockIdx.x * (kThreads * U) + threadIdx.x;
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++)
        if (i0 + (long)u * kThreads < n) v[u] = ld<U, NT>(a + i0 + (long)u * kThreads);
#pragma unroll
    for (int u = 0; u < U; u++)
        if (i0 + (long)u * kThreads < n) st<U, NT>(b + i0 + (long)u * kThreads, v[u]);
}

// a = b + s c
template <int U, bool NT>
__global__ void __launch_bounds__(kThreads) k_triad(float4 *__restrict__ a, const float4 *__restrict__ b,
                                                    const float4 *__restrict__ c, float s, long n)
{
    const long i0 = (long)blockIdx.x * (kThreads * U) + threadIdx.x;
    float4 x[U], y[U];
#pragma unroll
    for (int u = 0; u < U; u++)
        if (i0 + (long)u * kThreads < n) {
            x[u] = ld<U, NT>(b + i0 + (long)u * kThreads);
            y[u] = ld<U, NT>(c + i0 + (long)u * kThreads);
        }
#pragma unroll
    for (int u = 0; u < U; u++)
        if (i0 + (long)u * kThreads < n) {
            float4 r;
            r.x = x[u].x + s * y[u].x;
            r.y = x[u].y + s * y[u].y;
            r.z = x[u].z + s * y[u].z;
            r.w = x[u].w + s * y[u].w;
            st<U, NT>(a + i0 + (long)u * kThreads, r);
        }
}

// one partial sum per workgroup (the result keeps the loads live)
template <int U, bool NT>
__global__ void __launch_bounds__(kThreads) k_read(const float4 *__restrict__ a, float *__restrict__ part, long n)
{
    __shared__ float lds[kThreads / 64];
    const long i0 = (long)blockIdx.x * (kThreads * U) + threadIdx.x;
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++)
        v[u] = (i0 + (long)u * kThreads < n) ? ld<U, NT>(a + i0 + (long)u * kThreads) : float4{0, 0, 0, 0};
    float s = 0;
#pragma unroll
    for (int u = 0; u < U; u++) s += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

template <int U, bool NT>
__global__ void __launch_bounds__(kThreads) k_write(float4 *__restrict__ a, float v, long n)
{
    const long i0 = (long)blockIdx.x * (kThreads * U) + threadIdx.x;
#pragma unroll
    for (int u = 0; u < U; u++)
        if (i0 + (long)u * kThreads < n) st<U, NT>(a + i0 + (long)u * kThreads, float4{v, v, v, v});
}

template <int U, bool NT>
hipError_t launch(int kind, float4 *a, float4 *b, float4 *c, float *part, long n, hipStream_t s)
{
    const unsigned blocks = (unsigned)((n + (long)kThreads * U - 1) / ((long)kThreads * U));
    switch (kind) {
    case 0: hipLaunchKernelGGL((k_read<U, NT>), dim3(blocks), dim3(kThreads), 0, s, a, part, n); break;
    case 1: hipLaunchKernelGGL((k_write<U, NT>), dim3(blocks), dim3(kThreads), 0, s, a, 1.0f, n); break;
    case 2: hipLaunchKernelGGL((k_copy<U, NT>), dim3(blocks), dim3(kThreads), 0, s, a, b, n); break;
    default: hipLaunchKernelGGL((k_triad<U, NT>), dim3(blocks), dim3(kThreads), 0, s, a, b, c, 0.5f, n); break;
    }
    return hipGetLastError();
}

// b = a one float per thread: the width of the stencil kernels' own loads
template <bool NT>
__global__ void __launch_bounds__(kThreads) k_copy_dword(const float *__restrict__ a, float *__restrict__ b, long n)
{
    const long i = (long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    if constexpr (NT)
        __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
    else
        b[i] = a[i];
}

// the same with U floats per thread at a stride of one workgroup: as many
// workgroups (U = 4) as the 16-byte copy with one vector per thread, so the
// two differ only in the width of each load and store
template <int U>
__global__ void __launch_bounds__(kThreads) k_copy_dwords(const float *__restrict__ a, float *__restrict__ b, long n)
{
    const long i0 = (long)blockIdx.x * (kThreads * U) + threadIdx.x;
    float v[U];
#pragma unroll
    for (int u = 0; u < U; u++)
        if (i0 + (long)u * kThreads < n) v[u] = a[i0 + (long)u * kThreads];
#pragma unroll
    for (int u = 0; u < U; u++)
        if (i0 + (long)u * kThreads < n) b[i0 + (long)u * kThreads] = v[u];
}

hipError_t launch_variant(int variant, int kind, float4 *a, float4 *b, float4 *c, float *part, long n, hipStream_t s)
{
    if (variant >= 8) {  // dword copies (kind 2 only)
        const long m = 4 * n;
        const int per = variant == 10 ? 4 : (variant == 11 ? 16 : 1);
        const unsigned blocks = (unsigned)((m + (long)kThreads * per - 1) / ((long)kThreads * per));
        if (variant == 8)
            hipLaunchKernelGGL(k_copy_dword<false>, dim3(blocks), dim3(kThreads), 0, s, (const float *)a, (float *)b, m);
        else if (variant == 9)
            hipLaunchKernelGGL(k_copy_dword<true>, dim3(blocks), dim3(kThreads), 0, s, (const float *)a, (float *)b, m);
        else if (variant == 10)
            hipLaunchKernelGGL(k_copy_dwords<4>, dim3(blocks), dim3(kThreads), 0, s, (const float *)a, (float *)b, m);
        else
            hipLaunchKernelGGL(k_copy_dwords<16>, dim3(blocks), dim3(kThreads), 0, s, (const float *)a, (float *)b, m);
        return hipGetLastError();
    }
    switch (variant) {
    case 0: return launch<1, false>(kind, a, b, c, part, n, s);
    case 1: return launch<2, false>(kind, a, b, c, part, n, s);
    case 2: return launch<4, false>(kind, a, b, c, part, n, s);
    case 3: return launch<8, false>(kind, a, b, c, part, n, s);
    case 4: return launch<1, true>(kind, a, b, c, part, n, s);
    case 5: return launch<2, true>(kind, a, b, c, part, n, s);
    case 6: return launch<4, true>(kind, a, b, c, part, n, s);
    default: return launch<8, true>(kind, a, b, c, part, n, s);
    }
}

}  // namespace

extern "C" {

// Number of kernel variants lcs_stream_probe tries per kind: 16-byte vectors
// (U = 1, 2, 4, 8 per thread; plain, then nontemporal), then for copies one
// float per thread (plain, nontemporal) and 4 and 16 floats per thread.
int lcs_stream_variants(void) { return 12; }

// Times kind (0 read, 1 write, 2 copy b = a, 3 triad a = b + s c) over arrays
// of `bytes` each on `device`, variant by variant, `reps` launches each after
// one warm-up.  gbs[v] (12 entries) receives variant v's best rate in GB/s
// (bytes read + written per launch / fastest launch), ms[v] that launch's
// time; variants a kind does not have stay 0.  Returns 0 or a hipError_t.
int lcs_stream_probe(int device, int kind, size_t bytes, int reps, double *gbs, double *ms)
{
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    const long n = (long)(bytes / sizeof(float4));
    if (n < 1 || kind < 0 || kind > 3 || reps < 1) return (int)hipErrorInvalidValue;
    const int narr = kind == 3 ? 3 : (kind == 2 ? 2 : 1);
    float4 *arr[3] = {nullptr, nullptr, nullptr};
    float *part = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const long maxblocks = (n + kThreads - 1) / kThreads;
    for (int q = 0; q < narr && e == hipSuccess; q++) e = hipMalloc(&arr[q], (size_t)n * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&part, sizeof(float) * (size_t)maxblocks);
    if (e == hipSuccess) e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int q = 0; q < narr && e == hipSuccess; q++) e = hipMemsetAsync(arr[q], 0, (size_t)n * sizeof(float4), s);
    const double moved = (double)n * sizeof(float4) * (kind == 3 ? 3.0 : (kind == 2 ? 2.0 : 1.0));
    const int nvar = kind == 2 ? 12 : 8;
    for (int v = 0; v < 12; v++) gbs[v] = ms[v] = 0.0;
    for (int v = 0; v < nvar && e == hipSuccess; v++) {
        float best = 1e30f;
        for (int r = 0; r <= reps && e == hipSuccess; r++) {
            e = hipEventRecord(e0, s);
            if (e == hipSuccess) e = launch_variant(v, kind, arr[0], arr[1], arr[2], part, n, s);
            if (e == hipSuccess) e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (r > 0) best = std::min(best, t);
        }
        gbs[v] = moved / (best * 1e-3) / 1e9;
        ms[v] = best;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    for (int q = 0; q < 3; q++)
        if (arr[q]) (void)hipFree(arr[q]);
    if (part) (void)hipFree(part);
    return (int)e;
}

}  // extern "C"
