// kmx_gmres.hip -- the orthogonalisation step of a restarted GMRES on the device,
// for the matrix-free backward-Euler solve that replaces the direct solve of
// time_advance.m:328-349 ((I - dt A) f1 = f0 + dt (s + bc)).
//
// Step k of the Arnoldi process: w = s0 v_k + s1 u (u = A v_k, applied by the
// caller), then classical Gram-Schmidt against v_0 .. v_k twice (h = V w,
// w -= V^T h, again, the two coefficient vectors summed), then v_{k+1} = w / |w|.
// Four launches, each over the same G contiguous chunks of the vectors, one
// workgroup per chunk: the dot products of a pass leave one partial per
// (chunk, row), and every workgroup of the next launch sums the G partials of
// each row itself (G <= 128: at most 128 x 128 values, from L2, four threads per
// row each summing a quarter), so no launch waits on a separate reduction.  Every launch has a fixed shape for given (n, k) and
// allocates nothing, so the step can be captured in a graph.
#include "kmx_internal.h"

#include <algorithm>

namespace kmx {
namespace {

constexpr int AR_THREADS = 256;
constexpr int AR_MAX_G = 128;

struct ArnoldiArgs {
    const double* u;
    double* V;  // rows 0 .. k + 1, row j at V + j * n
    double* w;
    double* part1;  // [G][k + 1]
    double* part2;  // [G][k + 1]
    double* part3;  // [G]
    double* h;      // [k + 2]
    long long n, chunk;
    double s0, s1;
    int k, G;
};

__device__ __forceinline__ long long chunk_end(const ArnoldiArgs& a, long long i0) {
    return i0 + a.chunk < a.n ? i0 + a.chunk : a.n;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// partial dots of rows 0..k with w over this workgroup's chunk: wave q takes rows
// q, q + 4, q + 8, ..., AR_RB of them at a time, so a lane keeps AR_RB independent
// sums and loads in flight
constexpr int AR_RB = 8;
constexpr int AR_WAVES = AR_THREADS / 64;
__device__ __forceinline__ void chunk_dots(const ArnoldiArgs& a, long long i0, long long i1, double* part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j0 = wave; j0 <= a.k; j0 += AR_WAVES * AR_RB) {
        double s[AR_RB];
#pragma unroll
        for (int r = 0; r < AR_RB; ++r) s[r] = 0.0;
#pragma unroll 2
        for (long long i = i0 + lane; i < i1; i += 64) {
            const double wi = a.w[i];
#pragma unroll
            for (int r = 0; r < AR_RB; ++r)
                if (j0 + AR_WAVES * r <= a.k) s[r] = __builtin_fma(a.V[(long long)(j0 + AR_WAVES * r) * a.n + i], wi, s[r]);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int r = 0; r < AR_RB; ++r) s[r] += __shfl_xor(s[r], off);
        if (lane == 0)
#pragma unroll
            for (int r = 0; r < AR_RB; ++r)
                if (j0 + AR_WAVES * r <= a.k) part[(long long)blockIdx.x * (a.k + 1) + j0 + AR_WAVES * r] = s[r];
    }
}

// the G partials of each row summed, into LDS: rows 64 at a time, each by four
// threads that sum every fourth partial, their sums added in slice order
__device__ __forceinline__ void sum_partials(const ArnoldiArgs& a, const double* part, double* sh) {
    __shared__ double red[AR_WAVES][64];
    const int lane = threadIdx.x & 63, sl = threadIdx.x >> 6;
    for (int j0 = 0; j0 <= a.k; j0 += 64) {
        const int j = j0 + lane;
        double s = 0.0;
        if (j <= a.k) {
#pragma unroll 8
            for (int g = sl; g < a.G; g += AR_WAVES) s += part[(long long)g * (a.k + 1) + j];
        }
        red[sl][lane] = s;
        __syncthreads();
        if (sl == 0 && j <= a.k) {
            double t = red[0][lane];
#pragma unroll
            for (int q = 1; q < AR_WAVES; ++q) t += red[q][lane];
            sh[j] = t;
        }
        __syncthreads();
    }
}

// w -= V^T h over this workgroup's chunk
__device__ __forceinline__ void chunk_update(const ArnoldiArgs& a, long long i0, long long i1, const double* sh) {
    for (long long i = i0 + threadIdx.x; i < i1; i += AR_THREADS) {
        double s = a.w[i];
#pragma unroll 8
        for (int j = 0; j <= a.k; ++j) s = __builtin_fma(-a.V[(long long)j * a.n + i], sh[j], s);
        a.w[i] = s;
    }
    __syncthreads();  // the chunk's new w before the dots that read it (same workgroup)
}

__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_first(ArnoldiArgs a) {
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0);
    const double* vk = a.V + (long long)a.k * a.n;
    for (long long i = i0 + threadIdx.x; i < i1; i += AR_THREADS) a.w[i] = a.s0 * vk[i] + a.s1 * a.u[i];
    __syncthreads();
    chunk_dots(a, i0, i1, a.part1);
}

__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_second(ArnoldiArgs a) {
    __shared__ double sh[KMX_ARNOLDI_MAX_K + 1];
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0);
    sum_partials(a, a.part1, sh);
    chunk_update(a, i0, i1, sh);
    chunk_dots(a, i0, i1, a.part2);
}

__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_third(ArnoldiArgs a) {
    __shared__ double sh[KMX_ARNOLDI_MAX_K + 1];
    __shared__ double red[AR_THREADS / 64];
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0);
    sum_partials(a, a.part2, sh);
    chunk_update(a, i0, i1, sh);
    double s = 0.0;
    for (long long i = i0 + threadIdx.x; i < i1; i += AR_THREADS) s = __builtin_fma(a.w[i], a.w[i], s);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < AR_THREADS / 64; ++q) t += red[q];
        a.part3[blockIdx.x] = t;
    }
    if (blockIdx.x == 0) {  // h = the two passes' coefficients (part1, part2 are complete here)
        __shared__ double sh1[KMX_ARNOLDI_MAX_K + 1];
        sum_partials(a, a.part1, sh1);
        for (int j = threadIdx.x; j <= a.k; j += AR_THREADS) a.h[j] = sh1[j] + sh[j];
    }
}

__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_normalize(ArnoldiArgs a) {
    __shared__ double nrm;
    if (threadIdx.x < 64) {  // the G norm partials, two per lane, summed across the wave
        const int l = threadIdx.x;
        double t = (l < a.G ? a.part3[l] : 0.0) + (l + 64 < a.G ? a.part3[l + 64] : 0.0);
        t = wave_sum(t);
        if (l == 0) {
            nrm = sqrt(t);
            if (blockIdx.x == 0) a.h[a.k + 1] = nrm;
        }
    }
    __syncthreads();
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0);
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;  // breakdown: a zero row, h[k + 1] = 0 tells the caller
    double* vn = a.V + (long long)(a.k + 1) * a.n;
    for (long long i = i0 + threadIdx.x; i < i1; i += AR_THREADS) vn[i] = a.w[i] * inv;
}

}  // namespace
}  // namespace kmx

using namespace kmx;

extern "C" int kmx_arnoldi_step(int64_t n, int k, double* V, const double* u, double s0, double s1, double* h,
                                double* work, void* stream) {
    if (n <= 0 || k < 0 || k > KMX_ARNOLDI_MAX_K || !V || !u || !h || !work) {
        set_error("kmx_arnoldi_step: bad arguments (n > 0, 0 <= k <= KMX_ARNOLDI_MAX_K, device pointers)");
        return KMX_E_ARG;
    }
    ArnoldiArgs a;
    a.n = n;
    a.k = k;
    a.G = int(std::min<long long>(AR_MAX_G, (n + AR_THREADS - 1) / AR_THREADS));
    a.chunk = (n + a.G - 1) / a.G;
    a.G = int((n + a.chunk - 1) / a.chunk);  // every workgroup holds at least one element
    a.u = u;
    a.V = V;
    a.w = work;
    a.part1 = work + n;
    a.part2 = a.part1 + (long long)AR_MAX_G * (KMX_ARNOLDI_MAX_K + 1);
    a.part3 = a.part2 + (long long)AR_MAX_G * (KMX_ARNOLDI_MAX_K + 1);
    a.h = h;
    a.s0 = s0;
    a.s1 = s1;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid(unsigned(a.G)), block(AR_THREADS);
    hipLaunchKernelGGL(k_arnoldi_first, grid, block, 0, s, a);
    hipLaunchKernelGGL(k_arnoldi_second, grid, block, 0, s, a);
    hipLaunchKernelGGL(k_arnoldi_third, grid, block, 0, s, a);
    hipLaunchKernelGGL(k_arnoldi_normalize, grid, block, 0, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("kmx_arnoldi_step: ") + hipGetErrorString(e));
        return KMX_E_HIP;
    }
    return KMX_OK;
}
