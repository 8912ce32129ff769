// kmx_gmres.hip -- the orthogonalisation step of a restarted GMRES on the device,
// for the matrix-free backward-Euler solve that replaces the direct solve of
// time_advance.m:328-349 ((I - dt A) f1 = f0 + dt (s + bc)).
//
// Step k of the Arnoldi process: w = s0 v_k + s1 u (u = A v_k, applied by the
// caller), then classical Gram-Schmidt against v_0 .. v_k twice (h1 = V w,
// w1 = w - V^T h1; h2 = V w1, w2 = w1 - V^T h2; h = h1 + h2), and
// v_{k+1} = w2 / |w2| with |w2|^2 = |w1|^2 - |h2|^2 (V orthonormal: w2 is w1
// less its projection h2; after the first pass h2 is tiny against w1, so the
// difference does not cancel -- the Pythagorean norm of CGS2, Smoktunowicz,
// Barlow and Langou, Numer. Math. 105 (2006)).  Three launches, each over the same
// G contiguous chunks of the vectors, one workgroup of 16 waves per chunk: the dot
// products of a pass leave one partial per (chunk, row), |w1|^2 riding along with
// the second pass's, and every workgroup of the next launch sums the G partials of
// each row itself (G <= 128: 16 threads per row each summing at most 8), so no
// launch waits on a separate reduction.  Every launch has a fixed shape for given
// (n, k) and allocates nothing, so the step can be captured in a graph.
#include "kmx_internal.h"

#include <algorithm>

namespace kmx {
namespace {

constexpr int AR_THREADS = 1024;
constexpr int AR_WAVES = AR_THREADS / 64;
constexpr int AR_MAX_G = 128;
constexpr int AR_RB = 8;  // rows a lane keeps in flight in the dot products
constexpr int AR_ROWS = KMX_ARNOLDI_MAX_K + 2;  // rows of the second pass: v_0 .. v_k and w1
constexpr int AR_MIN_CHUNK = 256;  // elements per workgroup at least (while G < AR_MAX_G)

struct ArnoldiArgs {
    const double* u;
    double* V;  // rows 0 .. k + 1, row j at V + j * n
    double* w;
    double* part1;  // [G][k + 1]
    double* part2;  // [G][k + 2]: the second pass's dots and |w1|^2 last
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

// partial dots over this workgroup's chunk of rows 0 .. R - 1 with w, row j being
// v_j for j <= k and w itself for j = k + 1 (R = k + 2: |w|^2).  The 16 waves
// split into RGp row groups (wave q takes rows q % RGp + RGp r, AR_RB of them at a
// time) times 16 / RGp element slices (every (16 / RGp)-th run of 64 elements), so
// a short row list still has every wave loading; the slices' sums are added in
// slice order through LDS
__device__ __forceinline__ void chunk_dots(const ArnoldiArgs& a, long long i0, long long i1, int R, double* part) {
    __shared__ double red[AR_WAVES * AR_RB];  // [slice][row of the pass]: ES * RGp * AR_RB values
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rg_need = (R + AR_RB - 1) / AR_RB;
    int RGp = 1;
    while (RGp < rg_need && RGp < AR_WAVES) RGp <<= 1;
    const int ES = AR_WAVES / RGp, rg = wave % RGp, es = wave / RGp;
    const int per_pass = RGp * AR_RB;
    for (int j0 = 0; j0 < R; j0 += per_pass) {
        double s[AR_RB];
        const double* row[AR_RB];
#pragma unroll
        for (int r = 0; r < AR_RB; ++r) {
            s[r] = 0.0;
            const int j = j0 + rg + RGp * r;
            row[r] = j <= a.k ? a.V + (long long)j * a.n : a.w;  // j = k + 1: w (read only when j < R)
        }
#pragma unroll 2
        for (long long i = i0 + es * 64 + lane; i < i1; i += ES * 64) {
            const double wi = a.w[i];
#pragma unroll
            for (int r = 0; r < AR_RB; ++r)
                if (j0 + rg + RGp * r < R) s[r] = __builtin_fma(row[r][i], wi, s[r]);
        }
#pragma unroll
        for (int r = 0; r < AR_RB; ++r) s[r] = wave_sum(s[r]);
        if (lane == 0)
#pragma unroll
            for (int r = 0; r < AR_RB; ++r) red[es * per_pass + rg + RGp * r] = s[r];
        __syncthreads();
        if (int(threadIdx.x) < per_pass && j0 + int(threadIdx.x) < R) {
            double t = 0.0;
            for (int e = 0; e < ES; ++e) t += red[e * per_pass + threadIdx.x];
            part[(long long)blockIdx.x * R + j0 + threadIdx.x] = t;
        }
        __syncthreads();
    }
}

// the G partials of each of rows 0 .. R - 1 summed, into LDS: rows 64 at a time,
// each by 16 threads that sum every 16th partial, their sums added in slice order
__device__ __forceinline__ void sum_partials(const ArnoldiArgs& a, const double* part, int R, double* sh) {
    __shared__ double red[AR_WAVES][64];
    const int lane = threadIdx.x & 63, sl = threadIdx.x >> 6;
    for (int j0 = 0; j0 < R; j0 += 64) {
        const int j = j0 + lane;
        double s = 0.0;
        if (j < R) {
#pragma unroll
            for (int g = sl; g < AR_MAX_G; g += AR_WAVES)
                if (g < a.G) s += part[(long long)g * R + j];
        }
        red[sl][lane] = s;
        __syncthreads();
        if (sl == 0 && j < R) {
            double t = red[0][lane];
#pragma unroll
            for (int q = 1; q < AR_WAVES; ++q) t += red[q][lane];
            sh[j] = t;
        }
        __syncthreads();
    }
}

// out = (w - V^T c) * scale over this workgroup's chunk (c: rows 0..k, in LDS);
// 256 elements at a time, the rows split four ways (slice q takes rows q, q + 4,
// ...), the slices' sums added in slice order.  out may be w itself (each element
// is read and written by one thread).
__device__ __forceinline__ void chunk_update(const ArnoldiArgs& a, long long i0, long long i1, const double* c,
                                             double* out, double scale) {
    constexpr int EL = AR_THREADS / 4;
    __shared__ double red[4][EL];
    const int el = threadIdx.x % EL, q = threadIdx.x / EL;
    for (long long ib = i0; ib < i1; ib += EL) {
        const long long i = ib + el;
        double p = 0.0;
        if (i < i1) {
#pragma unroll 8
            for (int j = q; j <= a.k; j += 4) p = __builtin_fma(a.V[(long long)j * a.n + i], c[j], p);
        }
        red[q][el] = p;
        __syncthreads();
        if (q == 0 && i < i1) out[i] = (a.w[i] - (((red[0][el] + red[1][el]) + red[2][el]) + red[3][el])) * scale;
        __syncthreads();  // the chunk's new values before the dots that read them (same workgroup)
    }
}

__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_first(ArnoldiArgs a) {
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0);
    const double* vk = a.V + (long long)a.k * a.n;
    for (long long i = i0 + threadIdx.x; i < i1; i += AR_THREADS) a.w[i] = a.s0 * vk[i] + a.s1 * a.u[i];
    __syncthreads();
    chunk_dots(a, i0, i1, a.k + 1, a.part1);
}

__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_second(ArnoldiArgs a) {
    __shared__ double sh[AR_ROWS];
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0);
    sum_partials(a, a.part1, a.k + 1, sh);
    if (blockIdx.x == 0)  // h1 (the third launch adds h2)
        for (int j = threadIdx.x; j <= a.k; j += AR_THREADS) a.h[j] = sh[j];
    chunk_update(a, i0, i1, sh, a.w, 1.0);  // w1, in place
    chunk_dots(a, i0, i1, a.k + 2, a.part2);
}

__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_third(ArnoldiArgs a) {
    __shared__ double sh[AR_ROWS];
    __shared__ double nrm;
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0);
    sum_partials(a, a.part2, a.k + 2, sh);  // h2 in sh[0..k], |w1|^2 in sh[k + 1]
    if (threadIdx.x < 64) {
        double t = 0.0;
        for (int j = threadIdx.x; j <= a.k; j += 64) t = __builtin_fma(sh[j], sh[j], t);
        t = wave_sum(t);
        if (threadIdx.x == 0) {
            const double q = sh[a.k + 1] - t;
            nrm = q > 0.0 ? sqrt(q) : 0.0;
        }
    }
    __syncthreads();
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;  // breakdown: a zero row, h[k + 1] = 0 tells the caller
    chunk_update(a, i0, i1, sh, a.V + (long long)(a.k + 1) * a.n, inv);
    if (blockIdx.x == 0) {  // h = the two passes' coefficients (h1 from the second launch)
        for (int j = threadIdx.x; j <= a.k; j += AR_THREADS) a.h[j] += sh[j];
        if (threadIdx.x == 0) a.h[a.k + 1] = nrm;
    }
}

// Short vectors (chunk <= 256, n <= AR_MAX_G * 256) and k <= 126: the same three
// launches with each thread holding its element's values of the rows it takes in
// registers, loaded once per launch and used by both the update and the dot
// products, so each launch waits on one round of loads.  Thread t takes element
// t % 256 of the chunk and rows q, q + 4, ... (q = t / 256), RB of them
// (4 RB >= k + 2, the last row of the second pass being |w1|^2).
constexpr int AR_REG_EL = AR_THREADS / 4;

template <int RB>
struct RegRows {
    double v[RB];
};

template <int RB>
__device__ __forceinline__ void reg_load(const ArnoldiArgs& a, long long i, bool ok, int q, RegRows<RB>& r) {
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        const int j = q + 4 * b;
        r.v[b] = (ok && j <= a.k) ? a.V[(long long)j * a.n + i] : 0.0;
    }
}

// partials over the chunk of rows 0..R-1 into part[blockIdx.x * R + j]: row j <= k
// is v_j . w (wi: the thread's element of w), row k + 1 (when R = k + 2) is |w|^2.
// The products go through LDS 32 rows at a time ([row][element]); then 32 lanes
// per row each sum every 32nd element of it and the 32 lanes' sums are added
// across the half-wave, every row at once
template <int RB>
__device__ __forceinline__ void reg_dots(const ArnoldiArgs& a, const RegRows<RB>& r, double wi, int q, int el, int R,
                                         double* part) {
    __shared__ double prod[32][AR_REG_EL];
    constexpr int NP = (4 * RB + 31) / 32;
    const int row = threadIdx.x >> 5, s = threadIdx.x & 31;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        if (p) __syncthreads();  // the previous pass's reads are done
#pragma unroll
        for (int bb = 0; bb < 8; ++bb) {
            const int b = 8 * p + bb;  // rows 32 p .. 32 p + 31: this thread's q + 4 b
            if (b < RB) {
                const int j = q + 4 * b;
                prod[q + 4 * bb][el] = j <= a.k ? r.v[b] * wi : (j < R ? wi * wi : 0.0);
            }
        }
        __syncthreads();
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m) t += prod[row][s + 32 * m];
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) t += __shfl_xor(t, off);  // within the 32 lanes of the row
        const int j = 32 * p + row;
        if (s == 0 && j < R) part[(long long)blockIdx.x * R + j] = t;
    }
}

// x - sum_j v_j[i] c[j] for the thread's element, every slice getting it: x is
// slice 0's value (the only slice that reads or writes w), the rows' products
// summed per slice and the four slices' sums added in order
template <int RB>
__device__ __forceinline__ double reg_combine(const ArnoldiArgs& a, const RegRows<RB>& r, const double* c, int q,
                                              int el, double x) {
    __shared__ double red[4][AR_REG_EL];
    __shared__ double xs[AR_REG_EL];
    double p = 0.0;
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        const int j = q + 4 * b;
        if (j <= a.k) p = __builtin_fma(r.v[b], c[j], p);
    }
    red[q][el] = p;
    if (q == 0) xs[el] = x;
    __syncthreads();
    return xs[el] - (((red[0][el] + red[1][el]) + red[2][el]) + red[3][el]);
}

template <int RB>
__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_reg_first(ArnoldiArgs a) {
    const int el = threadIdx.x % AR_REG_EL, q = threadIdx.x / AR_REG_EL;
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0), i = i0 + el;
    const bool ok = i < i1;
    RegRows<RB> r;
    reg_load<RB>(a, i, ok, q, r);
    const double wi = ok ? a.s0 * a.V[(long long)a.k * a.n + i] + a.s1 * a.u[i] : 0.0;
    if (ok && q == 0) a.w[i] = wi;
    reg_dots<RB>(a, r, wi, q, el, a.k + 1, a.part1);
}

template <int RB>
__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_reg_second(ArnoldiArgs a) {
    __shared__ double sh[AR_ROWS];
    const int el = threadIdx.x % AR_REG_EL, q = threadIdx.x / AR_REG_EL;
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0), i = i0 + el;
    const bool ok = i < i1;
    RegRows<RB> r;
    reg_load<RB>(a, i, ok, q, r);
    const double wi = ok && q == 0 ? a.w[i] : 0.0;
    sum_partials(a, a.part1, a.k + 1, sh);  // h1
    if (blockIdx.x == 0)  // the third launch adds h2
        for (int j = threadIdx.x; j <= a.k; j += AR_THREADS) a.h[j] = sh[j];
    const double w1 = reg_combine<RB>(a, r, sh, q, el, wi);
    if (ok && q == 0) a.w[i] = w1;
    reg_dots<RB>(a, r, w1, q, el, a.k + 2, a.part2);
}

template <int RB>
__global__ void __launch_bounds__(AR_THREADS) k_arnoldi_reg_third(ArnoldiArgs a) {
    __shared__ double sh[AR_ROWS];
    __shared__ double nrm;
    const int el = threadIdx.x % AR_REG_EL, q = threadIdx.x / AR_REG_EL;
    const long long i0 = (long long)blockIdx.x * a.chunk, i1 = chunk_end(a, i0), i = i0 + el;
    const bool ok = i < i1;
    RegRows<RB> r;
    reg_load<RB>(a, i, ok, q, r);
    const double w1 = ok && q == 0 ? a.w[i] : 0.0;
    sum_partials(a, a.part2, a.k + 2, sh);  // h2 in sh[0..k], |w1|^2 in sh[k + 1]
    if (threadIdx.x < 64) {
        double t = 0.0;
        for (int j = threadIdx.x; j <= a.k; j += 64) t = __builtin_fma(sh[j], sh[j], t);
        t = wave_sum(t);
        if (threadIdx.x == 0) {
            const double d = sh[a.k + 1] - t;
            nrm = d > 0.0 ? sqrt(d) : 0.0;
        }
    }
    const double w2 = reg_combine<RB>(a, r, sh, q, el, w1);  // its barrier also publishes nrm
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    if (ok && q == 0) a.V[(long long)(a.k + 1) * a.n + i] = w2 * inv;
    if (blockIdx.x == 0) {  // h = h1 (from the second launch) + h2
        for (int j = threadIdx.x; j <= a.k; j += AR_THREADS) a.h[j] += sh[j];
        if (threadIdx.x == 0) a.h[a.k + 1] = nrm;
    }
}

template <int RB>
void launch_reg(const ArnoldiArgs& a, hipStream_t s) {
    const dim3 grid(unsigned(a.G)), block(AR_THREADS);
    hipLaunchKernelGGL(k_arnoldi_reg_first<RB>, grid, block, 0, s, a);
    hipLaunchKernelGGL(k_arnoldi_reg_second<RB>, grid, block, 0, s, a);
    hipLaunchKernelGGL(k_arnoldi_reg_third<RB>, grid, block, 0, s, a);
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
    a.G = int(std::min<long long>(AR_MAX_G, (n + AR_MIN_CHUNK - 1) / AR_MIN_CHUNK));
    a.chunk = (n + a.G - 1) / a.G;
    a.G = int((n + a.chunk - 1) / a.chunk);  // every workgroup holds at least one element
    a.u = u;
    a.V = V;
    a.w = work;
    a.part1 = work + n;  // AR_MAX_G (k + 1) <= AR_MAX_G (KMX_ARNOLDI_MAX_K + 1) values
    a.part2 = a.part1 + (long long)AR_MAX_G * (KMX_ARNOLDI_MAX_K + 1);  // then AR_MAX_G AR_ROWS
    a.h = h;
    a.s0 = s0;
    a.s1 = s1;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (a.chunk <= AR_REG_EL && k + 2 <= 128) {  // rows held in registers: 4 RB >= k + 2
        if (k + 2 <= 8) launch_reg<2>(a, s);
        else if (k + 2 <= 16) launch_reg<4>(a, s);
        else if (k + 2 <= 32) launch_reg<8>(a, s);
        else if (k + 2 <= 64) launch_reg<16>(a, s);
        else launch_reg<32>(a, s);
    } else {
        const dim3 grid(unsigned(a.G)), block(AR_THREADS);
        hipLaunchKernelGGL(k_arnoldi_first, grid, block, 0, s, a);
        hipLaunchKernelGGL(k_arnoldi_second, grid, block, 0, s, a);
        hipLaunchKernelGGL(k_arnoldi_third, grid, block, 0, s, a);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("kmx_arnoldi_step: ") + hipGetErrorString(e));
        return KMX_E_HIP;
    }
    return KMX_OK;
}
