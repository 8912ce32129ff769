// kmx_blocks.hip -- the level-block (sum-factorized) operator apply.
//
// apply_A.m:100-237 sums, over connected element pairs (i, j) and terms t,
// (A_{t,1}[I_1,J_1] x ... x A_{t,d}[I_d,J_d]) x_j into y_i.  Group the
// elements by their level multi-index (l_1..l_d): the elements of one group
// are cells of a d-dimensional tensor (2^max(l_k-1,0) cells along dim k, the
// 1D indices of a level being contiguous, lev_cell_to_1D_index.m:6).  When
// the pairs between a row group L and a column group L' are exactly the cells
// whose per-dimension 1D indices are connected (a tensor product of 1D masks,
// which is how connect_nD.m:63-106 builds them from connect_1D and what an
// all-pairs list gives trivially), the sum over those pairs is
//     Y_L += sum_t  X_{L'} x_1 Ahat_{t,1} x_2 ... x_d Ahat_{t,d}
// with Ahat the masked 1D sub-matrices (level rows of L, level columns of L'):
// d mode products on dense tensors instead of one Kronecker chain per pair.
// The terms summed are the same as pair by pair; only the order of the
// floating-point additions changes.
//
// Per plan the host verifies that structure (kmx_blocks_build), and per
// precision it splits the block pairs between this kernel and the
// pair-by-pair kernels by comparing their flop counts (kmx_blocks_split).
// An apply then runs: the pair kernel over the remaining pairs (writes y),
// k_blocks over tasks (runs of block pairs sharing a row block, one
// workgroup each, result into a per-task partial), and k_blocks_reduce
// (adds the partials of each row block into y in task order: deterministic).
#include "kmx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <type_traits>
#include <unordered_map>

namespace kmx {

namespace {

template <int B, int E>
struct IPow {
    static constexpr int v = B * IPow<B, E - 1>::v;
};
template <int B>
struct IPow<B, 0> {
    static constexpr int v = 1;
};

constexpr int kCap = 2048;  // largest tensor (X, intermediate, Y) of a factorized block pair, in values
constexpr int kNT = 256;

template <typename T>
struct BArgs {
    const T* mat;
    long long off[KMX_MAX_TERMS * KMX_MAX_DIM];
    int ld[KMX_MAX_DIM];
    T scale[KMX_MAX_TERMS];
    const int* binfo;
    const int* elems;
    const int4* tasks;
    const int2* items;
    const double* x;
    T* part;
    int nterms;
    int maxX, maxI, maxY;
};

template <typename T>
struct RArgs {
    const int* binfo;
    const int* eblk;
    const int* ecell;
    const int* btask;
    const int4* tasks;
    const T* part;
    double* y;
    int N;
    int accumulate;
};

__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

// 16x16x4 matrix-core product per wave.  Operands (both precisions): lane l
// feeds A[l % 16][l / 16] and B[l / 16][l % 16].  Result register r of lane l
// holds D[row(l, r)][l % 16]; the row map differs between the two
// instructions (checked on gfx950: FP32 4*(l/16)+r, FP64 l/16+4r).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    using V = dvec4;
    static __device__ __forceinline__ V run(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
    using V = fvec4;
    static __device__ __forceinline__ V run(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// One mode product on LDS tensors, as a GEMM over fibres f = (p, q):
//     out[p, i, q] (=|+= scale *) sum_j A(i, j) in[p, j, q]
// in of shape (pre, K, post), out (pre, M, post), A column-major with leading
// dimension lda (global memory).  A is staged through LDS (As, kAC values)
// in chunks of KC columns, the next chunk loading into registers while the
// matrix cores work on the current one.  Each wave owns up to 4 16x16 tiles
// of (i, f) per tile group; rows, fibres and columns beyond the stage's
// sizes enter as zeros and are not stored.
template <typename T>
constexpr int a_chunk() {
    return sizeof(T) == 8 ? 2048 : 4096;
}

template <typename T>
__device__ __forceinline__ void mode_mfma(const T* __restrict__ in, T* __restrict__ out, int pre, int K, int post,
                                          int M, const T* __restrict__ A, int lda, bool last, T scale,
                                          T* __restrict__ As) {
    constexpr int AC = a_chunk<T>();
    constexpr int PT = AC / kNT;  // staged values per thread
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int F = pre * post;
    const int Mp = (M + 15) & ~15;
    const int tm = Mp >> 4, tf = (F + 15) >> 4;
    const int li = lane & 15, lk = lane >> 4;
    int KC = (AC / Mp) & ~3;
    const int K4 = (K + 3) & ~3;
    if (KC > K4) KC = K4;
    const int nch = (K + KC - 1) / KC;
    const int ntile = tm * tf;
    for (int g0 = 0; g0 < ntile; g0 += 16) {
        typename Mfma<T>::V acc[4];
        int pb[4], qb[4], ib[4];
        bool on[4], bok[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc[u] = typename Mfma<T>::V{T(0), T(0), T(0), T(0)};
            const int tile = g0 + wave + 4 * u;
            on[u] = tile < ntile;
            const int tt = on[u] ? tile : 0;
            ib[u] = (tt % tm) * 16;
            const int fb = (tt / tm) * 16 + li;
            bok[u] = on[u] && fb < F;
            pb[u] = bok[u] ? fb / post : 0;
            qb[u] = bok[u] ? fb - pb[u] * post : 0;
        }
        T reg[PT];
        auto fetch = [&](int kc0) {
#pragma unroll
            for (int m = 0; m < PT; ++m) {
                const int v = threadIdx.x + m * kNT;
                const int jj = v / Mp;
                const int i = v - jj * Mp;
                const int j = kc0 + jj;
                const bool ok = jj < KC && i < M && j < K;
                const T val = A[(ok ? i : 0) + (long long)(ok ? j : 0) * lda];
                reg[m] = ok ? val : T(0);
            }
        };
        fetch(0);
        for (int c = 0; c < nch; ++c) {
            const int kc0 = c * KC;
            __syncthreads();  // the previous chunk's readers are done
#pragma unroll
            for (int m = 0; m < PT; ++m) {
                const int v = threadIdx.x + m * kNT;
                if (v < Mp * KC) As[v] = reg[m];
            }
            __syncthreads();
            if (c + 1 < nch) fetch(kc0 + KC);
            const int kn = min(KC, K4 - kc0);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (!on[u]) continue;  // wave-uniform
                const T* bp = in + pb[u] * K * post + qb[u];
                const T* ap = As + ib[u] + li;
                for (int s4 = 0; s4 < kn; s4 += 4) {
                    const int jj = s4 + lk;
                    const int j = kc0 + jj;
                    const T av = ap[jj * Mp];
                    const T bv = bp[(j < K ? j : 0) * post];
                    acc[u] = Mfma<T>::run(av, (bok[u] && j < K) ? bv : T(0), acc[u]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (!bok[u]) continue;
            T* op = out + pb[u] * M * post + qb[u];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = ib[u] + Mfma<T>::row(lane, r);
                if (i < M) {
                    if (last)
                        op[i * post] += scale * acc[u][r];
                    else
                        op[i * post] = acc[u][r];
                }
            }
        }
    }
}

// BF16 matrix-core product, FP32 accumulation: 16x16x16 per wave.  Lane l
// feeds A[l % 16][4 (l / 16) + e] and B[4 (l / 16) + e][l % 16], e = 0..3;
// result register r of lane l holds D[4 (l / 16) + r][l % 16] (checked on
// gfx950).  Operands are rounded to nearest-even from FP32.
typedef short svec4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) {
    unsigned u = __builtin_bit_cast(unsigned, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return short((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return short(u >> 16);
}

__device__ __forceinline__ void mode_mfma_bf16(const float* __restrict__ in, float* __restrict__ out, int pre, int K,
                                               int post, int M, const float* __restrict__ A, int lda, bool last,
                                               float scale, float* __restrict__ As) {
    constexpr int AC = a_chunk<float>();
    constexpr int PT = AC / kNT;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int F = pre * post;
    const int Mp = (M + 15) & ~15;
    const int tm = Mp >> 4, tf = (F + 15) >> 4;
    const int li = lane & 15, lk = lane >> 4;
    int KC = (AC / Mp) & ~15;
    const int K16 = (K + 15) & ~15;
    if (KC > K16) KC = K16;
    const int nch = (K + KC - 1) / KC;
    const int ntile = tm * tf;
    for (int g0 = 0; g0 < ntile; g0 += 16) {
        fvec4 acc[4];
        int pb[4], qb[4], ib[4];
        bool on[4], bok[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc[u] = fvec4{0.f, 0.f, 0.f, 0.f};
            const int tile = g0 + wave + 4 * u;
            on[u] = tile < ntile;
            const int tt = on[u] ? tile : 0;
            ib[u] = (tt % tm) * 16;
            const int fb = (tt / tm) * 16 + li;
            bok[u] = on[u] && fb < F;
            pb[u] = bok[u] ? fb / post : 0;
            qb[u] = bok[u] ? fb - pb[u] * post : 0;
        }
        float reg[PT];
        auto fetch = [&](int kc0) {
#pragma unroll
            for (int m = 0; m < PT; ++m) {
                const int v = threadIdx.x + m * kNT;
                const int jj = v / Mp;
                const int i = v - jj * Mp;
                const int j = kc0 + jj;
                const bool ok = jj < KC && i < M && j < K;
                const float val = A[(ok ? i : 0) + (long long)(ok ? j : 0) * lda];
                reg[m] = ok ? val : 0.f;
            }
        };
        fetch(0);
        for (int c = 0; c < nch; ++c) {
            const int kc0 = c * KC;
            __syncthreads();
#pragma unroll
            for (int m = 0; m < PT; ++m) {
                const int v = threadIdx.x + m * kNT;
                if (v < Mp * KC) As[v] = reg[m];
            }
            __syncthreads();
            if (c + 1 < nch) fetch(kc0 + KC);
            const int kn = min(KC, K16 - kc0);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (!on[u]) continue;  // wave-uniform
                const float* bp = in + pb[u] * K * post + qb[u];
                const float* ap = As + ib[u] + li;
                for (int s16 = 0; s16 < kn; s16 += 16) {
                    svec4 av, bv;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int jj = s16 + 4 * lk + e;
                        const int j = kc0 + jj;
                        av[e] = to_bf16(ap[jj * Mp]);
                        const float b = bp[(j < K ? j : 0) * post];
                        bv[e] = to_bf16((bok[u] && j < K) ? b : 0.f);
                    }
                    acc[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[u], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (!bok[u]) continue;
            float* op = out + pb[u] * M * post + qb[u];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = ib[u] + 4 * (lane >> 4) + r;
                if (i < M) {
                    if (last)
                        op[i * post] += scale * acc[u][r];
                    else
                        op[i * post] = acc[u][r];
                }
            }
        }
    }
}

// Stages with at most 8 output rows: a thread per fibre holds all M = R
// outputs, A(:, j) is the same for every thread of the workgroup.
template <typename T, int R>
__device__ __forceinline__ void mode_small(const T* __restrict__ in, T* __restrict__ out, int pre, int K, int post,
                                           const T* __restrict__ A, int lda, bool last, T scale) {
    const int F = pre * post;
    for (int f = threadIdx.x; f < F; f += kNT) {
        const int p = f / post;
        const int q = f - p * post;
        const T* ip = in + p * K * post + q;
        T s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = T(0);
#pragma unroll 3
        for (int j = 0; j < K; ++j) {
            const T v = ip[j * post];
            const T* aj = A + (long long)j * lda;
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] = fmadd(aj[r], v, s[r]);
        }
        T* op = out + p * R * post + q;
        if (last) {
#pragma unroll
            for (int r = 0; r < R; ++r) op[r * post] += scale * s[r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) op[r * post] = s[r];
        }
    }
}

template <typename T, int DEG>
__device__ __forceinline__ void mode_product(const T* in, T* out, int pre, int K, int post, int M, const T* A,
                                             int lda, bool last, T scale, T* As) {
    if (M == DEG) {
        mode_small<T, DEG>(in, out, pre, K, post, A, lda, last, scale);
    } else if constexpr (2 * DEG <= 8) {
        if (M == 2 * DEG) {
            mode_small<T, 2 * DEG>(in, out, pre, K, post, A, lda, last, scale);
            return;
        }
        if constexpr (4 * DEG <= 8) {
            if (M == 4 * DEG) {
                mode_small<T, 4 * DEG>(in, out, pre, K, post, A, lda, last, scale);
                return;
            }
        }
        mode_mfma<T>(in, out, pre, K, post, M, A, lda, last, scale, As);
    } else {
        mode_mfma<T>(in, out, pre, K, post, M, A, lda, last, scale, As);
    }
}

template <typename T, int D, int DEG, bool BF16>
__global__ void __launch_bounds__(kNT) k_blocks(BArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int BE = IPow<DEG, D>::v;
    T* X = reinterpret_cast<T*>(smem);
    T* I0 = X + a.maxX;
    T* I1 = I0 + a.maxI;
    T* Y = I1 + a.maxI;
    T* As = Y + a.maxY;  // a_chunk<T>() values
    const int4 task = a.tasks[blockIdx.x];
    const int* bi = a.binfo + task.x * (2 * D + 1);
    int rd[D];
    int ny = 1;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        rd[k] = DEG << bi[k];
        ny *= rd[k];
    }
    for (int v = threadIdx.x; v < ny; v += kNT) Y[v] = T(0);
    for (int it = task.y; it < task.z; ++it) {
        const int2 item = a.items[it];
        const int* bj = a.binfo + item.x * (2 * D + 1);
        int sd[D], st[D];
        int lcs = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            sd[k] = DEG << bj[k];
            lcs += bj[k];
        }
        st[D - 1] = 1;
#pragma unroll
        for (int k = D - 2; k >= 0; --k) st[k] = st[k + 1] * sd[k + 1];
        // gather x of the column block into its dense cell tensor
        const int* el = a.elems + bj[2 * D];
        const int nx = BE << lcs;
        __syncthreads();
        for (int v = threadIdx.x; v < nx; v += kNT) {
            int ci = v / BE;
            int ai = v - ci * BE;
            const int e = el[ci];
            const int av = ai;
            int pos = 0;
#pragma unroll
            for (int k = D - 1; k >= 0; --k) {
                const int c = ci & ((1 << bj[k]) - 1);
                ci >>= bj[k];
                const int ak = ai % DEG;
                ai /= DEG;
                pos += (c * DEG + ak) * st[k];
            }
            X[pos] = e >= 0 ? T(a.x[(long long)e * BE + av]) : T(0);
        }
        __syncthreads();
        for (int t = 0; t < a.nterms; ++t) {
            int dims[D];
#pragma unroll
            for (int k = 0; k < D; ++k) dims[k] = sd[k];
            const T* src = X;
            for (int s = 0; s < D; ++s) {
                const int k = (item.y >> (3 * s)) & 7;
                int pre = 1, post = 1;
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    if (m < k) pre *= dims[m];
                    if (m > k) post *= dims[m];
                }
                const T* A = a.mat + a.off[t * D + k] + bi[D + k] * DEG + (long long)(bj[D + k] * DEG) * a.ld[k];
                const bool last = s == D - 1;
                T* dst = last ? Y : (src == I0 ? I1 : I0);
                if constexpr (BF16)
                    mode_mfma_bf16(src, dst, pre, dims[k], post, rd[k], A, a.ld[k], last, a.scale[t], As);
                else
                    mode_product<T, DEG>(src, dst, pre, dims[k], post, rd[k], A, a.ld[k], last, a.scale[t], As);
                __syncthreads();
                dims[k] = rd[k];
                src = dst;
            }
        }
    }
    T* out = a.part + task.w;
    for (int v = threadIdx.x; v < ny; v += kNT) out[v] = Y[v];
}

template <typename T, int D, int DEG>
__global__ void __launch_bounds__(256) k_blocks_reduce(RArgs<T> a) {
    constexpr int BE = IPow<DEG, D>::v;
    const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    if (v >= (long long)a.N * BE) return;
    const int e = int(v / BE);
    int ai = int(v - (long long)e * BE);
    const int b = a.eblk[e];
    const int t0 = a.btask[b], t1 = a.btask[b + 1];
    double acc = a.accumulate ? a.y[v] : 0.0;
    if (t0 < t1) {
        const int* bi = a.binfo + b * (2 * D + 1);
        int ci = a.ecell[e];
        int pos = 0, st = 1;
#pragma unroll
        for (int k = D - 1; k >= 0; --k) {
            const int c = ci & ((1 << bi[k]) - 1);
            ci >>= bi[k];
            const int ak = ai % DEG;
            ai /= DEG;
            pos += (c * DEG + ak) * st;
            st *= DEG << bi[k];
        }
        // the partials of one row block are consecutive, |Y| values apart
        const T* pp = a.part + a.tasks[t0].w + pos;
        const int ysz = st;
        for (int t = 0; t < t1 - t0; ++t) acc += double(pp[(long long)t * ysz]);
    }
    a.y[v] = acc;
}

using BlocksFn = hipError_t (*)(const void* args, int ntasks, size_t smem, hipStream_t s);
using ReduceFn = hipError_t (*)(const void* args, long long nval, hipStream_t s);

template <typename T, int D, int DEG, bool BF16 = false>
hipError_t launch_blocks(const void* args, int ntasks, size_t smem, hipStream_t s) {
    auto kern = k_blocks<T, D, DEG, BF16>;
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(ntasks), dim3(kNT), smem, s, *reinterpret_cast<const BArgs<T>*>(args));
    return hipGetLastError();
}

template <typename T, int D, int DEG>
hipError_t launch_reduce(const void* args, long long nval, hipStream_t s) {
    const unsigned grid = unsigned((nval + 255) / 256);
    hipLaunchKernelGGL((k_blocks_reduce<T, D, DEG>), dim3(grid), dim3(256), 0, s,
                       *reinterpret_cast<const RArgs<T>*>(args));
    return hipGetLastError();
}

// deg 1..4, ndim 2..6
template <typename T, bool BF16, int D>
BlocksFn pick_blocks_d(int deg) {
    switch (deg) {
        case 1: return &launch_blocks<T, D, 1, BF16>;
        case 2: return &launch_blocks<T, D, 2, BF16>;
        case 3: return &launch_blocks<T, D, 3, BF16>;
        case 4: return &launch_blocks<T, D, 4, BF16>;
    }
    return nullptr;
}
template <typename T, bool BF16 = false>
BlocksFn pick_blocks(int D, int deg) {
    switch (D) {
        case 2: return pick_blocks_d<T, BF16, 2>(deg);
        case 3: return pick_blocks_d<T, BF16, 3>(deg);
        case 4: return pick_blocks_d<T, BF16, 4>(deg);
        case 5: return pick_blocks_d<T, BF16, 5>(deg);
        case 6: return pick_blocks_d<T, BF16, 6>(deg);
    }
    return nullptr;
}
template <typename T, int D>
ReduceFn pick_reduce_d(int deg) {
    switch (deg) {
        case 1: return &launch_reduce<T, D, 1>;
        case 2: return &launch_reduce<T, D, 2>;
        case 3: return &launch_reduce<T, D, 3>;
        case 4: return &launch_reduce<T, D, 4>;
    }
    return nullptr;
}
template <typename T>
ReduceFn pick_reduce(int D, int deg) {
    switch (D) {
        case 2: return pick_reduce_d<T, 2>(deg);
        case 3: return pick_reduce_d<T, 3>(deg);
        case 4: return pick_reduce_d<T, 4>(deg);
        case 5: return pick_reduce_d<T, 5>(deg);
        case 6: return pick_reduce_d<T, 6>(deg);
    }
    return nullptr;
}

// ------------------------------------------------------------- host helpers
struct HipErr : std::runtime_error {
    using std::runtime_error::runtime_error;
};
void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipErr(std::string(what) + ": " + hipGetErrorString(e));
}
template <typename T>
T* up(const T* h, size_t n) {
    if (n == 0) return nullptr;
    T* d = nullptr;
    if (hipMalloc(&d, sizeof(T) * n) != hipSuccess) throw std::runtime_error("device allocation failed");
    hcheck(hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}
template <typename T>
void dfree(T*& p) {
    if (p) hipFree(p);
    p = nullptr;
}
void free_split(kmx_split& s) {
    dfree(s.d_rp);
    dfree(s.d_cols);
    dfree(s.d_order);
    dfree(s.d_tasks);
    dfree(s.d_items);
    dfree(s.d_btask);
    dfree(s.d_part);
    s = kmx_split();
}

int level_of(int I0) {
    int l = 0;
    while ((1 << l) <= I0) ++l;  // I0 in [2^(l-1), 2^l) for l >= 1; I0 = 0 -> 0
    return l;
}

// Masked column-major copy of one 1D matrix: entries of unconnected cell pairs zeroed.
void masked_matrix(const double* M, int n, int deg, const std::vector<unsigned char>& mask, double* out) {
    const int nb = n / deg;
    for (int J = 0; J < nb; ++J)
        for (int b = 0; b < deg; ++b) {
            const size_t c = size_t(J) * deg + b;
            for (int I = 0; I < nb; ++I) {
                const bool on = mask[size_t(I) * nb + J] != 0;
                for (int a = 0; a < deg; ++a) {
                    const size_t r = size_t(I) * deg + a;
                    out[c * n + r] = on ? M[c * n + r] : 0.0;
                }
            }
        }
}

// Cheapest contraction order for one block pair (flops of one term), with all
// intermediates within kCap values.  Exhaustive for d <= 4, greedy above.
void best_order(int D, int deg, const int* r, const int* s, double& flops, int& order, int& maxI, bool& fits) {
    int perm[KMX_MAX_DIM];
    std::iota(perm, perm + D, 0);
    flops = 1e300;
    order = 0;
    maxI = 0;
    fits = false;
    auto eval = [&](const int* pm, double& f, int& mi) {
        long long cur[KMX_MAX_DIM];
        for (int k = 0; k < D; ++k) cur[k] = (long long)s[k] * deg;
        f = 0;
        mi = 0;
        bool ok = true;
        for (int st = 0; st < D; ++st) {
            const int k = pm[st];
            long long tot = 1;
            for (int m = 0; m < D; ++m) tot *= cur[m];
            const long long rest = tot / cur[k];
            f += 2.0 * double(rest) * double(r[k]) * deg * double(s[k]) * deg;
            cur[k] = (long long)r[k] * deg;
            if (st < D - 1) {
                const long long sz = rest * cur[k];
                if (sz > kCap) ok = false;
                mi = int(std::max<long long>(mi, std::min<long long>(sz, 1 << 30)));
            }
        }
        return ok;
    };
    if (D <= 4) {
        do {
            double f;
            int mi;
            const bool ok = eval(perm, f, mi);
            if (ok && f < flops) {
                flops = f;
                maxI = mi;
                fits = true;
                order = 0;
                for (int st = 0; st < D; ++st) order |= perm[st] << (3 * st);
            }
        } while (std::next_permutation(perm, perm + D));
        if (!fits) {  // report the flop count of the natural order anyway
            std::iota(perm, perm + D, 0);
            int mi;
            eval(perm, flops, mi);
        }
        return;
    }
    // greedy: at each stage contract the dim whose product is cheapest now
    bool used[KMX_MAX_DIM] = {false};
    long long cur[KMX_MAX_DIM];
    for (int k = 0; k < D; ++k) cur[k] = (long long)s[k] * deg;
    int pm[KMX_MAX_DIM];
    for (int st = 0; st < D; ++st) {
        int bk = -1;
        double bc = 0;
        for (int k = 0; k < D; ++k) {
            if (used[k]) continue;
            long long tot = 1;
            for (int m = 0; m < D; ++m) tot *= cur[m];
            const double c = 2.0 * double(tot / cur[k]) * double(r[k]) * deg * double(s[k]) * deg;
            if (bk < 0 || c < bc) {
                bk = k;
                bc = c;
            }
        }
        used[bk] = true;
        pm[st] = bk;
        cur[bk] = (long long)r[bk] * deg;
    }
    double f;
    int mi;
    fits = eval(pm, f, mi);
    flops = f;
    maxI = mi;
    order = 0;
    for (int st = 0; st < D; ++st) order |= pm[st] << (3 * st);
}

int num_cus() {
    static int n = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) return 256;
        return v;
    }();
    return n;
}

}  // namespace
}  // namespace kmx

using namespace kmx;

// --------------------------------------------------------------- structure
static void build_impl(kmx_plan* p, const std::vector<int>& idx, std::vector<long long>&& rp, std::vector<int>&& cols,
                       const double* const* mats) {
    const int D = p->D, deg = p->deg, N = p->N, T = p->T;
    kmx_blockset* bs = p->bs;
    auto fail = [&](const char* why) { bs->ok = false; bs->why = why; };
    if (D < 2 || deg > KMX_MAX_DEG || N == 0) return fail("ndim < 2 or empty plan");
    int Lmax[KMX_MAX_DIM];
    for (int k = 0; k < D; ++k) {
        const int nb = p->nb[k];
        if (nb & (nb - 1)) return fail("1D matrix size is not deg * 2^L");
        Lmax[k] = level_of(nb - 1);
    }
    // blocks by level multi-index
    std::unordered_map<long long, int> bid;
    std::vector<int> lev(size_t(N) * D), cell(size_t(N) * D);
    bs->eblk.resize(N);
    bs->ecell.resize(N);
    std::vector<std::vector<int>> blv;  // levels of each block
    for (int e = 0; e < N; ++e) {
        long long key = 0;
        for (int k = 0; k < D; ++k) {
            const int I0 = idx[size_t(e) * D + k];
            const int l = level_of(I0);
            lev[size_t(e) * D + k] = l;
            cell[size_t(e) * D + k] = l == 0 ? 0 : I0 - (1 << (l - 1));
            key = key * 64 + l;
        }
        auto it = bid.find(key);
        int b;
        if (it == bid.end()) {
            b = int(blv.size());
            bid.emplace(key, b);
            blv.emplace_back(lev.begin() + size_t(e) * D, lev.begin() + size_t(e + 1) * D);
            bs->blev.insert(bs->blev.end(), lev.begin() + size_t(e) * D, lev.begin() + size_t(e + 1) * D);
        } else {
            b = it->second;
        }
        bs->eblk[e] = b;
    }
    const int nblk = int(blv.size());
    bs->nblk = nblk;
    if ((long long)nblk * nblk > (1LL << 24)) return fail("too many level blocks");
    bs->binfo.assign(size_t(nblk) * (2 * D + 1), 0);
    std::vector<int> ncell(nblk);
    int eoff = 0;
    for (int b = 0; b < nblk; ++b) {
        int* bi = &bs->binfo[size_t(b) * (2 * D + 1)];
        int lcs = 0;
        for (int k = 0; k < D; ++k) {
            const int l = blv[b][k];
            bi[k] = l == 0 ? 0 : l - 1;
            bi[D + k] = l == 0 ? 0 : 1 << (l - 1);
            lcs += bi[k];
        }
        if (lcs > 24) return fail("level block too large");
        bi[2 * D] = eoff;
        ncell[b] = 1 << lcs;
        eoff += ncell[b];
    }
    bs->elems.assign(size_t(eoff), -1);
    for (int e = 0; e < N; ++e) {
        const int b = bs->eblk[e];
        const int* bi = &bs->binfo[size_t(b) * (2 * D + 1)];
        int ci = 0;
        for (int k = 0; k < D; ++k) ci = (ci << bi[k]) | cell[size_t(e) * D + k];
        bs->ecell[e] = ci;
        int& slot = bs->elems[size_t(bi[2 * D]) + ci];
        if (slot >= 0) return fail("two elements share one cell");
        slot = e;
    }
    std::vector<int> present(nblk, 0);
    for (int e = 0; e < N; ++e) present[bs->eblk[e]]++;

    // per-dimension connection masks and pair counts per block pair
    bs->mask.assign(D, std::vector<unsigned char>());
    for (int k = 0; k < D; ++k) bs->mask[k].assign(size_t(p->nb[k]) * p->nb[k], 0);
    std::vector<long long> cnt(size_t(nblk) * nblk, 0);
    const bool allp = p->all_pairs;
    if (allp) {
        for (int k = 0; k < D; ++k) {
            std::vector<unsigned char> has(p->nb[k], 0);
            for (int e = 0; e < N; ++e) has[idx[size_t(e) * D + k]] = 1;
            const int nb = p->nb[k];
            for (int I = 0; I < nb; ++I)
                if (has[I])
                    for (int J = 0; J < nb; ++J)
                        if (has[J]) bs->mask[k][size_t(I) * nb + J] = 1;
        }
        for (int a = 0; a < nblk; ++a)
            for (int b = 0; b < nblk; ++b) cnt[size_t(a) * nblk + b] = (long long)present[a] * present[b];
    } else {
        for (int i = 0; i < N; ++i) {
            const int bi = bs->eblk[i];
            for (long long q = rp[i]; q < rp[i + 1]; ++q) {
                const int j = cols[q];
                cnt[size_t(bi) * nblk + bs->eblk[j]]++;
                for (int k = 0; k < D; ++k)
                    bs->mask[k][size_t(idx[size_t(i) * D + k]) * p->nb[k] + idx[size_t(j) * D + k]] = 1;
            }
        }
    }
    // verify: pairs of each connected block pair == tensor product of the masks
    std::vector<std::vector<int>> members(nblk);
    for (int e = 0; e < N; ++e) members[bs->eblk[e]].push_back(e);
    double brute = 0;
    bs->bpid.assign(size_t(nblk) * nblk, -1);
    for (int a = 0; a < nblk; ++a)
        for (int b = 0; b < nblk; ++b) {
            const long long c = cnt[size_t(a) * nblk + b];
            if (c == 0) continue;
            long long expect = 0;
            if (present[a] == ncell[a] && present[b] == ncell[b]) {
                expect = 1;
                const int* ba = &bs->binfo[size_t(a) * (2 * D + 1)];
                const int* bb = &bs->binfo[size_t(b) * (2 * D + 1)];
                for (int k = 0; k < D; ++k) {
                    long long nz = 0;
                    const int nb = p->nb[k];
                    for (int I = 0; I < (1 << ba[k]); ++I)
                        for (int J = 0; J < (1 << bb[k]); ++J)
                            nz += bs->mask[k][size_t(ba[D + k] + I) * nb + (bb[D + k] + J)];
                    expect *= nz;
                }
            } else {
                brute += double(present[a]) * present[b];
                if (brute > 4e8) return fail("incomplete level blocks too large to verify");
                for (int i : members[a])
                    for (int j : members[b]) {
                        bool on = true;
                        for (int k = 0; k < D && on; ++k)
                            on = bs->mask[k][size_t(idx[size_t(i) * D + k]) * p->nb[k] + idx[size_t(j) * D + k]] != 0;
                        expect += on;
                    }
            }
            if (expect != c) return fail("pair list is not a per-block tensor product of 1D connections");
            kmx_blockset::BP bp;
            bp.bi = a;
            bp.bj = b;
            bp.pairs = c;
            int r[KMX_MAX_DIM], s[KMX_MAX_DIM];
            const int* ba = &bs->binfo[size_t(a) * (2 * D + 1)];
            const int* bb = &bs->binfo[size_t(b) * (2 * D + 1)];
            for (int k = 0; k < D; ++k) {
                r[k] = 1 << ba[k];
                s[k] = 1 << bb[k];
            }
            best_order(D, deg, r, s, bp.fflops, bp.order, bp.maxI, bp.fits);
            bp.fits = bp.fits && (long long)ncell[a] * p->B <= kCap && (long long)ncell[b] * p->B <= kCap;
            for (int k = 0; k < D; ++k)  // staged coefficient chunks hold at least 4 columns
                bp.fits = bp.fits && ((r[k] * deg + 15) & ~15) * 4 <= a_chunk<double>();
            bs->bpid[size_t(a) * nblk + b] = int(bs->bps.size());
            bs->bps.push_back(bp);
        }
    // unconnected block pairs: is their mask product empty as well?
    bs->full_product = true;
    for (int a = 0; a < nblk && bs->full_product; ++a)
        for (int b = 0; b < nblk && bs->full_product; ++b) {
            if (cnt[size_t(a) * nblk + b] != 0) continue;
            if (present[a] == ncell[a] && present[b] == ncell[b]) {
                const int* ba = &bs->binfo[size_t(a) * (2 * D + 1)];
                const int* bb = &bs->binfo[size_t(b) * (2 * D + 1)];
                bool empty = false;
                for (int k = 0; k < D && !empty; ++k) {
                    bool any = false;
                    const int nb = p->nb[k];
                    for (int I = 0; I < (1 << ba[k]) && !any; ++I)
                        for (int J = 0; J < (1 << bb[k]) && !any; ++J)
                            any = bs->mask[k][size_t(ba[D + k] + I) * nb + (bb[D + k] + J)] != 0;
                    empty = !any;
                }
                if (!empty) bs->full_product = false;
            } else {
                brute += double(present[a]) * present[b];
                if (brute > 4e8) {
                    bs->full_product = false;
                    break;
                }
                for (int i : members[a]) {
                    for (int j : members[b]) {
                        bool on = true;
                        for (int k = 0; k < D && on; ++k)
                            on = bs->mask[k][size_t(idx[size_t(i) * D + k]) * p->nb[k] + idx[size_t(j) * D + k]] != 0;
                        if (on) bs->full_product = false;
                    }
                    if (!bs->full_product) break;
                }
            }
        }
    (void)Lmax;
    // masked matrices
    std::vector<double> m64(size_t(p->mat_elems));
    for (int t = 0; t < T; ++t)
        for (int k = 0; k < D; ++k) masked_matrix(mats[t * D + k], p->n[k], deg, bs->mask[k], m64.data() + p->off[t * D + k]);
    bs->d_mat64 = up(m64.data(), m64.size());
    if (p->flags & KMX_PLAN_FP32) {
        std::vector<float> m32(m64.begin(), m64.end());
        bs->d_mat32 = up(m32.data(), m32.size());
    }
    bs->d_binfo = up(bs->binfo.data(), bs->binfo.size());
    bs->d_elems = up(bs->elems.data(), bs->elems.size());
    bs->d_eblk = up(bs->eblk.data(), bs->eblk.size());
    bs->d_ecell = up(bs->ecell.data(), bs->ecell.size());
    bs->rp = std::move(rp);
    bs->cols = std::move(cols);
    bs->ok = true;
}

// Detect the level-block structure; any failure (structure or resources)
// leaves the plan on the pair-by-pair path with the reason in bs->why.
void kmx_blocks_build(kmx_plan* p, const std::vector<int>& idx, std::vector<long long>&& rp, std::vector<int>&& cols,
                      const double* const* mats) {
    p->bs = new kmx_blockset();
    try {
        build_impl(p, idx, std::move(rp), std::move(cols), mats);
        if (p->bs->ok) kmx_sweep_build(p);
    } catch (const std::exception& e) {
        kmx_blocks_free(p);
        p->bs = new kmx_blockset();
        p->bs->why = e.what();
    }
    if (!p->bs->ok) {  // keep only the reason
        std::string why = p->bs->why;
        kmx_blocks_free(p);
        p->bs = new kmx_blockset();
        p->bs->why = why;
    }
}

double kmx_blocks_default_beta() {
    static double b = [] {
        const char* e = getenv("KMX_BLOCK_BETA");
        return e ? atof(e) : HUGE_VAL;  // off until set or tuned
    }();
    return b;
}

// Split the block pairs of one precision: factorized when its flops plus
// beta per stage undercut the pair-by-pair flops (beta < 0: every block pair
// that fits; beta = inf: none).
int kmx_blocks_split(kmx_plan* p, int slot, double beta) {
    kmx_blockset* bs = p->bs;
    if (!bs || !bs->ok) return KMX_OK;
    kmx_split& sp = bs->split[slot];
    try {
        free_split(sp);
        sp.beta = beta;
        if (std::isinf(beta) && beta > 0) return KMX_OK;
        if (p->deg > 4) return KMX_OK;  // the mode-product kernels take deg 1..4; the structure serves the sweep
        const int D = p->D, deg = p->deg, N = p->N, T = p->T, nblk = bs->nblk;
        double chain = 2.0 * D;
        for (int k = 0; k <= D; ++k) chain *= deg;  // 2 d deg^(d+1) per pair and term
        std::vector<char> fact(bs->bps.size(), 0);
        long long nf = 0, pf = 0;
        double ff = 0;
        for (size_t i = 0; i < bs->bps.size(); ++i) {
            const auto& bp = bs->bps[i];
            if (!bp.fits) continue;
            if (slot == 2) {  // BF16 chunks hold at least 16 columns of the staged coefficients
                const int* ba = &bs->binfo[size_t(bp.bi) * (2 * D + 1)];
                bool ok = true;
                for (int k = 0; k < D; ++k) ok = ok && (((deg << ba[k]) + 15) & ~15) * 16 <= a_chunk<float>();
                if (!ok) continue;
            }
            if (beta < 0 || bp.fflops + beta * D < double(bp.pairs) * chain) {
                fact[i] = 1;
                ++nf;
                pf += bp.pairs;
                ff += T * bp.fflops;
            }
        }
        if (nf == 0) return KMX_OK;
        // items sorted by row block, heaviest first; tasks = runs of one row block
        std::vector<int> it;
        for (size_t i = 0; i < fact.size(); ++i)
            if (fact[i]) it.push_back(int(i));
        auto cost = [&](int i) { return T * (bs->bps[i].fflops + 2000.0 * D); };
        std::stable_sort(it.begin(), it.end(), [&](int u, int v) {
            if (bs->bps[u].bi != bs->bps[v].bi) return bs->bps[u].bi < bs->bps[v].bi;
            return cost(u) > cost(v);
        });
        double total = 0, biggest = 0;
        for (int i : it) {
            total += cost(i);
            biggest = std::max(biggest, cost(i));
        }
        const double target = std::max(total / (4.0 * num_cus()), biggest);
        std::vector<int2> items;
        std::vector<int4> tasks;
        std::vector<int> btask(size_t(nblk) + 1, 0);
        long long part = 0;
        int maxX = 1, maxI = 1, maxY = 1;
        auto ysize = [&](int b) {
            const int* bi = &bs->binfo[size_t(b) * (2 * D + 1)];
            int lcs = 0;
            for (int k = 0; k < D; ++k) lcs += bi[k];
            return p->B << lcs;
        };
        size_t q = 0;
        for (int b = 0; b < nblk; ++b) {
            btask[b] = int(tasks.size());
            while (q < it.size() && bs->bps[it[q]].bi == b) {
                int4 tk;
                tk.x = b;
                tk.y = int(items.size());
                double c = 0;
                while (q < it.size() && bs->bps[it[q]].bi == b && (c == 0 || c + cost(it[q]) <= target)) {
                    const auto& bp = bs->bps[it[q]];
                    items.push_back(make_int2(bp.bj, bp.order));
                    maxX = std::max(maxX, ysize(bp.bj));
                    maxI = std::max(maxI, bp.maxI);
                    c += cost(it[q]);
                    ++q;
                }
                tk.z = int(items.size());
                if (part > (1LL << 31) - ysize(b)) throw std::runtime_error("partial buffer too large");
                tk.w = int(part);
                part += ysize(b);
                maxY = std::max(maxY, ysize(b));
                tasks.push_back(tk);
            }
        }
        btask[nblk] = int(tasks.size());
        // the pairs left to the pair-by-pair kernel
        std::vector<long long> srp(size_t(N) + 1, 0);
        std::vector<int> scols;
        for (int i = 0; i < N; ++i) {
            const int a = bs->eblk[i];
            if (p->all_pairs) {
                for (int j = 0; j < N; ++j)
                    if (!fact[bs->bpid[size_t(a) * nblk + bs->eblk[j]]]) scols.push_back(j);
            } else {
                for (long long r = bs->rp[i]; r < bs->rp[i + 1]; ++r) {
                    const int j = bs->cols[r];
                    if (!fact[bs->bpid[size_t(a) * nblk + bs->eblk[j]]]) scols.push_back(j);
                }
            }
            srp[i + 1] = (long long)scols.size();
        }
        sp.sub_pairs = srp[N];
        if (sp.sub_pairs > 0) {
            std::vector<int> ord(N);
            std::iota(ord.begin(), ord.end(), 0);
            std::stable_sort(ord.begin(), ord.end(),
                             [&](int u, int v) { return srp[u + 1] - srp[u] > srp[v + 1] - srp[v]; });
            sp.d_rp = up(srp.data(), srp.size());
            sp.d_cols = up(scols.data(), scols.size());
            sp.d_order = up(ord.data(), ord.size());
        }
        const size_t elem = slot == 0 ? sizeof(double) : sizeof(float);
        if (hipMalloc(&sp.d_part, size_t(part) * elem) != hipSuccess) throw std::runtime_error("device allocation failed");
        sp.part_elems = part;
        sp.d_tasks = up(tasks.data(), tasks.size());
        sp.d_items = up(items.data(), items.size());
        sp.d_btask = up(btask.data(), btask.size());
        sp.ntasks = int(tasks.size());
        sp.block_pairs = nf;
        sp.pairs_in_blocks = pf;
        sp.fact_flops = ff;
        sp.maxX = maxX;
        sp.maxI = maxI;
        sp.maxY = maxY;
        sp.active = true;
        return KMX_OK;
    } catch (const HipErr& e) {
        free_split(sp);
        set_error(std::string("level-block split: ") + e.what());
        return KMX_E_HIP;
    } catch (const std::exception& e) {
        free_split(sp);
        set_error(std::string("level-block split: ") + e.what());
        return KMX_E_NOMEM;
    }
}

int kmx_blocks_apply(kmx_plan* p, int slot, const double* x, double* y, hipStream_t s) {
    kmx_blockset* bs = p->bs;
    kmx_split& sp = bs->split[slot];
    const bool f64 = slot == 0, bf = slot == 2;
    if (bf && sp.sub_pairs > 0) {
        set_error("kmx_apply: BF16 needs every block pair of the plan on the level-block path");
        return KMX_E_UNSUPPORTED;
    }
    if (sp.sub_pairs > 0) {
        kmx_pairs pl;
        pl.rp = sp.d_rp;
        pl.cols = sp.d_cols;
        pl.order = sp.d_order;
        const int rc = kmx_launch_pairs(p, &pl, x, y, f64 ? KMX_PREC_FP64 : KMX_PREC_FP32, s);
        if (rc != KMX_OK) return rc;
    }
    const size_t elem = f64 ? sizeof(double) : sizeof(float);
    const size_t smem = size_t(sp.maxX + 2 * sp.maxI + sp.maxY + (f64 ? a_chunk<double>() : a_chunk<float>())) * elem;
    hipError_t err;
    auto fill = [&](auto& a, const auto* mat) {
        std::memset(&a, 0, sizeof(a));
        a.mat = mat;
        for (int i = 0; i < p->T * p->D; ++i) a.off[i] = p->off[i];
        for (int k = 0; k < p->D; ++k) a.ld[k] = p->n[k];
        using E = std::remove_reference_t<decltype(a.scale[0])>;
        for (int t = 0; t < p->T; ++t) a.scale[t] = E(p->scale[t]);
        a.binfo = bs->d_binfo;
        a.elems = bs->d_elems;
        a.tasks = sp.d_tasks;
        a.items = sp.d_items;
        a.x = x;
        a.nterms = p->T;
        a.maxX = sp.maxX;
        a.maxI = sp.maxI;
        a.maxY = sp.maxY;
    };
    auto fill_r = [&](auto& r, const auto* part) {
        r.binfo = bs->d_binfo;
        r.eblk = bs->d_eblk;
        r.ecell = bs->d_ecell;
        r.btask = sp.d_btask;
        r.tasks = sp.d_tasks;
        r.part = part;
        r.y = y;
        r.N = p->N;
        r.accumulate = sp.sub_pairs > 0;
    };
    const long long nval = (long long)p->N * p->B;
    if (f64) {
        if (!bs->d_mat64) return KMX_E_UNSUPPORTED;
        BArgs<double> a;
        fill(a, bs->d_mat64);
        a.part = static_cast<double*>(sp.d_part);
        err = pick_blocks<double>(p->D, p->deg)(&a, sp.ntasks, smem, s);
        if (err == hipSuccess) {
            RArgs<double> r;
            fill_r(r, static_cast<const double*>(sp.d_part));
            err = pick_reduce<double>(p->D, p->deg)(&r, nval, s);
        }
    } else {
        if (!bs->d_mat32) {
            set_error("kmx_apply: plan has no FP32 blocks (create it with KMX_PLAN_FP32)");
            return KMX_E_ARG;
        }
        BArgs<float> a;
        fill(a, bs->d_mat32);
        a.part = static_cast<float*>(sp.d_part);
        err = (bf ? pick_blocks<float, true>(p->D, p->deg) : pick_blocks<float>(p->D, p->deg))(&a, sp.ntasks, smem, s);
        if (err == hipSuccess) {
            RArgs<float> r;
            fill_r(r, static_cast<const float*>(sp.d_part));
            err = pick_reduce<float>(p->D, p->deg)(&r, nval, s);
        }
    }
    if (err != hipSuccess) {
        set_error(std::string("level-block launch: ") + hipGetErrorString(err));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

void kmx_blocks_update_term(kmx_plan* p, int term, int dim, const double* mat, hipStream_t s) {
    kmx_blockset* bs = p->bs;
    if (!bs || !bs->ok) return;
    const int n = p->n[dim];
    std::vector<double> m(size_t(n) * n);
    masked_matrix(mat, n, p->deg, bs->mask[dim], m.data());
    const long long off = p->off[term * p->D + dim];
    hcheck(hipMemcpyAsync(bs->d_mat64 + off, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice, s),
           "hipMemcpyAsync");
    std::vector<float> f;
    if (bs->d_mat32) {
        f.assign(m.begin(), m.end());
        hcheck(hipMemcpyAsync(bs->d_mat32 + off, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, s),
               "hipMemcpyAsync");
    }
    hcheck(hipStreamSynchronize(s), "hipStreamSynchronize");  // staging vectors die here
}

void kmx_blocks_free(kmx_plan* p) {
    kmx_blockset* bs = p->bs;
    if (!bs) return;
    free_split(bs->split[0]);
    free_split(bs->split[1]);
    free_split(bs->split[2]);
    dfree(bs->d_binfo);
    dfree(bs->d_elems);
    dfree(bs->d_eblk);
    dfree(bs->d_ecell);
    dfree(bs->d_mat64);
    dfree(bs->d_mat32);
    delete bs;
    p->bs = nullptr;
}
