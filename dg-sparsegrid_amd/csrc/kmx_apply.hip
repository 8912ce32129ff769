// kmx_apply.hip -- gfx950 kernels for y = A x with A = sum over element pairs
// (i, j) and terms t of kron_k A_{t,k}[I_k(i), J_k(j)]  (apply_A.m:100-237).
//
// Work decomposition (CDNA4, wave64):
//   * one workgroup per row element i (rows issued heaviest first), its NT
//     threads stride over i's pair list, one pair per lane;
//   * a lane keeps the deg^d block of x_j and its partial y_i in registers and
//     runs the d contractions of each term in place, fibre by fibre, dim d
//     first and dim 1 last (the order of kronmult{d} method 1,
//     Kronmult/kronmult3.m:73-89);
//   * the deg x deg coefficient blocks A_{t,k}[I_k(i), :] of the row are staged
//     once per workgroup into LDS when they fit (they are contiguous in the
//     plan's block-row-major layout), else read through L2;
//   * y_i is reduced across the workgroup with cross-lane shuffles and one LDS
//     pass, then written once: no atomics, deterministic.
// Blocks too large for registers go to a generic LDS kernel (one pair at a time,
// threads over block entries).
#include "kmx_internal.h"

#include <algorithm>
#include <cstring>

namespace kmx {

template <int A, int N>
struct Pow {
    static constexpr int v = A * Pow<A, N - 1>::v;
};
template <int A>
struct Pow<A, 0> {
    static constexpr int v = 1;
};

template <typename T, int N>
using RegArr = T[N];

__device__ __forceinline__ double madd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// In-place contraction of dim K (0 = slowest) of a register block with the
// deg x deg block m (row-major [a][b]): v[.., a, ..] = sum_b m[a][b] v[.., b, ..].
template <typename T, int D, int DEG, int K>
__device__ __forceinline__ void contract_dim(RegArr<T, Pow<DEG, D>::v>& v, const RegArr<T, DEG * DEG>& m) {
    constexpr int B = Pow<DEG, D>::v;
    constexpr int S = Pow<DEG, D - 1 - K>::v;
    constexpr int O = B / (S * DEG);
#pragma unroll
    for (int o = 0; o < O; ++o) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            T in[DEG];
#pragma unroll
            for (int b = 0; b < DEG; ++b) in[b] = v[o * S * DEG + b * S + s];
#pragma unroll
            for (int a = 0; a < DEG; ++a) {
                T acc = m[a * DEG] * in[0];
#pragma unroll
                for (int b = 1; b < DEG; ++b) acc = madd(m[a * DEG + b], in[b], acc);
                v[o * S * DEG + a * S + s] = acc;
            }
        }
    }
}

// Per-row context shared by the chain steps.
template <typename T, int D, int DEG, bool USE_LDS>
struct RowCtx {
    const KArgs<T>* a;
    const T* strip;           // LDS strips of this row (USE_LDS)
    int I[D];                 // row's 0-based 1D indices
    int soff[D];              // LDS offset of (t=0, k) strip
    int sterm;                // LDS elements per term

    __device__ __forceinline__ void load_block(int t, int k, int J, RegArr<T, DEG * DEG>& m) const {
        const T* p;
        if (USE_LDS) {
            p = strip + t * sterm + soff[k] + J * (DEG * DEG);
        } else {
            p = a->blk + a->off[t * D + k] + ((long long)I[k] * a->nb[k] + J) * (DEG * DEG);
        }
#pragma unroll
        for (int e = 0; e < DEG * DEG; ++e) m[e] = p[e];
    }
};

template <typename T, int D, int DEG, bool USE_LDS, int K>
struct Chain {
    __device__ __forceinline__ static void run(RegArr<T, Pow<DEG, D>::v>& v, const RowCtx<T, D, DEG, USE_LDS>& c, int t,
                                               const RegArr<int, D>& J) {
        T m[DEG * DEG];
        c.load_block(t, K, J[K], m);
        contract_dim<T, D, DEG, K>(v, m);
        Chain<T, D, DEG, USE_LDS, K - 1>::run(v, c, t, J);
    }
};
template <typename T, int D, int DEG, bool USE_LDS>
struct Chain<T, D, DEG, USE_LDS, -1> {
    __device__ __forceinline__ static void run(RegArr<T, Pow<DEG, D>::v>&, const RowCtx<T, D, DEG, USE_LDS>&, int,
                                               const RegArr<int, D>&) {}
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One row: NT lanes stride over the pair list; y_i reduced and written once.
template <typename T, int D, int DEG, int NT, bool USE_LDS, bool KEEPX>
__device__ __forceinline__ void apply_row(const KArgs<T>& a, const double* __restrict__ x, double* __restrict__ y,
                                          int row, unsigned char* smem) {
    constexpr int B = Pow<DEG, D>::v;
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x;

    RowCtx<T, D, DEG, USE_LDS> c;
    c.a = &a;
#pragma unroll
    for (int k = 0; k < D; ++k) c.I[k] = a.idx1d[(long long)row * D + k];

    if (USE_LDS) {
        T* strip = reinterpret_cast<T*>(smem);
        int so = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            c.soff[k] = so;
            so += a.nb[k] * DEG * DEG;
        }
        c.sterm = so;
        for (int t = 0; t < a.nterms; ++t) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const T* src = a.blk + a.off[t * D + k] + (long long)c.I[k] * a.nb[k] * DEG * DEG;
                T* dst = strip + t * so + c.soff[k];
                const int len = a.nb[k] * DEG * DEG;
                for (int e = tid; e < len; e += NT) dst[e] = src[e];
            }
        }
        __syncthreads();
        c.strip = strip;
    }

    long long pb, pe;
    if (a.row_ptr) {
        pb = a.row_ptr[row];
        pe = a.row_ptr[row + 1];
    } else {
        pb = 0;
        pe = a.N;
    }

    T acc[B];
#pragma unroll
    for (int q = 0; q < B; ++q) acc[q] = T(0);

    for (long long q = pb + tid; q < pe; q += NT) {
        const int j = a.row_ptr ? a.cols[q] : int(q);
        int J[D];
#pragma unroll
        for (int k = 0; k < D; ++k) J[k] = a.idx1d[(long long)j * D + k];
        const double* xj = x + (long long)j * B;
        T xv[KEEPX ? B : 1];
        if (KEEPX) {
#pragma unroll
            for (int e = 0; e < B; ++e) xv[e] = T(xj[e]);
        }
        for (int t = 0; t < a.nterms; ++t) {
            T v[B];
#pragma unroll
            for (int e = 0; e < B; ++e) v[e] = KEEPX ? xv[e] : T(xj[e]);
            Chain<T, D, DEG, USE_LDS, D - 1>::run(v, c, t, J);
            const T s = a.scale[t];
#pragma unroll
            for (int e = 0; e < B; ++e) acc[e] = madd(s, v[e], acc[e]);
        }
    }

    // workgroup reduction of acc[B]
    if (USE_LDS) __syncthreads();  // strips are dead; reuse LDS
    T* red = reinterpret_cast<T*>(smem);
    const int lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int e = 0; e < B; ++e) {
        T s = wave_sum(acc[e]);
        if (lane == 0) red[w * B + e] = s;
    }
    __syncthreads();
    for (int e = tid; e < B; e += NT) {
        T s = red[e];
#pragma unroll
        for (int ww = 1; ww < NW; ++ww) s += red[ww * B + e];
        y[(long long)row * B + e] = double(s);
    }
}

template <typename T, int D, int DEG, int NT, bool USE_LDS, bool KEEPX>
__global__ void __launch_bounds__(NT) k_apply_reg(KArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int row = a.row_order[blockIdx.x];
    apply_row<T, D, DEG, NT, USE_LDS, KEEPX>(a, a.x, a.y, row, smem);
}

// ------------------------------------------------------------ many instances
// One persistent launch over a work list of (instance, row) items, heaviest
// first, striding by the grid; each instance keeps its own plan (levels, deg,
// pair list, coefficient blocks).  The shape switch selects the fully unrolled
// row body for the instance's (d, deg).
#define KMX_BATCH_MAX 128
struct BatchIO {
    const double* x[KMX_BATCH_MAX];
    double* y[KMX_BATCH_MAX];
};

template <typename T, int D, int DEG>
__device__ __forceinline__ void batch_row(const KArgs<T>& a, const double* x, double* y, int row, unsigned char* sm) {
    if constexpr (Pow<DEG, D>::v * int(sizeof(T)) <= 256) {
        constexpr bool KX = 3 * Pow<DEG, D>::v * int(sizeof(T)) / 4 <= 144;
        apply_row<T, D, DEG, 256, false, KX>(a, x, y, row, sm);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_apply_batch(const KArgs<T>* __restrict__ args, const int2* __restrict__ items,
                                                     int nitems, const int* __restrict__ shape, BatchIO io) {
    extern __shared__ __align__(16) unsigned char smem[];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 w = items[it];
        const KArgs<T>& a = args[w.x];
        const double* x = io.x[w.x];
        double* y = io.y[w.x];
        switch (shape[w.x]) {
            case 0x11: batch_row<T, 1, 1>(a, x, y, w.y, smem); break;
            case 0x12: batch_row<T, 1, 2>(a, x, y, w.y, smem); break;
            case 0x13: batch_row<T, 1, 3>(a, x, y, w.y, smem); break;
            case 0x14: batch_row<T, 1, 4>(a, x, y, w.y, smem); break;
            case 0x21: batch_row<T, 2, 1>(a, x, y, w.y, smem); break;
            case 0x22: batch_row<T, 2, 2>(a, x, y, w.y, smem); break;
            case 0x23: batch_row<T, 2, 3>(a, x, y, w.y, smem); break;
            case 0x24: batch_row<T, 2, 4>(a, x, y, w.y, smem); break;
            case 0x31: batch_row<T, 3, 1>(a, x, y, w.y, smem); break;
            case 0x32: batch_row<T, 3, 2>(a, x, y, w.y, smem); break;
            case 0x33: batch_row<T, 3, 3>(a, x, y, w.y, smem); break;
            case 0x34: batch_row<T, 3, 4>(a, x, y, w.y, smem); break;
            default: break;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ generic
// One workgroup per row, one pair at a time, threads over the block entries;
// any (d, deg) whose three block buffers fit in LDS.
template <typename T>
__global__ void __launch_bounds__(256) k_apply_generic(KArgs<T> a, int D, int DEG) {
    extern __shared__ __align__(16) unsigned char smem[];
    int B = 1;
    for (int k = 0; k < D; ++k) B *= DEG;
    T* acc = reinterpret_cast<T*>(smem);
    T* v = acc + B;
    T* w = v + B;
    const int row = a.row_order[blockIdx.x];
    const int tid = threadIdx.x;
    int I[KMX_MAX_DIM], J[KMX_MAX_DIM];
    for (int k = 0; k < D; ++k) I[k] = a.idx1d[(long long)row * D + k];
    for (int e = tid; e < B; e += 256) acc[e] = T(0);
    long long pb = a.row_ptr ? a.row_ptr[row] : 0;
    long long pe = a.row_ptr ? a.row_ptr[row + 1] : a.N;
    for (long long q = pb; q < pe; ++q) {
        const int j = a.row_ptr ? a.cols[q] : int(q);
        for (int k = 0; k < D; ++k) J[k] = a.idx1d[(long long)j * D + k];
        for (int t = 0; t < a.nterms; ++t) {
            __syncthreads();
            for (int e = tid; e < B; e += 256) v[e] = T(a.x[(long long)j * B + e]);
            __syncthreads();
            T* vin = v;
            T* vout = w;
            int stride = 1;
            for (int k = D - 1; k >= 0; --k) {
                const T* m = a.blk + a.off[t * D + k] + ((long long)I[k] * a.nb[k] + J[k]) * (DEG * DEG);
                for (int e = tid; e < B; e += 256) {
                    int ai = (e / stride) % DEG;
                    int base = e - ai * stride;
                    T s = m[ai * DEG] * vin[base];
                    for (int b = 1; b < DEG; ++b) s = madd(m[ai * DEG + b], vin[base + b * stride], s);
                    vout[e] = s;
                }
                __syncthreads();
                T* tmp = vin;
                vin = vout;
                vout = tmp;
                stride *= DEG;
            }
            const T s = a.scale[t];
            for (int e = tid; e < B; e += 256) acc[e] = madd(s, vin[e], acc[e]);
        }
    }
    __syncthreads();
    for (int e = tid; e < B; e += 256) a.y[(long long)row * B + e] = double(acc[e]);
}

// -------------------------------------------------------------- grouped FP32
// Pairs of a row are pre-sorted so that pairs sharing (J_1..J_{d-1}) are
// adjacent ("prefix groups").  A lane takes one group: it contracts dim d of
// every member's x_j per term and sums them (W_t), then applies dims d-1..1
// once per group and term.  Same sum as the pair-by-pair chain, fewer flops.
template <int D, int DEG, int NT, int TMAX>
__global__ void __launch_bounds__(NT) k_apply_grouped(KArgs<float> a) {
    constexpr int B = Pow<DEG, D>::v;
    constexpr int NW = NT / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    const int row = a.row_order[blockIdx.x];
    const int tid = threadIdx.x;
    RowCtx<float, D, DEG, true> c;
    c.a = &a;
#pragma unroll
    for (int k = 0; k < D; ++k) c.I[k] = a.idx1d[(long long)row * D + k];
    float* strip = reinterpret_cast<float*>(smem);
    {
        int so = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            c.soff[k] = so;
            so += a.nb[k] * DEG * DEG;
        }
        c.sterm = so;
        for (int t = 0; t < a.nterms; ++t) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float* src = a.blk + a.off[t * D + k] + (long long)c.I[k] * a.nb[k] * DEG * DEG;
                float* dst = strip + t * so + c.soff[k];
                const int len = a.nb[k] * DEG * DEG;
                for (int e = tid; e < len; e += NT) dst[e] = src[e];
            }
        }
        __syncthreads();
        c.strip = strip;
    }
    float acc[B];
#pragma unroll
    for (int e = 0; e < B; ++e) acc[e] = 0.f;

    const long long gb = a.grp_row_ptr[row], ge = a.grp_row_ptr[row + 1] - 1;  // groups gb..ge-1
    for (long long g = gb + tid; g < ge; g += NT) {
        const int p0 = a.grp_ptr[g], p1 = a.grp_ptr[g + 1];
        int J[D];
        {
            const int j0 = a.cols[p0];
#pragma unroll
            for (int k = 0; k < D - 1; ++k) J[k] = a.idx1d[(long long)j0 * D + k];
        }
        float W[TMAX][B];
#pragma unroll
        for (int t = 0; t < TMAX; ++t)
#pragma unroll
            for (int e = 0; e < B; ++e) W[t][e] = 0.f;
        for (int p = p0; p < p1; ++p) {
            const int j = a.cols[p];
            const int Jd = a.idx1d[(long long)j * D + (D - 1)];
            float xv[B];
            const double* xj = a.x + (long long)j * B;
#pragma unroll
            for (int e = 0; e < B; ++e) xv[e] = float(xj[e]);
#pragma unroll
            for (int t = 0; t < TMAX; ++t) {
                if (t < a.nterms) {
                    float m[DEG * DEG];
                    c.load_block(t, D - 1, Jd, m);
                    float v[B];
#pragma unroll
                    for (int e = 0; e < B; ++e) v[e] = xv[e];
                    contract_dim<float, D, DEG, D - 1>(v, m);
#pragma unroll
                    for (int e = 0; e < B; ++e) W[t][e] += v[e];
                }
            }
        }
#pragma unroll
        for (int t = 0; t < TMAX; ++t) {
            if (t < a.nterms) {
                Chain<float, D, DEG, true, D - 2>::run(W[t], c, t, J);
                const float s = a.scale[t];
#pragma unroll
                for (int e = 0; e < B; ++e) acc[e] = madd(s, W[t][e], acc[e]);
            }
        }
    }
    __syncthreads();
    float* red = reinterpret_cast<float*>(smem);
    const int lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int e = 0; e < B; ++e) {
        float s = wave_sum(acc[e]);
        if (lane == 0) red[w * B + e] = s;
    }
    __syncthreads();
    for (int e = tid; e < B; e += NT) {
        float s = red[e];
#pragma unroll
        for (int ww = 1; ww < NW; ++ww) s += red[ww * B + e];
        a.y[(long long)row * B + e] = double(s);
    }
}

// ---------------------------------------------------------------- dispatch
using LaunchFn = hipError_t (*)(const void* args, int nrows, size_t smem, hipStream_t s);

template <typename T, int D, int DEG, int NT, bool USE_LDS, bool KEEPX>
hipError_t launch_reg(const void* args, int nrows, size_t smem, hipStream_t s) {
    auto kern = k_apply_reg<T, D, DEG, NT, USE_LDS, KEEPX>;
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nrows), dim3(NT), smem, s, *reinterpret_cast<const KArgs<T>*>(args));
    return hipGetLastError();
}

template <typename T, int D, int DEG>
constexpr bool reg_ok() {
    return Pow<DEG, D>::v * int(sizeof(T)) <= 256;  // B <= 32 (FP64) / 64 (FP32)
}
template <typename T, int D, int DEG>
constexpr bool keep_x() {
    return 3 * Pow<DEG, D>::v * int(sizeof(T)) / 4 <= 144;
}

template <typename T, int D, int DEG>
LaunchFn pick_reg(bool lds) {
    if constexpr (reg_ok<T, D, DEG>()) {
        constexpr bool KX = keep_x<T, D, DEG>();
        return lds ? &launch_reg<T, D, DEG, 256, true, KX> : &launch_reg<T, D, DEG, 256, false, KX>;
    } else {
        (void)lds;
        return nullptr;
    }
}

template <typename T, int D>
LaunchFn pick_deg(int deg, bool lds) {
    switch (deg) {
        case 1: return pick_reg<T, D, 1>(lds);
        case 2: return pick_reg<T, D, 2>(lds);
        case 3: return pick_reg<T, D, 3>(lds);
        case 4: return pick_reg<T, D, 4>(lds);
        case 5: return pick_reg<T, D, 5>(lds);
        case 6: return pick_reg<T, D, 6>(lds);
        case 7: return pick_reg<T, D, 7>(lds);
        case 8: return pick_reg<T, D, 8>(lds);
        default: return nullptr;
    }
}

template <typename T>
LaunchFn pick(int D, int deg, bool lds) {
    switch (D) {
        case 1: return pick_deg<T, 1>(deg, lds);
        case 2: return pick_deg<T, 2>(deg, lds);
        case 3: return pick_deg<T, 3>(deg, lds);
        case 4: return pick_deg<T, 4>(deg, lds);
        case 5: return pick_deg<T, 5>(deg, lds);
        case 6: return pick_deg<T, 6>(deg, lds);
        default: return nullptr;
    }
}

// grouped FP32 kernels are built for the configurations that pay off
template <int D, int DEG, int TMAX>
hipError_t launch_grouped(const void* args, int nrows, size_t smem, hipStream_t s) {
    auto kern = k_apply_grouped<D, DEG, 256, TMAX>;
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nrows), dim3(256), smem, s, *reinterpret_cast<const KArgs<float>*>(args));
    return hipGetLastError();
}

LaunchFn pick_grouped(int D, int deg, int T) {
    if (T > 4) return nullptr;
    if (D == 2 && deg == 2) return &launch_grouped<2, 2, 4>;
    if (D == 2 && deg == 3) return &launch_grouped<2, 3, 4>;
    if (D == 2 && deg == 4) return &launch_grouped<2, 4, 4>;
    if (D == 3 && deg == 2) return &launch_grouped<3, 2, 4>;
    if (D == 3 && deg == 3) return T <= 3 ? &launch_grouped<3, 3, 3> : &launch_grouped<3, 3, 4>;
    return nullptr;
}

}  // namespace kmx

using namespace kmx;

static size_t lds_limit() {
    static size_t lim = [] {
        const char* e = getenv("KMX_LDS_LIMIT");
        return e ? size_t(atol(e)) : size_t(160 * 1024);
    }();
    return lim;
}

template <typename T>
static void fill_args(KArgs<T>& a, const kmx_plan* p, bool transpose, const double* x, double* y) {
    std::memset(&a, 0, sizeof(a));
    if (std::is_same<T, double>::value)
        a.blk = reinterpret_cast<const T*>(transpose ? p->d_blkT64 : p->d_blk64);
    else
        a.blk = reinterpret_cast<const T*>(transpose ? p->d_blkT32 : p->d_blk32);
    for (int i = 0; i < p->T * p->D; ++i) a.off[i] = p->off[i];
    for (int k = 0; k < p->D; ++k) a.nb[k] = p->nb[k];
    for (int t = 0; t < p->T; ++t) a.scale[t] = T(p->scale[t]);
    a.idx1d = p->d_idx1d;
    a.row_ptr = transpose ? p->d_col_ptr : p->d_row_ptr;
    a.cols = transpose ? p->d_rows : p->d_cols;
    a.row_order = transpose ? p->d_col_order : p->d_row_order;
    a.x = x;
    a.y = y;
    a.N = p->N;
    a.nterms = p->T;
    int se = 0;
    for (int k = 0; k < p->D; ++k) se += p->nb[k] * p->deg * p->deg;
    a.strip_elems = se * p->T;
}

int kmx_launch_apply(kmx_plan* p, bool transpose, const double* x, double* y, int prec, hipStream_t s) {
    if (p->N == 0) return KMX_OK;
    const int D = p->D, deg = p->deg, B = p->B;
    hipError_t err = hipSuccess;
    if (prec == KMX_PREC_FP32_GROUPED && (transpose || !p->d_gcols)) prec = KMX_PREC_FP32;  // no groups: same sum pair by pair
    if (prec == KMX_PREC_FP32_GROUPED) {
        KArgs<float> a;
        fill_args(a, p, false, x, y);
        a.cols = p->d_gcols;
        a.grp_ptr = p->d_grp_ptr;
        a.grp_row_ptr = p->d_grp_row_ptr;
        LaunchFn f = pick_grouped(D, deg, p->T);
        size_t smem = std::max(size_t(a.strip_elems) * sizeof(float), size_t(4 * B) * sizeof(float));
        if (f && smem <= lds_limit()) {
            err = f(&a, p->N, smem, s);
            if (err != hipSuccess) {
                set_error(std::string("grouped apply launch: ") + hipGetErrorString(err));
                return KMX_E_HIP;
            }
            return KMX_OK;
        }
        prec = KMX_PREC_FP32;  // no grouped kernel for this shape: same sum pair by pair
    }
    if (prec == KMX_PREC_FP64) {
        KArgs<double> a;
        fill_args(a, p, transpose, x, y);
        size_t strip = size_t(a.strip_elems) * sizeof(double);
        bool lds = strip <= lds_limit();
        LaunchFn f = pick<double>(D, deg, lds);
        if (f) {
            size_t smem = std::max(lds ? strip : size_t(0), size_t(4 * B) * sizeof(double));
            err = f(&a, p->N, smem, s);
        } else {
            size_t smem = size_t(3 * B) * sizeof(double);
            if (smem > 160 * 1024) {
                set_error("kmx_apply: block too large for the generic kernel");
                return KMX_E_UNSUPPORTED;
            }
            if (smem > 65536)
                hipFuncSetAttribute(reinterpret_cast<const void*>(k_apply_generic<double>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
            hipLaunchKernelGGL(k_apply_generic<double>, dim3(p->N), dim3(256), smem, s, a, D, deg);
            err = hipGetLastError();
        }
    } else if (prec == KMX_PREC_FP32) {
        if (!(transpose ? p->d_blkT32 : p->d_blk32)) {
            set_error("kmx_apply: plan has no FP32 blocks (create it with KMX_PLAN_FP32)");
            return KMX_E_ARG;
        }
        KArgs<float> a;
        fill_args(a, p, transpose, x, y);
        size_t strip = size_t(a.strip_elems) * sizeof(float);
        bool lds = strip <= lds_limit();
        LaunchFn f = pick<float>(D, deg, lds);
        if (f) {
            size_t smem = std::max(lds ? strip : size_t(0), size_t(4 * B) * sizeof(float));
            err = f(&a, p->N, smem, s);
        } else {
            size_t smem = size_t(3 * B) * sizeof(float);
            if (smem > 160 * 1024) {
                set_error("kmx_apply: block too large for the generic kernel");
                return KMX_E_UNSUPPORTED;
            }
            if (smem > 65536)
                hipFuncSetAttribute(reinterpret_cast<const void*>(k_apply_generic<float>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
            hipLaunchKernelGGL(k_apply_generic<float>, dim3(p->N), dim3(256), smem, s, a, D, deg);
            err = hipGetLastError();
        }
    } else {
        set_error("kmx_apply: unknown precision");
        return KMX_E_ARG;
    }
    if (err != hipSuccess) {
        set_error(std::string("apply launch: ") + hipGetErrorString(err));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

// ------------------------------------------------------------------ batches
struct kmx_batch {
    std::vector<kmx_plan*> plans;
    std::vector<int> in_kernel;  // plan indices served by the persistent kernel (slot order)
    std::vector<int> solo;       // plans launched on their own (shape not in the switch)
    KArgs<double>* d_args64 = nullptr;
    KArgs<float>* d_args32 = nullptr;
    int* d_shape = nullptr;
    int2* d_items = nullptr;
    int nitems = 0;
    int maxB = 1;
    int grid64 = 0, grid32 = 0;
    bool has32 = true;
};

static bool batch_shape_ok(const kmx_plan* p, bool fp64) {
    if (p->D > 3 || p->deg > 4) return false;
    return p->B * int(fp64 ? sizeof(double) : sizeof(float)) <= 256;
}

extern "C" int kmx_batch_create(kmx_plan* const* plans, int count, kmx_batch** out) {
    if (!plans || count < 0 || !out) {
        set_error("kmx_batch_create: bad arguments");
        return KMX_E_ARG;
    }
    *out = nullptr;
    kmx_batch* b = new kmx_batch();
    b->plans.assign(plans, plans + count);
    std::vector<KArgs<double>> a64;
    std::vector<KArgs<float>> a32;
    std::vector<int> shape;
    struct Item {
        double cost;
        int slot, row;
    };
    std::vector<Item> items;
    std::vector<long long> rp;
    for (int i = 0; i < count; ++i) {
        kmx_plan* p = plans[i];
        if (!p) {
            delete b;
            set_error("kmx_batch_create: null plan");
            return KMX_E_ARG;
        }
        if (!p->d_blk32) b->has32 = false;
        // a plan joins the persistent kernel only if both precisions fit its switch
        if (!batch_shape_ok(p, true) || !batch_shape_ok(p, false) || int(b->in_kernel.size()) >= KMX_BATCH_MAX) {
            b->solo.push_back(i);
            continue;
        }
        const int slot = int(b->in_kernel.size());
        b->in_kernel.push_back(i);
        KArgs<double> x64;
        fill_args(x64, p, false, nullptr, nullptr);
        a64.push_back(x64);
        KArgs<float> x32;
        fill_args(x32, p, false, nullptr, nullptr);
        a32.push_back(x32);
        shape.push_back(p->D * 16 + p->deg);
        b->maxB = std::max(b->maxB, p->B);
        // per-row pair counts from the device row_ptr are not needed: recompute
        rp.assign(size_t(p->N) + 1, 0);
        if (p->all_pairs) {
            for (int r = 0; r <= p->N; ++r) rp[r] = (long long)r * p->N;
        } else if (p->N > 0) {
            hipMemcpy(rp.data(), p->d_row_ptr, sizeof(long long) * (p->N + 1), hipMemcpyDeviceToHost);
        }
        const double unit = double(p->T) * p->B * p->D * p->deg;
        for (int r = 0; r < p->N; ++r) items.push_back({double(rp[r + 1] - rp[r]) * unit, slot, r});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
    std::vector<int2> it(items.size());
    for (size_t q = 0; q < items.size(); ++q) it[q] = make_int2(items[q].slot, items[q].row);
    b->nitems = int(it.size());
    auto up = [](const void* h, size_t bytes, void** d) {
        if (bytes == 0) return hipSuccess;
        hipError_t e = hipMalloc(d, bytes);
        if (e != hipSuccess) return e;
        return hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice);
    };
    hipError_t e = hipSuccess;
    if (!a64.empty()) {
        e = up(a64.data(), a64.size() * sizeof(KArgs<double>), reinterpret_cast<void**>(&b->d_args64));
        if (e == hipSuccess) e = up(a32.data(), a32.size() * sizeof(KArgs<float>), reinterpret_cast<void**>(&b->d_args32));
        if (e == hipSuccess) e = up(shape.data(), shape.size() * sizeof(int), reinterpret_cast<void**>(&b->d_shape));
        if (e == hipSuccess) e = up(it.data(), it.size() * sizeof(int2), reinterpret_cast<void**>(&b->d_items));
    }
    if (e != hipSuccess) {
        set_error(std::string("kmx_batch_create: ") + hipGetErrorString(e));
        kmx_batch_destroy(b);
        return KMX_E_HIP;
    }
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int occ64 = 1, occ32 = 1;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ64, reinterpret_cast<const void*>(k_apply_batch<double>), 256,
                                                 size_t(4 * b->maxB) * sizeof(double));
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ32, reinterpret_cast<const void*>(k_apply_batch<float>), 256,
                                                 size_t(4 * b->maxB) * sizeof(float));
    b->grid64 = std::max(1, std::min(b->nitems, cus * std::max(1, occ64)));
    b->grid32 = std::max(1, std::min(b->nitems, cus * std::max(1, occ32)));
    *out = b;
    return KMX_OK;
}

extern "C" int kmx_batch_apply(kmx_batch* b, const double* const* xs, double* const* ys, int prec, void* stream) {
    if (!b || (!b->plans.empty() && (!xs || !ys))) {
        set_error("kmx_batch_apply: bad arguments");
        return KMX_E_ARG;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (prec == KMX_PREC_FP32_GROUPED) prec = KMX_PREC_FP32;
    if (prec == KMX_PREC_FP32 && !b->has32) {
        set_error("kmx_batch_apply: a plan has no FP32 blocks");
        return KMX_E_ARG;
    }
    if (prec != KMX_PREC_FP32 && prec != KMX_PREC_FP64) {
        set_error("kmx_batch_apply: unknown precision");
        return KMX_E_ARG;
    }
    if (b->nitems > 0) {
        BatchIO io;
        std::memset(&io, 0, sizeof(io));
        for (size_t q = 0; q < b->in_kernel.size(); ++q) {
            io.x[q] = xs[b->in_kernel[q]];
            io.y[q] = ys[b->in_kernel[q]];
            if (!io.x[q] || !io.y[q]) {
                set_error("kmx_batch_apply: null vector");
                return KMX_E_ARG;
            }
        }
        if (prec == KMX_PREC_FP64)
            hipLaunchKernelGGL(k_apply_batch<double>, dim3(b->grid64), dim3(256), size_t(4 * b->maxB) * sizeof(double),
                               s, b->d_args64, b->d_items, b->nitems, b->d_shape, io);
        else
            hipLaunchKernelGGL(k_apply_batch<float>, dim3(b->grid32), dim3(256), size_t(4 * b->maxB) * sizeof(float),
                               s, b->d_args32, b->d_items, b->nitems, b->d_shape, io);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error(std::string("kmx_batch_apply: ") + hipGetErrorString(e));
            return KMX_E_HIP;
        }
    }
    for (int i : b->solo) {
        int rc = kmx_launch_apply(b->plans[i], false, xs[i], ys[i], prec, s);
        if (rc != KMX_OK) return rc;
    }
    return KMX_OK;
}

extern "C" int kmx_batch_destroy(kmx_batch* b) {
    if (!b) return KMX_OK;
    if (b->d_args64) hipFree(b->d_args64);
    if (b->d_args32) hipFree(b->d_args32);
    if (b->d_shape) hipFree(b->d_shape);
    if (b->d_items) hipFree(b->d_items);
    delete b;
    return KMX_OK;
}
