// kmx_kernels.h -- gfx950 device code for y = A x with A = sum over element
// pairs (i, j) and terms t of kron_k A_{t,k}[I_k(i), J_k(j)]  (apply_A.m:100-237).
//
// Work decomposition (CDNA4, wave64):
//   * one workgroup per row element i (rows issued heaviest first); its NT
//     threads stride over i's pair list, one pair per lane;
//   * a lane keeps the deg^d block of x_j and its partial y_i in registers and
//     runs the d contractions of each term in place, fibre by fibre, dim d
//     first and dim 1 last (the order of kronmult{d} method 1,
//     Kronmult/kronmult3.m:73-89);
//   * the deg x deg coefficient blocks A_{t,k}[I_k(i), :] of the row are staged
//     once per workgroup into LDS when they fit (they are contiguous in the
//     plan's block-row-major layout), else read through L2;
//   * y_i is reduced across the workgroup -- a halving exchange between lanes
//     (each step a lane keeps half of its live entries and adds its partner's
//     copy of that half), then one LDS pass over the waves -- and written once:
//     no atomics, deterministic.
// Blocks too large for registers go to a generic LDS kernel (one pair at a
// time, threads over block entries).
#pragma once

#include "kmx_internal.h"

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

constexpr int next_pow2(int v) {
    int p = 1;
    while (p < v) p *= 2;
    return p;
}
constexpr int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

__device__ __forceinline__ double madd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// In-place contraction of dim K (0 = slowest) of a register block with the
// deg x deg block m (row-major [a][b]): v[.., a, ..] = sum_b m[a][b] v[.., b, ..].
template <typename T, int D, int DEG, int K>
__device__ __forceinline__ void contract_dim(RegArr<T, Pow<DEG, D>::v>& v, const RegArr<T, DEG * DEG>& m) {
    constexpr int S = Pow<DEG, D - 1 - K>::v;
    constexpr int O = Pow<DEG, D>::v / (S * DEG);
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
template <typename T, int D, int DEG, int USE_LDS>
struct RowCtx {
    const KArgs<T>* a;
    const T* strip;  // LDS strips of this row (USE_LDS)
    int I[D];        // row's 0-based 1D indices
    int soff[D];     // LDS offset of (t=0, k) strip
    int sterm;       // LDS elements per term

    __device__ __forceinline__ void load_block(int t, int k, int J, RegArr<T, DEG * DEG>& m) const {
        const T* p;
        if (USE_LDS == 1 || (USE_LDS == 2 && k != 0)) {  // mode 2: dim 1's blocks come from L2
            p = strip + t * sterm + soff[k] + J * (DEG * DEG);
        } else {
            p = a->blk + a->off[t * D + k] + ((long long)I[k] * a->nb[k] + J) * (DEG * DEG);
        }
#pragma unroll
        for (int e = 0; e < DEG * DEG; ++e) m[e] = p[e];
    }
};

template <typename T, int D, int DEG, int USE_LDS, int K>
struct Chain {
    __device__ __forceinline__ static void run(RegArr<T, Pow<DEG, D>::v>& v, const RowCtx<T, D, DEG, USE_LDS>& c,
                                               int t, const RegArr<int, D>& J) {
        T m[DEG * DEG];
        c.load_block(t, K, J[K], m);
        contract_dim<T, D, DEG, K>(v, m);
        Chain<T, D, DEG, USE_LDS, K - 1>::run(v, c, t, J);
    }
};
template <typename T, int D, int DEG, int USE_LDS>
struct Chain<T, D, DEG, USE_LDS, -1> {
    __device__ __forceinline__ static void run(RegArr<T, Pow<DEG, D>::v>&, const RowCtx<T, D, DEG, USE_LDS>&, int,
                                               const RegArr<int, D>&) {}
};

// Contractions of dims D-1..1 of a lane's dim-1 slice (a (D-1)-dim block of
// DEG^(D-1) values): dim k of the block is dim k-1 of the slice.
template <typename T, int D, int DEG, int USE_LDS, int K>
struct SliceChain {
    __device__ __forceinline__ static void run(RegArr<T, Pow<DEG, D - 1>::v>& v, const RowCtx<T, D, DEG, USE_LDS>& c,
                                               int t, const RegArr<int, D>& J) {
        T m[DEG * DEG];
        c.load_block(t, K, J[K], m);
        contract_dim<T, D - 1, DEG, K - 1>(v, m);
        SliceChain<T, D, DEG, USE_LDS, K - 1>::run(v, c, t, J);
    }
};
template <typename T, int D, int DEG, int USE_LDS>
struct SliceChain<T, D, DEG, USE_LDS, 0> {
    __device__ __forceinline__ static void run(RegArr<T, Pow<DEG, D - 1>::v>&, const RowCtx<T, D, DEG, USE_LDS>&,
                                               int, const RegArr<int, D>&) {}
};

template <typename T, int G, int BS, int SRC>
struct GatherAll;

// Copy `len` elements global -> LDS with 16-byte accesses when aligned.
template <typename T, int NT>
__device__ __forceinline__ void stage(T* __restrict__ dst, const T* __restrict__ src, int len, int tid) {
    const bool vec = (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0) &&
                     ((len * int(sizeof(T))) & 15) == 0;
    if (vec) {
        const int n4 = len * int(sizeof(T)) / 16;
        const float4* s4 = reinterpret_cast<const float4*>(src);
        float4* d4 = reinterpret_cast<float4*>(dst);
        for (int e = tid; e < n4; e += NT) d4[e] = s4[e];
    } else {
        for (int e = tid; e < len; e += NT) dst[e] = src[e];
    }
}

// Sum acc[B] over the 64 lanes of the wave and store the B sums at red[0..B).
// Halving exchange: with BP = next_pow2(B) live entries, step s (mask 32>>s)
// keeps the half selected by lane bit (5-s) and adds the partner's copy of it,
// so the shuffles total BP/2 + BP/4 + ... + 1 (+ the plain steps left when
// BP < 64) instead of 6*B.  Lane L ends with entry idx(L) = sum over steps of
// bit_s(L) * BP >> (s+1).
template <typename T, int BP, int S, int K>
struct Halve {
    __device__ __forceinline__ static void run(RegArr<T, BP>& v, int lane, int& idx) {
        constexpr int mask = 32 >> S;
        constexpr int half = BP >> (S + 1);
        const bool hi = (lane & mask) != 0;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const T send = hi ? v[i] : v[i + half];
            const T keep = hi ? v[i + half] : v[i];
            v[i] = keep + __shfl_xor(send, mask, 64);
        }
        idx += hi ? half : 0;
        Halve<T, BP, S + 1, K>::run(v, lane, idx);
    }
};
template <typename T, int BP, int K>
struct Halve<T, BP, K, K> {
    __device__ __forceinline__ static void run(RegArr<T, BP>&, int, int&) {}
};

template <typename T, int B>
__device__ __forceinline__ void wave_reduce_store(const RegArr<T, B>& acc, T* red, int lane) {
    constexpr int BP = next_pow2(B) > 64 ? 64 : next_pow2(B);
    constexpr int K = ilog2(BP);
    static_assert(B <= 64, "register path blocks hold at most 64 values");
    T v[BP];
#pragma unroll
    for (int e = 0; e < BP; ++e) v[e] = (e < B) ? acc[e] : T(0);
    int idx = 0;
    Halve<T, BP, 0, K>::run(v, lane, idx);
    T r = v[0];
#pragma unroll
    for (int s = K; s < 6; ++s) r += __shfl_xor(r, 32 >> s, 64);
    // the lanes sharing the top K lane bits now hold the same sum; one stores it
    const int low = (64 >> K) - 1;
    if ((lane & low) == 0 && idx < B) red[idx] = r;
}

// One row: NT lanes stride over the pair list; y_i reduced and written once.
template <typename T, int D, int DEG, int NT, int USE_LDS, bool KEEPX>
__device__ __forceinline__ void apply_row(const KArgs<T>& a, const double* __restrict__ x, double* __restrict__ y,
                                          int row, unsigned char* smem) {
    constexpr int B = Pow<DEG, D>::v;
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x;

    RowCtx<T, D, DEG, USE_LDS> c;
    c.a = &a;
#pragma unroll
    for (int k = 0; k < D; ++k) c.I[k] = a.idx1d[(long long)row * D + k];

    if (USE_LDS) {  // mode 1: every dim's strip; mode 2: all but dim 1's
        T* strip = reinterpret_cast<T*>(smem);
        constexpr int K0 = USE_LDS == 2 ? 1 : 0;
        int so = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            c.soff[k] = so;
            if (k >= K0) so += a.nb[k] * DEG * DEG;
        }
        c.sterm = so;
        for (int t = 0; t < a.nterms; ++t) {
#pragma unroll
            for (int k = K0; k < D; ++k)
                stage<T, NT>(strip + t * so + c.soff[k], a.blk + a.off[t * D + k] + (long long)c.I[k] * a.nb[k] * DEG * DEG,
                             a.nb[k] * DEG * DEG, tid);
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
    wave_reduce_store<T, B>(acc, red + w * B, lane);
    __syncthreads();
    for (int e = tid; e < B; e += NT) {
        T s = red[e];
#pragma unroll
        for (int ww = 1; ww < NW; ++ww) s += red[ww * B + e];
        y[(long long)row * B + e] = double(s);
    }
}

template <typename T, int D, int DEG, int NT, int USE_LDS, bool KEEPX>
__global__ void __launch_bounds__(NT) k_apply_reg(KArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int row = a.row_order[blockIdx.x];
    apply_row<T, D, DEG, NT, USE_LDS, KEEPX>(a, a.x, a.y, row, smem);
}

// --------------------------------------------------------------- lane split
// Blocks too large for one lane (64 values in FP64): G = DEG lanes share a
// pair, lane r of the group holding the dim-1 slice b_1 = r (DEG^(D-1)
// contiguous values, dim 1 being the slowest).  Dims D..2 contract inside the
// slice; the dim-1 contraction reads the other slices of the group through
// DPP quad permutes (G = 2 or 4 lanes of one quad), so no LDS traffic.
// Value of lane (group base + SRC) for groups of G lanes inside a DPP quad:
// quad_perm [SRC,SRC,SRC,SRC] for G = 4, [SRC,SRC,2+SRC,2+SRC] for G = 2.
template <int G, int SRC>
constexpr int quad_ctrl() {
    return G == 4 ? SRC * 85 : (SRC | (SRC << 2) | ((2 + SRC) << 4) | ((2 + SRC) << 6));
}

template <typename T, int G, int SRC>
__device__ __forceinline__ T quad_from(T v) {
    constexpr int ctrl = quad_ctrl<G, SRC>();
    if constexpr (sizeof(T) == 8) {
        const long long bits = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_update_dpp(0, int(bits), ctrl, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, int(bits >> 32), ctrl, 0xF, 0xF, false);
        return __builtin_bit_cast(T, (long long)(unsigned)lo | ((long long)hi << 32));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), ctrl, 0xF, 0xF, false));
    }
}

template <typename T, int G, int BS, int SRC>
struct GatherFMA {
    // out[q] += m[r][SRC] * (slice SRC)[q]
    __device__ __forceinline__ static void run(RegArr<T, BS>& out, const RegArr<T, BS>& v, T coef) {
#pragma unroll
        for (int q = 0; q < BS; ++q) out[q] = madd(coef, quad_from<T, G, SRC>(v[q]), out[q]);
    }
};

template <typename T, int G, int BS, int SRC>
struct GatherAll {
    __device__ __forceinline__ static void run(RegArr<T, BS>& out, const RegArr<T, BS>& v, const RegArr<T, G>& coef) {
        GatherFMA<T, G, BS, SRC>::run(out, v, coef[SRC]);
        GatherAll<T, G, BS, SRC + 1>::run(out, v, coef);
    }
};
template <typename T, int BS>
struct GatherAll<T, 2, BS, 2> {
    __device__ __forceinline__ static void run(RegArr<T, BS>&, const RegArr<T, BS>&, const RegArr<T, 2>&) {}
};
template <typename T, int BS>
struct GatherAll<T, 4, BS, 4> {
    __device__ __forceinline__ static void run(RegArr<T, BS>&, const RegArr<T, BS>&, const RegArr<T, 4>&) {}
};

// Sum acc[BS] over the lanes of the wave that share the low LB lane bits (the
// slice id) and store slice r's sums at red[r*BS + ..].
template <typename T, int BP, int S, int K, int LB>
struct HalveS {
    __device__ __forceinline__ static void run(RegArr<T, BP>& v, int lane, int& idx) {
        constexpr int mask = 32 >> S;
        constexpr int half = BP >> (S + 1);
        const bool hi = (lane & mask) != 0;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const T send = hi ? v[i] : v[i + half];
            const T keep = hi ? v[i + half] : v[i];
            v[i] = keep + __shfl_xor(send, mask, 64);
        }
        idx += hi ? half : 0;
        HalveS<T, BP, S + 1, K, LB>::run(v, lane, idx);
    }
};
template <typename T, int BP, int K, int LB>
struct HalveS<T, BP, K, K, LB> {
    __device__ __forceinline__ static void run(RegArr<T, BP>&, int, int&) {}
};

template <typename T, int BS, int LB>
__device__ __forceinline__ void wave_reduce_store_slices(const RegArr<T, BS>& acc, T* red, int lane) {
    constexpr int BP = next_pow2(BS);
    constexpr int LEVELS = 6 - LB;  // lane bits to reduce over
    constexpr int K = ilog2(BP) < LEVELS ? ilog2(BP) : LEVELS;
    T v[BP];
#pragma unroll
    for (int e = 0; e < BP; ++e) v[e] = (e < BS) ? acc[e] : T(0);
    int idx = 0;
    HalveS<T, BP, 0, K, LB>::run(v, lane, idx);
    // after K halvings each lane holds BP >> K entries; finish with plain steps
    constexpr int LIVE = BP >> K;
#pragma unroll
    for (int s = K; s < LEVELS; ++s)
#pragma unroll
        for (int i = 0; i < LIVE; ++i) v[i] += __shfl_xor(v[i], 32 >> s, 64);
    constexpr int plain = ((64 >> K) - 1) & ~((1 << LB) - 1);
    const int r = lane & ((1 << LB) - 1);
    if ((lane & plain) == 0) {
#pragma unroll
        for (int i = 0; i < LIVE; ++i)
            if (idx + i < BS) red[r * BS + idx + i] = v[i];
    }
}

template <typename T, int D, int DEG, int NT, int USE_LDS>
__device__ __forceinline__ void apply_row_split(const KArgs<T>& a, const double* __restrict__ x,
                                                double* __restrict__ y, int row, unsigned char* smem) {
    constexpr int B = Pow<DEG, D>::v;
    constexpr int BS = B / DEG;  // slice of one lane
    constexpr int G = DEG;
    constexpr int LB = ilog2(G);
    constexpr int NW = NT / 64;
    static_assert(G == 2 || G == 4, "lane split needs 2 or 4 lanes per pair");
    const int tid = threadIdx.x;
    const int r = tid & (G - 1);

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
            for (int k = 0; k < D; ++k)
                stage<T, NT>(strip + t * so + c.soff[k], a.blk + a.off[t * D + k] + (long long)c.I[k] * a.nb[k] * DEG * DEG,
                             a.nb[k] * DEG * DEG, tid);
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
    T acc[BS];
#pragma unroll
    for (int q = 0; q < BS; ++q) acc[q] = T(0);
    // every lane of a group runs the same trip count (the DPP reads need the
    // partner lanes active); lanes past the end work on a clamped pair and
    // drop the result
    const long long npairs = pe - pb;
    const long long iters = (npairs + (NT / G) - 1) / (NT / G);
    for (long long it = 0; it < iters; ++it) {
        const long long qq = it * (NT / G) + tid / G;
        const bool live = qq < npairs;
        const long long q = pb + (live ? qq : 0);
        if (npairs == 0) break;
        const int j = a.row_ptr ? a.cols[q] : int(q);
        int J[D];
#pragma unroll
        for (int k = 0; k < D; ++k) J[k] = a.idx1d[(long long)j * D + k];
        const double* xs = x + (long long)j * B + r * BS;
        for (int t = 0; t < a.nterms; ++t) {
            T v[BS];
#pragma unroll
            for (int e = 0; e < BS; ++e) v[e] = T(xs[e]);
            // dims D..2 inside the slice (a (D-1)-dim block)
            SliceChain<T, D, DEG, USE_LDS, D - 1>::run(v, c, t, J);
            // dim 1 across the group
            T m[DEG * DEG];
            c.load_block(t, 0, J[0], m);
            // acc[r-slice] += scale_t * sum_b m[r][b] * (slice b): the term scale
            // is folded into the coefficients so no extra slice buffer is live
            const T s = live ? a.scale[t] : T(0);
            T coef[G];
#pragma unroll
            for (int b = 0; b < G; ++b) coef[b] = s * m[0 * DEG + b];
#pragma unroll
            for (int a0 = 1; a0 < G; ++a0)
                if (r == a0)
#pragma unroll
                    for (int b = 0; b < G; ++b) coef[b] = s * m[a0 * DEG + b];
            GatherAll<T, G, BS, 0>::run(acc, v, coef);
        }
    }
    if (USE_LDS) __syncthreads();
    T* red = reinterpret_cast<T*>(smem);
    const int lane = tid & 63, w = tid >> 6;
    wave_reduce_store_slices<T, BS, LB>(acc, red + w * B, lane);
    __syncthreads();
    for (int e = tid; e < B; e += NT) {
        T sum = red[e];
#pragma unroll
        for (int ww = 1; ww < NW; ++ww) sum += red[ww * B + e];
        y[(long long)row * B + e] = double(sum);
    }
}

template <typename T, int D, int DEG, int NT, int USE_LDS>
__global__ void __launch_bounds__(NT) k_apply_split(KArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    apply_row_split<T, D, DEG, NT, USE_LDS>(a, a.x, a.y, a.row_order[blockIdx.x], smem);
}

// ------------------------------------------------------------ many instances
// One persistent launch over a work list of (instance, row) items, heaviest
// first, striding by the grid; each instance keeps its own plan (levels, deg,
// pair list, coefficient blocks).  The shape switch selects the fully unrolled
// row body for the instance's (d, deg).  (BatchIO: kmx_internal.h)

template <typename T, int D, int DEG>
__device__ __forceinline__ void batch_row(const KArgs<T>& a, const double* x, double* y, int row, unsigned char* sm) {
    if constexpr (Pow<DEG, D>::v <= 27) {
        constexpr bool KX = 3 * Pow<DEG, D>::v * int(sizeof(T)) / 4 <= 144;
        apply_row<T, D, DEG, 256, false, KX>(a, x, y, row, sm);
    } else if constexpr (sizeof(T) == 8 && (DEG == 2 || DEG == 4)) {
        // larger FP64 blocks: DEG lanes share a pair (its registers fit the FP64 switch's
        // budget; in FP32 the case would cost the switch occupancy, so those run alone)
        apply_row_split<T, D, DEG, 256, 0>(a, x, y, row, sm);
    }
}

// (FP64: at least two waves per SIMD, the register budget that keeps its largest cases unspilled)
template <typename T>
__global__ void __launch_bounds__(256, sizeof(T) == 8 ? 2 : 1)
    k_apply_batch(const KArgs<T>* __restrict__ args, const int2* __restrict__ items, int nitems,
                  const int* __restrict__ shape, BatchIO io) {
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

// ---------------------------------------------------------------- launchers
using LaunchFn = hipError_t (*)(const void* args, int nrows, size_t smem, hipStream_t s);

template <typename T, int D, int DEG, int NT, int USE_LDS, bool KEEPX>
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

// Workgroup sizes offered to the autotuner: 128/256/512 threads.
template <typename T, int D, int DEG>
LaunchFn pick_reg(int nt, int lds) {
    if constexpr (reg_ok<T, D, DEG>()) {
        constexpr bool KX = keep_x<T, D, DEG>();
        if (lds == 2 && D < 2) return nullptr;
        switch (nt) {
            case 128:
                return lds == 2 ? &launch_reg<T, D, DEG, 128, 2, KX>
                                : lds ? &launch_reg<T, D, DEG, 128, 1, KX> : &launch_reg<T, D, DEG, 128, 0, KX>;
            case 256:
                return lds == 2 ? &launch_reg<T, D, DEG, 256, 2, KX>
                                : lds ? &launch_reg<T, D, DEG, 256, 1, KX> : &launch_reg<T, D, DEG, 256, 0, KX>;
            case 512:
                return lds == 2 ? &launch_reg<T, D, DEG, 512, 2, KX>
                                : lds ? &launch_reg<T, D, DEG, 512, 1, KX> : &launch_reg<T, D, DEG, 512, 0, KX>;
            default: return nullptr;
        }
    } else {
        (void)nt;
        (void)lds;
        return nullptr;
    }
}

template <typename T, int D>
LaunchFn pick_deg(int deg, int nt, int lds) {
    switch (deg) {
        case 1: return pick_reg<T, D, 1>(nt, lds);
        case 2: return pick_reg<T, D, 2>(nt, lds);
        case 3: return pick_reg<T, D, 3>(nt, lds);
        case 4: return pick_reg<T, D, 4>(nt, lds);
        case 5: return pick_reg<T, D, 5>(nt, lds);
        case 6: return pick_reg<T, D, 6>(nt, lds);
        case 7: return pick_reg<T, D, 7>(nt, lds);
        case 8: return pick_reg<T, D, 8>(nt, lds);
        default: return nullptr;
    }
}

template <typename T>
LaunchFn pick_impl(int D, int deg, int nt, int lds) {
    switch (D) {
        case 1: return pick_deg<T, 1>(deg, nt, lds);
        case 2: return pick_deg<T, 2>(deg, nt, lds);
        case 3: return pick_deg<T, 3>(deg, nt, lds);
        case 4: return pick_deg<T, 4>(deg, nt, lds);
        case 5: return pick_deg<T, 5>(deg, nt, lds);
        case 6: return pick_deg<T, 6>(deg, nt, lds);
        default: return nullptr;
    }
}

template <typename T, int D, int DEG, int NT, int USE_LDS>
hipError_t launch_split(const void* args, int nrows, size_t smem, hipStream_t s) {
    auto kern = k_apply_split<T, D, DEG, NT, USE_LDS>;
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nrows), dim3(NT), smem, s, *reinterpret_cast<const KArgs<T>*>(args));
    return hipGetLastError();
}

template <typename T, int D, int DEG>
LaunchFn pick_split_shape(int nt, bool lds) {
    switch (nt) {
        case 256: return lds ? &launch_split<T, D, DEG, 256, true> : &launch_split<T, D, DEG, 256, false>;
        case 512: return lds ? &launch_split<T, D, DEG, 512, true> : &launch_split<T, D, DEG, 512, false>;
        default: return nullptr;
    }
}

// shapes with 32-64 value blocks split over 2 or 4 lanes
template <typename T>
LaunchFn pick_split_impl(int D, int deg, int nt, bool lds) {
    if (D == 6 && deg == 2) return pick_split_shape<T, 6, 2>(nt, lds);
    if (D == 5 && deg == 2) return pick_split_shape<T, 5, 2>(nt, lds);
    if (D == 3 && deg == 4) return pick_split_shape<T, 3, 4>(nt, lds);
    return nullptr;
}

// per-precision entry points (instantiated in kmx_apply_f64.hip / _f32.hip)
LaunchFn pick_f64(int D, int deg, int nt, int lds);
LaunchFn pick_f32(int D, int deg, int nt, int lds);
LaunchFn pick_split_f64(int D, int deg, int nt, bool lds);
LaunchFn pick_split_f32(int D, int deg, int nt, bool lds);
hipError_t launch_generic_f64(const KArgs<double>& a, int D, int deg, int nrows, size_t smem, hipStream_t s);
hipError_t launch_generic_f32(const KArgs<float>& a, int D, int deg, int nrows, size_t smem, hipStream_t s);
hipError_t launch_batch_f64(const KArgs<double>* args, const int2* items, int nitems, const int* shape,
                            const BatchIO& io, int grid, size_t smem, hipStream_t s);
hipError_t launch_batch_f32(const KArgs<float>* args, const int2* items, int nitems, const int* shape,
                            const BatchIO& io, int grid, size_t smem, hipStream_t s);
int batch_occupancy_f64(size_t smem);
int batch_occupancy_f32(size_t smem);

}  // namespace kmx
