// kmx_sweep.hip -- the operator apply as dimension sweeps that stay on the sparse grid.
//
// apply_A.m:100-237 sums, for every target element and every connected source
// element, (A_{t,1}[I_1,J_1] x ... x A_{t,d}[I_d,J_d]) x_j.  When the pair
// list is, between every pair of level tuples, exactly the tensor product of
// per-dimension 1D connections (verified at plan time, as connect_nD.m:63-106
// builds it; an all-pairs list trivially) and the set of level tuples is
// downward closed (every sparse or full grid is), the sum factorizes by
// dimension.  Each 1D operator splits by level into a "down" part (target
// level <= source level) and an "up" part (target level > source level).
// Expanding the product over these parts and applying, for every choice, all
// down parts before any up part keeps every intermediate on the grid's own
// level tuples: a down step lowers a level below the source's, an up step
// raises one towards the target's.  Sharing the common work of the choices
// gives the recursion
//
//   F(z, d)   = F(Down_d z, d+1) + Up_d(F(z, d+1))     (d < D-1)
//   F(z, D-1) = A_{D-1} z                              (the whole 1D operator)
//
// run breadth-first: D-1 down launches (branch b of depth d reads slot b and
// writes slot b | 2^d), one launch of whole operators on the 2^(D-1) branches,
// and D-1 up launches that halve the branches and add, the last of them
// summing the terms and writing y.  Same terms as the pair-by-pair sum, added
// in another order (22x fewer flops on a 3D level-7 sparse grid, degree 3).
//
// Data: every intermediate is a set of dense boxes, one per level tuple L,
// holding prod_k deg*cells(L_k) values with dimension 1 slowest.  An op on
// dimension d groups the boxes that agree off dimension d (rest tuple R):
// their dimension-d positions, levels 0 to m(R) in order, are exactly the
// hierarchical 1D positions 0 .. deg*2^m - 1, so the op is a GEMM with the
// leading block of the masked 1D matrix, rows and columns cut by level for the
// down and up parts.  Output tiles of 16x16 run on one wave each (or on up to
// four waves splitting a long K range, summed in a fixed order) on the matrix
// cores (16x16x4 FP64 / FP32); every output is written by exactly one wave.
// Everything a wave needs about its tile (row bases and levels of the output
// rows, the box offset of every source level) is precomputed into a 512-byte
// work record that the wave reads in one coalesced load.
#include "kmx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace kmx {
namespace {

typedef double sw_d4 __attribute__((ext_vector_type(4)));
typedef float sw_f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct SwMfma;
template <>
struct SwMfma<double> {
    using V = sw_d4;
    static __device__ __forceinline__ V run(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct SwMfma<float> {
    using V = sw_f4;
    static __device__ __forceinline__ V run(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// BF16 values are kept as their bit patterns (the upper half of the FP32 pattern).
typedef unsigned short bf16_bits;
typedef short sw_s4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN stays a quiet NaN).
__device__ __forceinline__ short sw_to_bf16(float f) {
    unsigned u = __builtin_bit_cast(unsigned, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return short((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return short(u >> 16);
}

// Four consecutive steps of a separable tile: lane (li, lk) holds A(row li,
// position lk of step j) in av[j] and B(position lk of step j, column li) in bv[j],
// both already zero where masked.  FP64 / FP32: one 16x16x4 product per step.
// BF16 storage (TA = bf16_bits, T = float): the four steps are one 16x16x16 BF16
// product with FP32 accumulation -- lane (li, lk) feeds k = 4 lk + j of row li and
// of column li, so A and B pair up as in the per-step form and the result rows
// are the same (D[4 (lane / 16) + r][lane % 16]).
template <typename T, typename TA>
struct SepMma {
    using V = typename SwMfma<T>::V;
    static __device__ __forceinline__ V run4(const TA* av, const T* bv, V acc) {
#pragma unroll
        for (int j = 0; j < 4; ++j) acc = SwMfma<T>::run(av[j], bv[j], acc);
        return acc;
    }
};
template <>
struct SepMma<float, bf16_bits> {
    using V = sw_f4;
    static __device__ __forceinline__ V run4(const bf16_bits* av, const float* bv, V acc) {
        sw_s4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            a[j] = short(av[j]);
            b[j] = sw_to_bf16(bv[j]);
        }
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
};

// Loads and stores through global-address-space pointers: every buffer of these
// kernels is device memory, and generic pointers would compile to flat accesses
// (which also occupy the LDS counter and check the LDS aperture).
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
    *(__attribute__((address_space(1))) T*)p = v;
}
__device__ __forceinline__ int2 gld(const int2* p) {
    typedef int v2 __attribute__((ext_vector_type(2)));
    const v2 v = *(const __attribute__((address_space(1))) v2*)p;
    return make_int2(v.x, v.y);
}
__device__ __forceinline__ int4 gld(const int4* p) {
    typedef int v4 __attribute__((ext_vector_type(4)));
    const v4 v = *(const __attribute__((address_space(1))) v4*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}

// Four consecutive values of one lane, in one vector load (the box values of a step
// are only element-aligned: gfx950 loads unaligned vectors)
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
    typedef float v __attribute__((ext_vector_type(4), aligned(4)));
};
template <>
struct Vec4<double> {
    typedef double v __attribute__((ext_vector_type(4), aligned(8)));
};
template <>
struct Vec4<unsigned short> {
    typedef unsigned short v __attribute__((ext_vector_type(4), aligned(2)));
};
template <typename T>
__device__ __forceinline__ void gld4(const T* p, T out[4]) {
    typedef typename Vec4<T>::v V;
    const V x = *(const __attribute__((address_space(1))) V*)p;
    out[0] = x[0];
    out[1] = x[1];
    out[2] = x[2];
    out[3] = x[3];
}
// four values at an address aligned to four of them (the RK3 stage workspace)
template <typename T>
struct Vec4A {
    typedef T v __attribute__((ext_vector_type(4), aligned(4 * sizeof(T))));
};
template <typename T>
__device__ __forceinline__ void gld4a(const T* p, T out[4]) {
    typedef typename Vec4A<T>::v V;
    const V x = *(const __attribute__((address_space(1))) V*)p;
    out[0] = x[0];
    out[1] = x[1];
    out[2] = x[2];
    out[3] = x[3];
}
template <typename T>
__device__ __forceinline__ void gst4a(T* p, const T v[4]) {
    typedef typename Vec4A<T>::v V;
    V x;
    x[0] = v[0];
    x[1] = v[1];
    x[2] = v[2];
    x[3] = v[3];
    *(__attribute__((address_space(1))) V*)p = x;
}

// positions of level l along a dimension: deg * cells(l)
__device__ __forceinline__ int deg_cells(int deg, int l) { return deg * (l ? 1 << (l - 1) : 1); }

// UD_SEP: one dimension's op of a separable sweep: the whole operators of its terms,
// summed, from the box copy of x (slot in_base) into the dimension's own slot
// (out_base); the slots are summed into y afterwards.
enum { UD_DOWN = 0, UD_FULL = 1, UD_UP = 2, UD_SEP = 3 };

// Work record of one wave (UD_REC ints, stored so that lane j reads ints j and
// 64 + j as one int2).  The tile: rows [m0, m0+16) of the group's positions
// along d, columns [f0, f0+16) of its (o, i) pairs (O x I), branch b, output
// term tout; its K range: source levels [la, lb], sper steps of four
// positions per term.  This wave's share: steps [slo, ...) of term tlo up to
// step shi of term tlo + tn - 1.  The gsize waves from wave gbase of the
// workgroup hold the same tile (gsize 0: empty slot).  Per output row i:
// box_l + p * I, nl * I and level l (-1 past the group); per source level l:
// box offset.
enum {
    R_M0, R_F0, R_B, R_TOUT, R_O, R_I, R_LA, R_LB, R_SPER, R_GSIZE, R_GBASE, R_MAGIC, R_TLO, R_SLO, R_TN, R_SHI,
    R_ROWBASE = 16,
    R_ROWNL = 32,
    R_ROWLEV = 48,
    UD_BOX = 64,
    UD_REC = 128
};
// a separable op's record: its panel's runs in the plan's run table instead of a level range
// (R_RUN0 / R_NRUN for the host's checks; the kernels read the panel's steps from
// R_STEP0, its first step inside the op's storage)
enum { R_RUN0 = R_LA, R_NRUN = R_LB, R_NCT = R_TLO, R_STEP0 = R_TOUT };
// a separable tile whose panel has at most SEP_KMAX steps may sweep R_NCT column
// tiles of its group with the panel's A values held in registers
// (8: the packed body's registers set the launch's occupancy -- 92 VGPRs at 16
// steps, 68 at 8; an FP32 RK3 step of the 64-instance batch took 0.297 ms at 16,
// 0.288 at 12, 0.279 at 8 and 4)
constexpr int SEP_KMAX = 8;
// zero steps after the last of a separable storage (a tile's last load batch, or a
// packed tile's SEP_KMAX steps, reads up to 7 steps past its own)
constexpr int SEP_PAD = 8;
// A step's code: level l (bits 0-4), max(l - 1, 0) (bits 5-9: the shift that turns
// a column-o offset at level 1 into level l's), valid positions (bits 10-12; 0 for
// no step), first position inside the level (bits 13-30).
__host__ __device__ constexpr int sep_code(int l, int valid, int pos) {
    return l | (l > 1 ? l - 1 : 0) << 5 | valid << 10 | pos << 13;
}
constexpr int UD_MAXLEV = UD_REC - UD_BOX - 1;  // highest level a record holds a box for
// Waves per workgroup.  Two: a workgroup holds its slots until its slowest wave
// ends (the split tiles' partial sums meet at a barrier), so small workgroups keep
// the CUs full; on the 64-instance batch an FP32 RK3 step took 0.372 ms with eight
// waves, 0.319 with four, 0.305 with two and 0.336 with one (no tile split).
constexpr int UD_WAVES = 2;

// Slot b of term t of an intermediate starts at base + (b * T + t) * nb; slot 0
// of an input with x0 set is x itself, read through the box map.  part: which
// source levels a row takes (UD_DOWN / UD_FULL / UD_UP, the last adding slot
// b | add_bit); final: sum the terms, scaled, into y.
struct UdOp {
    int d, deg, lda, T;
    int part, final, x0, out_bit, add_bit;
    long long nb;
    long long in_base, out_base, add_base;
    long long a_off[KMX_MAX_TERMS];  // column-major masked matrix of (term, d)
    double scale[KMX_MAX_TERMS];
};

// Steps per load batch: one A value and NBR B values per step stay in registers (FP64
// values take two).  Tiles are split over waves by the FP64 batch, so FP32 waves finish
// theirs in fewer batches.
template <typename T, int NBR>
constexpr int ud_batch() {
    return sizeof(T) == 4 || NBR == 1 ? 8 : NBR == 2 ? 6 : 4;
}
int ud_split_batch(int nbr) {
    return nbr == 1 ? ud_batch<double, 1>() : nbr == 2 ? ud_batch<double, 2>() : ud_batch<double, 4>();
}

// One wave: one output tile of NBR consecutive branches (same group, rows, columns and
// source range; only their slots differ), so the A operand, the level cursor and the
// offsets are shared.  Lane (li, lk) feeds row m0 + li of A and column f0 + li of B at
// position 4 step + lk of the current level; the cursor is wave-uniform.  rv: this
// lane's two ints of the wave's work record; part: LDS for the partial tiles of waves
// 1.. of a split tile, (UD_WAVES - 1) x NBR x 4 x 64 values.  Every wave of the
// workgroup calls it (it holds one workgroup barrier).  Down, whole and up parts of
// the general sweep; a separable op has its own body (sep_tile).
template <typename T, int PART, bool FINAL, int NBR>
__device__ __forceinline__ void ud_tile(const T* __restrict__ mat, T* __restrict__ arena, const int2 rv, const int w,
                                        const int lane, const UdOp& op, const double* __restrict__ x,
                                        const int* __restrict__ xmap, double* __restrict__ y,
                                        const int* __restrict__ ymap, T* __restrict__ part) {
    using V = typename SwMfma<T>::V;
    constexpr int NB = ud_batch<T, NBR>();
    const int gsize = __builtin_amdgcn_readlane(rv.x, R_GSIZE);
    const int gbase = __builtin_amdgcn_readlane(rv.x, R_GBASE);
    const bool leader = gsize > 0 && w == gbase;
    V acc[NBR];
    T addv[NBR][4];
#pragma unroll
    for (int j = 0; j < NBR; ++j) {
        acc[j] = V{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int q = 0; q < 4; ++q) addv[j][q] = T(0);
    }
    int oaddr[4] = {0, 0, 0, 0};
    bool ook[4] = {false, false, false, false};
    int yi[4] = {-1, -1, -1, -1};
    const int b = __builtin_amdgcn_readlane(rv.x, R_B);  // first branch of the tile
    const int tout = __builtin_amdgcn_readlane(rv.x, R_TOUT);
    if (gsize > 0) {
        const int m0 = __builtin_amdgcn_readlane(rv.x, R_M0), f0 = __builtin_amdgcn_readlane(rv.x, R_F0);
        const int O = __builtin_amdgcn_readlane(rv.x, R_O), I = __builtin_amdgcn_readlane(rv.x, R_I);
        const int li = lane & 15, lk = lane >> 4;
        const int lr = __shfl(rv.x, R_ROWLEV + li);  // level of A row m0 + li (-1: past the group)
        const bool rok = lr >= 0;
        const int r0 = rok ? m0 + li : 0;
        const int f = f0 + li;
        const bool fok = f < O * I;
        const int fc = fok ? f : 0;
        const unsigned magic = unsigned(__builtin_amdgcn_readlane(rv.x, R_MAGIC));
        const int fo = I == 1 ? fc : magic ? int(__umulhi(unsigned(fc), magic)) : fc / I;
        const int fi = fc - fo * I;

        if (leader) {  // output addresses of this lane's four rows; addends and y indices fetched now
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = SwMfma<T>::row(lane, q);
                const int rb = __shfl(rv.x, R_ROWBASE + row), rn = __shfl(rv.x, R_ROWNL + row);
                const int rl = __shfl(rv.x, R_ROWLEV + row);
                ook[q] = fok && rl >= 0;
                oaddr[q] = ook[q] ? rb + fo * rn + fi : 0;
            }
            if (PART == UD_UP) {
                const int ta = FINAL ? 0 : tout, tb = FINAL ? op.T : tout + 1;
                for (int t = ta; t < tb; ++t)
#pragma unroll
                    for (int j = 0; j < NBR; ++j) {
                        const T* add = arena + op.add_base + ((long long)((b + j) | op.add_bit) * op.T + t) * op.nb;
                        const T sc = FINAL ? T(op.scale[t]) : T(1);
#pragma unroll
                        for (int q = 0; q < 4; ++q) addv[j][q] += sc * add[oaddr[q]];
                    }
            }
            if (FINAL)
#pragma unroll
                for (int q = 0; q < 4; ++q) yi[q] = ymap[oaddr[q]];
        }

        const int la = __builtin_amdgcn_readlane(rv.x, R_LA);
        const int sper = __builtin_amdgcn_readlane(rv.x, R_SPER), tlo = __builtin_amdgcn_readlane(rv.x, R_TLO);
        const int slo = __builtin_amdgcn_readlane(rv.x, R_SLO), tn = __builtin_amdgcn_readlane(rv.x, R_TN);
        const int shi = __builtin_amdgcn_readlane(rv.x, R_SHI);
        const bool xin0 = op.x0 && b == 0;  // branch b (j = 0) is x, read through the box map
        for (int k = 0; k < tn; ++k) {
            const int t = tlo + k;
            const int s0 = k == 0 ? slo : 0, s1 = k == tn - 1 ? shi : sper;
            const T* abase = mat + op.a_off[t];
            const T* bbase[NBR];
#pragma unroll
            for (int j = 0; j < NBR; ++j)
                bbase[j] = arena + ((j == 0 && xin0) ? 0 : op.in_base + ((long long)(b + j) * op.T + t) * op.nb);
            const T sc = FINAL ? T(op.scale[t]) : T(1);
            int l = la, rem = s0;  // level and step of step s0 (wave-uniform)
            for (;;) {
                const int ns = (op.deg * (l ? 1 << (l - 1) : 1) + 3) >> 2;
                if (rem < ns) break;
                rem -= ns;
                ++l;
            }
            int p = 4 * rem;
            int nl = 0, aidx = 0, bidx = 0;
            bool aon = false, newlev = true;
            for (int q = s0; q < s1; q += NB) {
                // 1. per-lane offsets of the batch (0 where masked); the cursor moves on the scalar unit
                int ai[NB], bi[NB];
                bool ma[NB], mb[NB];
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    if (newlev) {
                        const int st = l ? 1 << (l - 1) : 0;
                        nl = op.deg * (l ? st : 1);
                        const int box = __builtin_amdgcn_readlane(rv.y, min(l, UD_MAXLEV));
                        aon = rok && (PART == UD_FULL || (PART == UD_DOWN ? lr <= l : lr > l));
                        aidx = r0 + (op.deg * st + p + lk) * op.lda;
                        bidx = box + (fo * nl + p + lk) * I + fi;
                        newlev = false;
                    }
                    const bool on = q + c < s1;
                    const bool pin = on && p + lk < nl;
                    ai[c] = pin ? aidx : 0;
                    bi[c] = pin ? bidx : 0;
                    ma[c] = pin && aon;  // A element (row li, position): row and level masks
                    mb[c] = pin && fok;  // B element (position, column li)
                    if (on) {
                        p += 4;
                        aidx += 4 * op.lda;
                        bidx += 4 * I;
                        if (p >= nl) {
                            p = 0;
                            ++l;
                            newlev = true;
                        }
                    }
                }
                // 2. every load of the batch in flight at once
                T av[NB], bv[NBR][NB];
#pragma unroll
                for (int c = 0; c < NB; ++c) av[c] = abase[ai[c]];
#pragma unroll
                for (int j = 0; j < NBR; ++j) {
                    if (j == 0 && xin0) {
                        int xm[NB];
#pragma unroll
                        for (int c = 0; c < NB; ++c) xm[c] = xmap[bi[c]];
#pragma unroll
                        for (int c = 0; c < NB; ++c) {
                            const double v = x[xm[c] >= 0 ? xm[c] : 0];
                            bv[0][c] = xm[c] >= 0 ? T(v) : T(0);  // absent cell: x is 0 there
                        }
                    } else {
#pragma unroll
                        for (int c = 0; c < NB; ++c) bv[j][c] = bbase[j][bi[c]];
                    }
                }
                // 3. matrix operations (masked steps feed zeros)
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const T a = ma[c] ? av[c] * sc : T(0);
#pragma unroll
                    for (int j = 0; j < NBR; ++j) acc[j] = SwMfma<T>::run(a, mb[c] ? bv[j][c] : T(0), acc[j]);
                }
            }
        }
        if (gsize > 1 && !leader)
#pragma unroll
            for (int j = 0; j < NBR; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) part[(((w - 1) * NBR + j) * 4 + q) * 64 + lane] = acc[j][q];
    }
    __syncthreads();  // every wave of every workgroup, split tiles or not
    if (!leader) return;
    for (int v = 1; v < gsize; ++v)  // the tile's partial sums in wave order
#pragma unroll
        for (int j = 0; j < NBR; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[j][q] += part[(((w + v - 1) * NBR + j) * 4 + q) * 64 + lane];
    if (FINAL) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ook[q] && yi[q] >= 0) y[yi[q]] = double(acc[0][q] + addv[0][q]);
    } else {
#pragma unroll
        for (int j = 0; j < NBR; ++j) {
            T* out = arena + op.out_base + ((long long)((b + j) | op.out_bit) * op.T + tout) * op.nb;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (ook[q]) out[oaddr[q]] = acc[j][q] + addv[j][q];
        }
    }
}

// One op of one plan: UD_WAVES work records per workgroup.
template <typename T, int PART, bool FINAL, int NBR>
__global__ void __launch_bounds__(UD_WAVES * 64) k_ud_op(const T* __restrict__ mat, T* __restrict__ arena,
                                                          const int2* __restrict__ recs, UdOp op,
                                                          const double* __restrict__ x, const int* __restrict__ xmap,
                                                          double* __restrict__ y, const int* __restrict__ ymap) {
    __shared__ T part[(UD_WAVES - 1) * NBR * 4 * 64];
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int2 rv = recs[((long long)blockIdx.x * UD_WAVES + w) * (UD_REC / 2) + lane];
    ud_tile<T, PART, FINAL, NBR>(mat, arena, rv, w, lane, op, x, xmap, y, ymap, part);
}

// Box copies of x and slot sums of every separable instance of a batch, one
// launch each: job j covers workgroups [blk0_j, blk0_{j+1}), 256 positions each.
// Every workgroup reads its own copy of its job (one load, no index to chase).
struct BoxJob {
    const int* xmap;
    const int* ymap;
    void* arena[2];  // FP64, FP32 arena of the plan: slot 0 x's box copy, slots 1..nslot the ops' sums
    long long nb;
    long long koff;  // the instance's first box value in the batch's RK3 stage workspace
    int nslot, inst, blk0;
};
// x / y of a batch of one plan: a 16-byte kernel argument instead of BatchIO's 2 KB
// (a lone plan's apply is bound by its launches)
struct IoOne {
    const double* x;
    double* y;
};
__device__ __forceinline__ const double* io_x(const BatchIO& io, int i) { return io.x[i]; }
__device__ __forceinline__ double* io_y(const BatchIO& io, int i) { return io.y[i]; }
__device__ __forceinline__ const double* io_x(const IoOne& io, int) { return io.x; }
__device__ __forceinline__ double* io_y(const IoOne& io, int) { return io.y; }

// The block of the box list a workgroup of the gather / slot-sum launches takes:
// runs of BOX_XCD_RUN consecutive blocks on one XCD (workgroup b runs on XCD b % 8;
// in each span of 8 BOX_XCD_RUN workgroups, XCD y takes blocks y BOX_XCD_RUN ..
// (y + 1) BOX_XCD_RUN - 1; a last partial span in order).  Box order splits an
// element's x / y values into runs of deg, the rest of whose lines the next blocks
// read: on one XCD they come from its L2.
constexpr int BOX_XCD_RUN = 16;
__device__ __forceinline__ int box_block() {
    const int b = int(blockIdx.x), span = 8 * BOX_XCD_RUN;
    const int base = b - b % span;
    if (base + span > int(gridDim.x)) return b;
    return base + (b % 8) * BOX_XCD_RUN + (b % span) / 8;
}

template <typename T, typename IO>
__global__ void __launch_bounds__(256) k_box_gather_batch(const BoxJob* __restrict__ wg_jobs, IO io) {
    const int blk = box_block();
    const BoxJob& j = wg_jobs[blk];
    const long long i = (long long)(blk - j.blk0) * 256 + threadIdx.x;
    if (i < j.nb) {
        const int m = gld(j.xmap + i);
        T* xb = static_cast<T*>(j.arena[sizeof(T) == 8 ? 0 : 1]);
        gst(xb + i, m >= 0 ? T(gld(io_x(io, j.inst) + m)) : T(0));
    }
}
// y = the ops' slots summed (in op order), at the present cells
// (the slots are read before the map, so both loads are in flight together)
template <typename T, typename IO>
__global__ void __launch_bounds__(256) k_box_sum_batch(const BoxJob* __restrict__ wg_jobs, IO io) {
    const int blk = box_block();
    const BoxJob& j = wg_jobs[blk];
    const long long i = (long long)(blk - j.blk0) * 256 + threadIdx.x;
    if (i < j.nb) {
        const T* sl = static_cast<const T*>(j.arena[sizeof(T) == 8 ? 0 : 1]) + j.nb + i;
        double v = 0.0;
        for (int k = 0; k < j.nslot; ++k) v += double(gld(sl + (long long)k * j.nb));
        const int m = gld(j.ymap + i);
        if (m >= 0) gst(io_y(io, j.inst) + m, v);
    }
}

// Explicit RK3 (time_advance.m:151-204: a21 = 1/2, a31 = -1, a32 = 2,
// b = 1/6, 2/3, 1/6) fused around the merged separable sweeps.  Per stage one
// launch of the ops and one of this kernel, which, at every present cell of every
// instance, sums the ops' slots (dimension order, FP64) into the stage derivative
// k_s and writes what the next launch reads: stage 1 k1 and the box copy of
// f + dt a21 k1, stage 2 k2 and the box copy of f + dt (a31 k1 + a32 k2), stage 3
// the new state f + dt (b1 k1 + b2 k2 + b3 k3) and (next_box) its box copy for the
// next step; stage 0 only copies f into the box.  Absent cells of the box copy are
// never written (zero since the plan's set-up).  During one call the state itself
// is kept in box order too (fb, FP64): stage 0 reads the caller's state in, every
// later stage reads and writes it coalesced with the slots, and the last step's
// stage 3 writes the caller's state back; present cells are flagged by one byte
// each.  k1 and k2 are kept in box order in T, the precision the ops computed them
// in; the combinations are FP64.
constexpr double RK_A21 = 0.5, RK_A31 = -1.0, RK_A32 = 2.0;
constexpr double RK_B1 = 1.0 / 6.0, RK_B2 = 2.0 / 3.0, RK_B3 = 1.0 / 6.0;
struct Rk3Args {
    double* f[KMX_BATCH_MAX];
    double dt[KMX_BATCH_MAX];
    void* k1;  // stage workspace (T), the instance's box values at koff
    void* k2;
    double* fb;                   // the state in box order (FP64), at koff
    const unsigned char* present;  // 1 at the present cells, at koff
};
// stages 1..3 at box value i of job j (a present cell, or nothing is written)
template <typename T, int STAGE>
__device__ __forceinline__ void rk3_cell(const BoxJob& j, const Rk3Args& a, int next_box, long long i) {
    T* box = static_cast<T*>(j.arena[sizeof(T) == 8 ? 0 : 1]);
    double* fb = a.fb + j.koff;
    double k = 0.0;  // the slots first, so their loads and the flag's are in flight together
    const T* sl = box + j.nb + i;
    for (int q = 0; q < j.nslot; ++q) k += double(gld(sl + (long long)q * j.nb));
    if (!gld(a.present + j.koff + i)) return;
    const double fm = gld(fb + i);
    const double dt = a.dt[j.inst];
    T* k1 = static_cast<T*>(a.k1) + j.koff;
    T* k2 = static_cast<T*>(a.k2) + j.koff;
    if (STAGE == 1) {
        gst(k1 + i, T(k));
        gst(box + i, T(fm + dt * RK_A21 * k));
    } else if (STAGE == 2) {
        gst(k2 + i, T(k));
        gst(box + i, T(fm + dt * (RK_A31 * double(gld(k1 + i)) + RK_A32 * k)));
    } else {
        const double fn = fm + dt * (RK_B1 * double(gld(k1 + i)) + RK_B2 * double(gld(k2 + i)) + RK_B3 * k);
        if (next_box) {  // another step follows: the state stays in box order
            gst(fb + i, fn);
            gst(box + i, T(fn));
        } else {  // the call's last step: the caller's state
            gst(a.f[j.inst] + gld(j.ymap + i), fn);
        }
    }
}

template <typename T, int STAGE>
__global__ void __launch_bounds__(256) k_box_rk3(const BoxJob* __restrict__ wg_jobs, Rk3Args a, int next_box) {
    const BoxJob& j = wg_jobs[blockIdx.x];
    const long long i = (long long)(int(blockIdx.x) - j.blk0) * 256 + threadIdx.x;
    if (i >= j.nb) return;
    if (STAGE == 0) {  // the caller's state in, and its box copy
        T* box = static_cast<T*>(j.arena[sizeof(T) == 8 ? 0 : 1]);
        const int m = gld(j.ymap + i);
        if (m < 0) return;
        const double fm = gld(a.f[j.inst] + m);
        gst(a.fb + j.koff + i, fm);
        gst(box + i, T(fm));
        return;
    }
    rk3_cell<T, STAGE>(j, a, next_box, i);
}

// Stages 1..3 four box values per thread (jobs in blocks of 1024 values): the
// stage workspace is aligned to four values (every instance's koff is a multiple
// of four), so k1, k2, the box-order state, the box copy and four present flags
// move as one vector each; the slots (at multiples of nb) as unaligned vectors.
// The arithmetic of every value is rk3_cell's; absent cells of the box copy are
// never written (k1, k2 and the box-order state at absent cells are, and never
// read).  An instance's last 1..3 values go one at a time.
template <typename T, int STAGE>
__global__ void __launch_bounds__(256) k_box_rk3v(const BoxJob* __restrict__ wg_jobs, Rk3Args a, int next_box) {
    static_assert(STAGE >= 1 && STAGE <= 3, "stages 1..3");
    const BoxJob& j = wg_jobs[blockIdx.x];
    const long long i0 = ((long long)(int(blockIdx.x) - j.blk0) * 256 + threadIdx.x) * 4;
    if (i0 >= j.nb) return;
    if (i0 + 4 > j.nb) {
        for (long long i = i0; i < j.nb; ++i) rk3_cell<T, STAGE>(j, a, next_box, i);
        return;
    }
    T* box = static_cast<T*>(j.arena[sizeof(T) == 8 ? 0 : 1]);
    double* fb = a.fb + j.koff;
    double k[4] = {0.0, 0.0, 0.0, 0.0};
    const T* sl = box + j.nb + i0;
    for (int q = 0; q < j.nslot; ++q) {
        T v[4];
        gld4(sl + (long long)q * j.nb, v);
#pragma unroll
        for (int e = 0; e < 4; ++e) k[e] += double(v[e]);
    }
    const unsigned pf = gld(reinterpret_cast<const unsigned*>(a.present + j.koff + i0));
    if (pf == 0u) return;
    double fm[4];
    gld4a(fb + i0, fm);
    const double dt = a.dt[j.inst];
    T* k1 = static_cast<T*>(a.k1) + j.koff;
    T* k2 = static_cast<T*>(a.k2) + j.koff;
    T kk[4], bx[4];
    if (STAGE == 1) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            kk[e] = T(k[e]);
            bx[e] = T(fm[e] + dt * RK_A21 * k[e]);
        }
        gst4a(k1 + i0, kk);
    } else if (STAGE == 2) {
        T k1v[4];
        gld4a(k1 + i0, k1v);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            kk[e] = T(k[e]);
            bx[e] = T(fm[e] + dt * (RK_A31 * double(k1v[e]) + RK_A32 * k[e]));
        }
        gst4a(k2 + i0, kk);
    } else {
        T k1v[4], k2v[4];
        gld4a(k1 + i0, k1v);
        gld4a(k2 + i0, k2v);
        double fn[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            fn[e] = fm[e] + dt * (RK_B1 * double(k1v[e]) + RK_B2 * double(k2v[e]) + RK_B3 * k[e]);
        if (!next_box) {  // the call's last step: the caller's state at the present cells
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if ((pf >> (8 * e)) & 0xffu) gst(a.f[j.inst] + gld(j.ymap + i0 + e), fn[e]);
            return;
        }
        gst4a(fb + i0, fn);
#pragma unroll
        for (int e = 0; e < 4; ++e) bx[e] = T(fn[e]);
    }
    if (pf == 0x01010101u) {
        gst4a(box + i0, bx);
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if ((pf >> (8 * e)) & 0xffu) gst(box + i0 + e, bx[e]);
    }
}

// The present-cell flags of every member (1 where the box maps to a value of y).
__global__ void __launch_bounds__(256) k_box_present(const BoxJob* __restrict__ wg_jobs, unsigned char* present) {
    const BoxJob& j = wg_jobs[blockIdx.x];
    const long long i = (long long)(int(blockIdx.x) - j.blk0) * 256 + threadIdx.x;
    if (i < j.nb) gst(present + j.koff + i, (unsigned char)(gld(j.ymap + i) >= 0 ? 1 : 0));
}

// Op kinds of the batched launch: (part, final, branches per tile).
enum {
    UK_DOWN1, UK_DOWN2, UK_DOWN4, UK_FULL1, UK_FULL2, UK_FULL4, UK_UP1, UK_UP2, UK_UP4, UK_FULL_FINAL, UK_UP_FINAL,
    UK_SEP
};
int ud_kind(int part, bool fin, int nbr) {
    const int nb = nbr == 1 ? 0 : nbr == 2 ? 1 : 2;
    if (part == UD_SEP) return UK_SEP;
    return part == UD_DOWN ? UK_DOWN1 + nb
           : part == UD_FULL ? (fin ? int(UK_FULL_FINAL) : UK_FULL1 + nb)
                             : (fin ? int(UK_UP_FINAL) : UK_UP1 + nb);
}
int ud_kind_nbr(int kind) {
    return kind >= UK_FULL_FINAL ? 1 : 1 << ((kind - UK_DOWN1) % 3);
}

// One op of one instance of a batch: its parameters (term scales as at batch
// set-up), its plan's matrices, arena and box maps per precision, and its slot
// in the batch's x / y pointer arrays.
struct UdBatchOp {
    UdOp a;
    const void* mat[3];  // FP64, FP32 blocks of the plan; BF16 values (separable ops)
    void* arena[2];
    const int* xmap;
    const int* ymap;
    const int* codes;  // separable ops: the plan's step codes (forward or transposed)
    int inst;
    int kind;
    int quad;  // separable op of the last dimension in the quad storage (sep_tile_quad)
};

template <typename T, typename TA, int NB>
__device__ __forceinline__ void sep_tile(const TA* __restrict__ mat, const int* __restrict__ codes,
                                         T* __restrict__ arena, const int2 rv, const int w, const int lane,
                                         const UdOp& op, T* __restrict__ part);
template <typename T, typename TA>
__device__ __forceinline__ void sep_tile_cols(const TA* __restrict__ mat, const int* __restrict__ codes,
                                              T* __restrict__ arena, const int2 rv, const int lane, const UdOp& op);
template <typename T, typename TA, int NG>
__device__ __forceinline__ void sep_tile_quad(const TA* __restrict__ mat, const int* __restrict__ codes,
                                              T* __restrict__ arena, const int2 rv, const int w, const int lane,
                                              const UdOp& op, T* __restrict__ part);
template <typename T, typename TA>
__device__ __forceinline__ void sep_tile_quad_cols(const TA* __restrict__ mat, const int* __restrict__ codes,
                                                   T* __restrict__ arena, const int2 rv, const int lane,
                                                   const UdOp& op);

// One phase of a batch: ops of many instances, each on whole workgroups
// (wg_op[workgroup] = its op); the workgroup's op kind picks the tile body.
// MAXNBR bounds the branches per tile of the phase (LDS for the partial tiles).
template <typename T, int MAXNBR>
__global__ void __launch_bounds__(UD_WAVES * 64) k_ud_batch(const UdBatchOp* __restrict__ ops,
                                                             const int* __restrict__ wg_op,
                                                             const int2* __restrict__ recs, BatchIO io) {
    __shared__ T part[(UD_WAVES - 1) * MAXNBR * 4 * 64];
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int2 rv = recs[((long long)blockIdx.x * UD_WAVES + w) * (UD_REC / 2) + lane];
    const UdBatchOp& o = ops[wg_op[blockIdx.x]];
    constexpr int sl = sizeof(T) == 8 ? 0 : 1;
    const T* mat = static_cast<const T*>(o.mat[sl]);
    T* arena = static_cast<T*>(o.arena[sl]);
    const double* x = io.x[o.inst];
    double* y = io.y[o.inst];
    switch (o.kind) {
        case UK_DOWN1: ud_tile<T, UD_DOWN, false, 1>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part); break;
        case UK_DOWN2: ud_tile<T, UD_DOWN, false, 2>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part); break;
        case UK_FULL1: ud_tile<T, UD_FULL, false, 1>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part); break;
        case UK_FULL2: ud_tile<T, UD_FULL, false, 2>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part); break;
        case UK_UP1: ud_tile<T, UD_UP, false, 1>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part); break;
        case UK_UP2: ud_tile<T, UD_UP, false, 2>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part); break;
        case UK_FULL_FINAL:
            ud_tile<T, UD_FULL, true, 1>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part);
            break;
        case UK_UP_FINAL: ud_tile<T, UD_UP, true, 1>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part); break;
        case UK_SEP:
            if (o.quad && __builtin_amdgcn_readlane(rv.x, R_GSIZE) > 0 && __builtin_amdgcn_readlane(rv.x, R_NCT) > 1)
                sep_tile_quad_cols<T, T>(mat, o.codes, arena, rv, lane, o.a);
            else if (o.quad)
                sep_tile_quad<T, T, 2>(mat, o.codes, arena, rv, w, lane, o.a, part);
            else if (__builtin_amdgcn_readlane(rv.x, R_GSIZE) > 0 && __builtin_amdgcn_readlane(rv.x, R_NCT) > 1)
                sep_tile_cols<T, T>(mat, o.codes, arena, rv, lane, o.a);
            else
                sep_tile<T, T, 8>(mat, o.codes, arena, rv, w, lane, o.a, part);
            break;
        default:
            if constexpr (MAXNBR == 4) {
                if (o.kind == UK_DOWN4)
                    ud_tile<T, UD_DOWN, false, 4>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part);
                else if (o.kind == UK_FULL4)
                    ud_tile<T, UD_FULL, false, 4>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part);
                else if (o.kind == UK_UP4)
                    ud_tile<T, UD_UP, false, 4>(mat, arena, rv, w, lane, o.a, x, o.xmap, y, o.ymap, part);
            }
            break;
    }
}

// One wave of a separable op (UD_SEP): one 16 x 16 output tile of one group (rows
// m0 .. m0 + 15 of the group's positions along the op dimension, columns f0 .. f0 + 15
// of its O x I columns) from the compact row-panel storage of the dimension's
// combined operator (its terms summed, scales included).  The storage keeps, per
// 16-row panel, only the steps of four source positions that meet a kept block,
// consecutive from the panel's first step (record R_STEP0, inside the op's
// storage), levels in order; the group's steps are the panel's first R_SPER.
// Element (row li, position 4 s + lk) of step s sits at offset 64 s + 16 lk + li =
// 64 s + lane, so each step is one contiguous 64-value read per wave; the step's
// code (level, valid positions, first position inside the level) gives the box
// offset of B at the lane's column: box[l] + (fo * nl + p + lk) * I + fi, where
// fo * nl * I = (fo * deg * I) << max(l - 1, 0).  The wave's share is steps
// [slo, shi); NB steps are loaded per batch, then multiplied.  Rows past the
// group read the panel's other rows (finite; never stored), rows past the matrix
// read the storage's zero padding; B is zero at positions past a step's valid ones
// and at the steps past shi of the last batch (their A is a step of the panel).
// TA: the storage type of A (T, or BF16 bit patterns with T = float; NB is a
// multiple of 4, the steps of one BF16 product).
template <typename T, typename TA, int NB>
__device__ __forceinline__ void sep_tile(const TA* __restrict__ mat, const int* __restrict__ codes,
                                         T* __restrict__ arena, const int2 rv, const int w, const int lane,
                                         const UdOp& op, T* __restrict__ part) {
    static_assert(NB % 4 == 0, "steps come in groups of four");
    using V = typename SwMfma<T>::V;
    const int gsize = __builtin_amdgcn_readlane(rv.x, R_GSIZE);
    const int gbase = __builtin_amdgcn_readlane(rv.x, R_GBASE);
    const bool leader = gsize > 0 && w == gbase;
    V acc = V{T(0), T(0), T(0), T(0)};
    int oaddr[4] = {0, 0, 0, 0};
    bool ook[4] = {false, false, false, false};
    if (gsize > 0) {
        const int f0 = __builtin_amdgcn_readlane(rv.x, R_F0);
        const int O = __builtin_amdgcn_readlane(rv.x, R_O), I = __builtin_amdgcn_readlane(rv.x, R_I);
        const int li = lane & 15, lk = lane >> 4;
        const int f = f0 + li;
        const bool fok = f < O * I;
        const int fc = fok ? f : 0;
        const unsigned magic = unsigned(__builtin_amdgcn_readlane(rv.x, R_MAGIC));
        const int fo = I == 1 ? fc : magic ? int(__umulhi(unsigned(fc), magic)) : fc / I;
        const int fi = fc - fo * I;
        if (leader)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = SwMfma<T>::row(lane, q);
                const int rb = __shfl(rv.x, R_ROWBASE + row), rn = __shfl(rv.x, R_ROWNL + row);
                const int rl = __shfl(rv.x, R_ROWLEV + row);
                ook[q] = fok && rl >= 0;
                oaddr[q] = ook[q] ? rb + fo * rn + fi : 0;
            }
        const int tn = __builtin_amdgcn_readlane(rv.x, R_TN);
        if (tn > 0) {
            const int slo = __builtin_amdgcn_readlane(rv.x, R_SLO), shi = __builtin_amdgcn_readlane(rv.x, R_SHI);
            const long long g0 = op.a_off[0] / 64 + __builtin_amdgcn_readlane(rv.x, R_STEP0) + slo;
            const TA* amat = mat + 64 * g0 + lane;  // the storage runs SEP_PAD steps past its last
            const int* cd = codes + g0;
            const T* bsrc = arena + op.in_base;
            int lane_off = lk * I + fi;          // the lane's part of a B offset
            asm("" : "+v"(lane_off));            // (kept whole: not refolded into (p + lk) * I)
            const int fo_dI = fo * op.deg * I;   // fo * nl * I at levels 0 and 1
            const int ns = shi - slo;
            for (int q0 = 0; q0 < ns; q0 += 64) {
                const int qn = min(64, ns - q0);
                const int cv = lane < qn ? gld(cd + q0 + lane) : 0;  // lane t: the code of step q0 + t
                for (int q = 0; q < qn; q += NB) {
                    const TA* ab = amat + 64 * (q0 + q);
                    TA av[NB];
                    T bv[NB];
#pragma unroll
                    for (int c = 0; c < NB; ++c) av[c] = gld(ab + 64 * c);
#pragma unroll
                    for (int c = 0; c < NB; ++c) {
                        const int code = __builtin_amdgcn_readlane(cv, q + c);
                        const int nv = (code >> 10) & 7;  // valid positions (0 past the wave's steps)
                        const int sb = __builtin_amdgcn_readlane(rv.y, code & 31) + (code >> 13) * I;
                        const unsigned ad = unsigned(sb + (fo_dI << ((code >> 5) & 31)) + lane_off);
                        const bool ok = lk < nv;
                        const T b = gld(bsrc + (ok ? ad : 0u));
                        bv[c] = ok ? b : T(0);
                    }
#pragma unroll
                    for (int c = 0; c < NB; c += 4) acc = SepMma<T, TA>::run4(av + c, bv + c, acc);
                }
            }
        }
        if (gsize > 1 && !leader)
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) part[((w - 1) * 4 + qq) * 64 + lane] = acc[qq];
    }
    __syncthreads();  // every wave of the workgroup
    if (!leader) return;
    for (int v = 1; v < gsize; ++v)  // the tile's partial sums in wave order
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) acc[qq] += part[((w + v - 1) * 4 + qq) * 64 + lane];
    T* out = arena + op.out_base;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
        if (ook[qq]) gst(out + oaddr[qq], acc[qq]);
}

// One wave of a separable op of the last dimension (inner extent I = 1, the op's
// "quad" storage): the tile of sep_tile with the matrix operations' k index taken
// as (step of a group of four, position inside the step) instead of (position,
// step).  A panel's steps come in groups of four (zero steps pad the panel); the
// storage keeps lane (li, q)'s values A(row li, positions 0..3 of step q) of a group
// together, so one vector load per lane brings a group's operator values, and at
// I = 1 a step's four positions are consecutive box values of one column, so one
// vector load per lane (column li, step q) brings its B values: operation j of the
// four multiplies position j of every step of the group.  The wave's share
// [slo, shi) starts at a group.
template <typename T, typename TA, int NG>
__device__ __forceinline__ void sep_tile_quad(const TA* __restrict__ mat, const int* __restrict__ codes,
                                              T* __restrict__ arena, const int2 rv, const int w, const int lane,
                                              const UdOp& op, T* __restrict__ part) {
    using V = typename SwMfma<T>::V;
    const int gsize = __builtin_amdgcn_readlane(rv.x, R_GSIZE);
    const int gbase = __builtin_amdgcn_readlane(rv.x, R_GBASE);
    const bool leader = gsize > 0 && w == gbase;
    V acc = V{T(0), T(0), T(0), T(0)};
    int oaddr[4] = {0, 0, 0, 0};
    bool ook[4] = {false, false, false, false};
    if (gsize > 0) {
        const int f0 = __builtin_amdgcn_readlane(rv.x, R_F0);
        const int O = __builtin_amdgcn_readlane(rv.x, R_O);
        const int li = lane & 15, lk = lane >> 4;
        const int f = f0 + li;
        const bool fok = f < O;
        const int fo = fok ? f : 0;  // I = 1: the column is o
        if (leader)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = SwMfma<T>::row(lane, q);
                const int rb = __shfl(rv.x, R_ROWBASE + row), rn = __shfl(rv.x, R_ROWNL + row);
                const int rl = __shfl(rv.x, R_ROWLEV + row);
                ook[q] = fok && rl >= 0;
                oaddr[q] = ook[q] ? rb + fo * rn : 0;
            }
        const int tn = __builtin_amdgcn_readlane(rv.x, R_TN);
        if (tn > 0) {
            const int slo = __builtin_amdgcn_readlane(rv.x, R_SLO), shi = __builtin_amdgcn_readlane(rv.x, R_SHI);
            const long long g0 = op.a_off[0] / 64 + __builtin_amdgcn_readlane(rv.x, R_STEP0) + slo;
            const TA* amat = mat + 64 * g0 + 4 * lane;  // group g of the share at + 256 g
            const int* cd = codes + g0;
            const T* bsrc = arena + op.in_base;
            const int fo_d = fo * op.deg;  // fo * nl at levels 0 and 1
            const int ns = shi - slo;
            for (int q0 = 0; q0 < ns; q0 += 64) {
                const int qn = min(64, ns - q0);
                const int cv = lane < qn ? gld(cd + q0 + lane) : 0;  // lane t: the code of step q0 + t
                for (int q = 0; q < qn; q += 4 * NG) {
                    TA av[NG][4];
                    T bv[NG][4];
#pragma unroll
                    for (int g = 0; g < NG; ++g) gld4(amat + 64 * (q0 + q + 4 * g), av[g]);
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        const int code = __shfl(cv, q + 4 * g + lk);  // step lk of group g
                        const int nv = (code >> 10) & 7;
                        const int box = __shfl(rv.y, code & 31);
                        const unsigned ad = unsigned(box + (fo_d << ((code >> 5) & 31)) + (code >> 13));
                        gld4(bsrc + (nv > 0 ? ad : 0u), bv[g]);
#pragma unroll
                        for (int j = 0; j < 4; ++j) bv[g][j] = j < nv ? bv[g][j] : T(0);
                    }
#pragma unroll
                    for (int g = 0; g < NG; ++g) acc = SepMma<T, TA>::run4(av[g], bv[g], acc);
                }
            }
        }
        if (gsize > 1 && !leader)
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) part[((w - 1) * 4 + qq) * 64 + lane] = acc[qq];
    }
    __syncthreads();  // every wave of the workgroup
    if (!leader) return;
    for (int v = 1; v < gsize; ++v)  // the tile's partial sums in wave order
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) acc[qq] += part[((w + v - 1) * 4 + qq) * 64 + lane];
    T* out = arena + op.out_base;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
        if (ook[qq]) gst(out + oaddr[qq], acc[qq]);
}

// sep_tile_cols in the quad storage: a panel of at most SEP_KMAX steps (two groups
// of four) keeps its operator values and its B offsets without the column term in
// registers and sweeps R_NCT column tiles of its group, two at a time.
template <typename T, typename TA>
__device__ __forceinline__ void sep_tile_quad_cols(const TA* __restrict__ mat, const int* __restrict__ codes,
                                                   T* __restrict__ arena, const int2 rv, const int lane,
                                                   const UdOp& op) {
    static_assert(SEP_KMAX == 8, "two groups of four steps");
    using V = typename SwMfma<T>::V;
    const int f0 = __builtin_amdgcn_readlane(rv.x, R_F0), nct = __builtin_amdgcn_readlane(rv.x, R_NCT);
    const int O = __builtin_amdgcn_readlane(rv.x, R_O);
    const int li = lane & 15, lk = lane >> 4;
    int rb[4], rn[4];
    bool rok[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = SwMfma<T>::row(lane, q);
        rb[q] = __shfl(rv.x, R_ROWBASE + row);
        rn[q] = __shfl(rv.x, R_ROWNL + row);
        rok[q] = __shfl(rv.x, R_ROWLEV + row) >= 0;
    }
    const int ns = __builtin_amdgcn_readlane(rv.x, R_SHI);  // the panel's steps (from step 0)
    const long long g0 = op.a_off[0] / 64 + __builtin_amdgcn_readlane(rv.x, R_STEP0);
    const int cv = lane < ns ? gld(codes + g0 + lane) : 0;
    const T* bsrc = arena + op.in_base;
    TA av[2][4];
    int bp[2], sh[2], nv[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        gld4(mat + 64 * (g0 + 4 * g) + 4 * lane, av[g]);
        const int code = __shfl(cv, 4 * g + lk);
        nv[g] = (code >> 10) & 7;
        bp[g] = __shfl(rv.y, code & 31) + (code >> 13);
        sh[g] = (code >> 5) & 31;
    }
    T* out = arena + op.out_base;
    for (int ct = 0; ct < nct; ct += 2) {
        const bool two = ct + 1 < nct;
        int fo[2];
        bool fok[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int f = f0 + 16 * (ct + u) + li;
            fok[u] = (u == 0 || two) && f < O;
            fo[u] = fok[u] ? f : 0;
        }
        V acc[2] = {V{T(0), T(0), T(0), T(0)}, V{T(0), T(0), T(0), T(0)}};
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            if (4 * g < ns) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    T bv[4];
                    const bool ok = nv[g] > 0 && fok[u];
                    gld4(bsrc + (ok ? unsigned(bp[g] + ((fo[u] * op.deg) << sh[g])) : 0u), bv);
#pragma unroll
                    for (int j = 0; j < 4; ++j) bv[j] = ok && j < nv[g] ? bv[j] : T(0);
                    acc[u] = SepMma<T, TA>::run4(av[g], bv, acc[u]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (fok[u] && rok[q]) gst(out + rb[q] + fo[u] * rn[q], acc[u][q]);
    }
    __syncthreads();  // the one barrier every wave of the workgroup takes (sep_tile's partial sums)
}

// One wave of a separable op over R_NCT consecutive 16-column tiles of one group
// and one row panel whose kept steps (at most SEP_KMAX, not split over waves) are
// loaded once: the panel's A values and the lane's B offsets without the column
// term stay in registers, and each column tile then costs its B loads, its matrix
// operations and its stores.  Two column tiles are in flight at a time.  Holds
// one workgroup barrier, like sep_tile.
template <typename T, typename TA>
__device__ __forceinline__ void sep_tile_cols(const TA* __restrict__ mat, const int* __restrict__ codes,
                                              T* __restrict__ arena, const int2 rv, const int lane, const UdOp& op) {
    static_assert(SEP_KMAX % 4 == 0, "steps come in groups of four");
    using V = typename SwMfma<T>::V;
    const int f0 = __builtin_amdgcn_readlane(rv.x, R_F0), nct = __builtin_amdgcn_readlane(rv.x, R_NCT);
    const int O = __builtin_amdgcn_readlane(rv.x, R_O), I = __builtin_amdgcn_readlane(rv.x, R_I);
    const unsigned magic = unsigned(__builtin_amdgcn_readlane(rv.x, R_MAGIC));
    const int li = lane & 15, lk = lane >> 4;
    int rb[4], rn[4];
    bool rok[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = SwMfma<T>::row(lane, q);
        rb[q] = __shfl(rv.x, R_ROWBASE + row);
        rn[q] = __shfl(rv.x, R_ROWNL + row);
        rok[q] = __shfl(rv.x, R_ROWLEV + row) >= 0;
    }
    const int ns = __builtin_amdgcn_readlane(rv.x, R_SHI);  // the panel's steps (from step 0)
    const long long g0 = op.a_off[0] / 64 + __builtin_amdgcn_readlane(rv.x, R_STEP0);
    const TA* amat = mat + 64 * g0 + lane;  // the storage runs SEP_PAD steps past its last
    const int cv = lane < ns ? gld(codes + g0 + lane) : 0;  // lane t: the code of step t
    const T* bsrc = arena + op.in_base;
    TA av[SEP_KMAX];
    int bp[SEP_KMAX], sh[SEP_KMAX];
    bool pb[SEP_KMAX];
#pragma unroll
    for (int c = 0; c < SEP_KMAX; ++c) {
        const int code = __builtin_amdgcn_readlane(cv, c);
        const int nv = (code >> 10) & 7;
        pb[c] = lk < nv;
        av[c] = gld(amat + 64 * c);
        bp[c] = __builtin_amdgcn_readlane(rv.y, code & 31) + ((code >> 13) + lk) * I;
        sh[c] = (code >> 5) & 31;  // fo * nl * I = (fo * deg * I) << sh
    }
    T* out = arena + op.out_base;
    for (int ct = 0; ct < nct; ct += 2) {
        const bool two = ct + 1 < nct;
        int fo[2], fi[2], fod[2];
        bool fok[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int f = f0 + 16 * (ct + u) + li;
            fok[u] = (u == 0 || two) && f < O * I;
            const int fc = fok[u] ? f : 0;
            fo[u] = I == 1 ? fc : magic ? int(__umulhi(unsigned(fc), magic)) : fc / I;
            fi[u] = fc - fo[u] * I;
            fod[u] = fo[u] * op.deg * I;
        }
        V acc[2] = {V{T(0), T(0), T(0), T(0)}, V{T(0), T(0), T(0), T(0)}};
#pragma unroll
        for (int c0 = 0; c0 < SEP_KMAX; c0 += 4) {
            if (c0 < ns) {
                T bv[2][4];
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const bool ok = pb[c0 + j] && fok[u];
                        const unsigned ad = ok ? unsigned(bp[c0 + j] + fi[u] + (fod[u] << sh[c0 + j])) : 0u;
                        const T v = gld(bsrc + ad);
                        bv[u][j] = ok ? v : T(0);
                    }
                acc[0] = SepMma<T, TA>::run4(av + c0, bv[0], acc[0]);
                acc[1] = SepMma<T, TA>::run4(av + c0, bv[1], acc[1]);
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (fok[u] && rok[q]) gst(out + rb[q] + fo[u] * rn[q] + fi[u], acc[u][q]);
    }
    __syncthreads();  // the one barrier every wave of the workgroup takes (sep_tile's partial sums)
}

// A phase of separable ops only (every op UD_SEP): the separable tile bodies
// alone, so the launch keeps the occupancy their registers allow.  TA = T, or
// BF16 values on the FP32 arena (T = float).  (The packed body's registers set the
// launch's occupancy; running the packed workgroups as a launch of their own, so
// that the plain one reaches 8 waves per SIMD, measured 20 % slower on the
// 64-instance batch: the packed launch's tail leaves the device idle.)
template <typename T, typename TA>
__global__ void __launch_bounds__(UD_WAVES * 64) k_sep_batch(const UdBatchOp* __restrict__ ops,
                                                              const int* __restrict__ wg_op,
                                                              const int2* __restrict__ recs) {
    __shared__ T part[(UD_WAVES - 1) * 4 * 64];
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int2 rv = gld(recs + ((long long)blockIdx.x * UD_WAVES + w) * (UD_REC / 2) + lane);
    const UdBatchOp& o = ops[wg_op[blockIdx.x]];
    constexpr int sl = sizeof(T) == 8 ? 0 : 1;            // arena
    constexpr int ml = sizeof(TA) == 2 ? 2 : sl;          // values
    const TA* mat = static_cast<const TA*>(o.mat[ml]);
    if (o.quad && __builtin_amdgcn_readlane(rv.x, R_GSIZE) > 0 && __builtin_amdgcn_readlane(rv.x, R_NCT) > 1)
        sep_tile_quad_cols<T, TA>(mat, o.codes, static_cast<T*>(o.arena[sl]), rv, lane, o.a);
    else if (o.quad)
        sep_tile_quad<T, TA, 2>(mat, o.codes, static_cast<T*>(o.arena[sl]), rv, w, lane, o.a, part);
    else if (__builtin_amdgcn_readlane(rv.x, R_GSIZE) > 0 && __builtin_amdgcn_readlane(rv.x, R_NCT) > 1)
        sep_tile_cols<T, TA>(mat, o.codes, static_cast<T*>(o.arena[sl]), rv, lane, o.a);
    else
        sep_tile<T, TA, 8>(mat, o.codes, static_cast<T*>(o.arena[sl]), rv, w, lane, o.a, part);
}

}  // namespace
}  // namespace kmx

using namespace kmx;

namespace {

using Lev = std::vector<int>;

long long cells_of(int l) { return l == 0 ? 1 : 1LL << (l - 1); }  // cells of level l
long long start_of(int l) { return l == 0 ? 0 : 1LL << (l - 1); }  // first 1D index of level l

int host_level(long long s, int deg) {
    const long long idx = s / deg;
    int l = 0;
    while ((1LL << l) <= idx) ++l;
    return l;  // idx in [2^(l-1), 2^l)
}

// steps of four source positions of levels la..lb
int steps_of(int la, int lb, int deg) {
    int n = 0;
    for (int l = la; l <= lb; ++l) n += int((deg * cells_of(l) + 3) / 4);
    return n;
}

// The boxes R u {d: l}, l = 0..m, of one op dimension d: O / I = values of the
// dimensions before / after d; box = their offsets, level 0 first.
struct Group {
    int m, O, I;
    std::vector<int> box;
    unsigned magic;  // s / I = umulhi(s, magic) for every column s < O * I (0: divide)
};

struct HostTile {
    int m0, f0, b, t, group, la, lb, nterm, sper, steps, gsize;
    int nct = 1;  // separable: column tiles this tile's wave sweeps (packed panels)
};

}  // namespace

// Compact row-panel storage of the combined operators of a separable sweep, one
// orientation (0: the operator, 1: its transpose, for the adjoint apply).  Op k's
// panel p (rows 16 p .. 16 p + 15 of its n x n combined matrix) has runs
// runs[run_base[k] + prun[k][p] ..  run_base[k] + prun[k][p + 1]) in level order:
// (level, first position inside the level, positions, storage offset from
// a_base[k]); a run holds 4 * ceil(positions / 4) positions x 16 rows, position
// major, zero past the run's positions and past the matrix rows.  codes: per step
// of the storage (its offset / 64), sep_code(level, valid positions, first position
// inside the level) -- what the kernels read per step.
struct SepSide {
    std::vector<int4> runs;
    std::vector<int> codes;
    // per op: the quad storage (the last dimension's op, I = 1): each panel's steps
    // padded to a multiple of four, a group of four steps stored [lane][position]
    std::vector<char> quad;
    std::vector<int> run_base;
    std::vector<std::vector<int>> prun;
    std::vector<long long> a_base;
    long long a_elems = 0;
    int* d_codes = nullptr;
    double* d_a64 = nullptr;
    float* d_a32 = nullptr;
    unsigned short* d_a16 = nullptr;  // BF16 bit patterns of the same values (with the FP32 copy)
};

struct kmx_sweep {
    int D = 0;
    long long nb = 0;     // values per intermediate (every box of every level tuple)
    long long arena = 0;  // elements
    struct Op {
        int part;
        bool final;
        UdOp a;
        int wg0, wg1;  // workgroups (UD_WAVES work records each) of this op
        int nbr;       // branches per tile
        std::vector<int> terms;  // plan term of each of the op's terms (scales of a final op)
    };
    std::vector<Op> ops;
    int* d_xmap = nullptr;
    int* d_ymap = nullptr;
    int* d_recs = nullptr;
    void* d_arena[2] = {nullptr, nullptr};  // FP64, FP32
    long long flops = 0;                    // useful flops per apply (all terms)
    long long mfma_flops = 0;               // matrix-core flops issued per apply (tiles and steps as run)
    // what the work records are emitted from (a batch emits them again with its own split)
    struct OpDesc {
        int part;
        bool fin;
        UdOp a;
        long long nbranch;
        std::vector<int> terms;
    };
    std::vector<OpDesc> desc;
    // separable form: every term acts in at most one dimension (its other 1D
    // factors are the identity, or declared so), so the apply is one whole-operator
    // op per dimension; identity[t * D + k] records the factors taken as the
    // identity, and stale is set when kmx_plan_update_term replaces one of them by
    // another matrix.  fp64 is false when some such factor is only within
    // identity_tol of the identity (then only FP32 applies may use the sweep);
    // identity_dev is the largest max |A - I| over the undeclared ones.
    bool separable = false;
    std::vector<char> identity;
    bool stale = false;
    bool fp64 = true;
    double identity_dev = 0.0;
    double drop_tol = 0.0;
    long long kept_blocks = 0, total_blocks = 0;
    // separable: per op, the kept deg x deg blocks of its combined matrix (nb_d x nb_d)
    std::vector<std::vector<unsigned char>> keep;
    SepSide side[2];
    // storage and launch tables replaced by a rebuild, kept until the plan is
    // destroyed (a graph captured earlier may still read them)
    std::vector<SepSide> retired;
    std::vector<kmx_sweep_batch*> retired_batches;
    kmx_sweep_batch* self = nullptr;   // separable: the plan's own launches (a batch of one)
    kmx_sweep_batch* selfT = nullptr;  // separable: the adjoint apply's launches
    std::vector<Group> groups;
    std::vector<int> group0;  // groups of op dimension d: [group0[d], group0[d+1])
    int deg = 0, T = 0, max_nbr = 2;
};

namespace {

// Load batches per wave before a tile is split over more waves: `dflt`, or
// KMX_SWEEP_WAVE_BATCHES when set.
int sweep_wave_batches(int dflt) {
    if (const char* e = std::getenv("KMX_SWEEP_WAVE_BATCHES")) return std::max(1, std::atoi(e));
    return dflt;
}

// The quad storage for the separable op of the last dimension (KMX_SEP_QUAD=0: the
// plain storage for every op).
bool sep_quad_on() {
    const char* e = std::getenv("KMX_SEP_QUAD");
    return !(e && e[0] == '0');
}

// Workgroups of a merged separable launch dealt to one XCD in runs of this many
// consecutive entries of the work list (KMX_SEP_XCD_BLOCK; 1: the list order as is).
// 64-instance batch, FP32 (alternating runs on one GPU, results identical): fused RK3 step 0.228 ms in
// list order, 0.220 at runs of 8, 0.214-0.217 at 16, 0.216-0.218 at 32, 0.215 at 64,
// 0.219 at 128; batched apply 0.095-0.098 -> 0.094.
int sep_xcd_block() {
    if (const char* e = std::getenv("KMX_SEP_XCD_BLOCK")) return std::max(1, std::atoi(e));
    return 16;
}

// The RK3 stage launches four values per thread (KMX_RK3_VEC=0: one).
bool rk3_vec_on() {
    const char* e = std::getenv("KMX_RK3_VEC");
    return !(e && e[0] == '0');
}

double env_double(const char* name, double dflt) {
    if (const char* e = std::getenv(name)) return std::atof(e);
    return dflt;
}

// Runs of panel p of separable op k (orientation `side`) whose level is <= m: a
// prefix of the panel's runs.  Returns their count and total steps.
void panel_prefix(const SepSide& sd, int k, int p, int m, int& nrun, int& steps) {
    nrun = 0;
    steps = 0;
    const int r0 = sd.run_base[size_t(k)];
    for (int q = sd.prun[size_t(k)][size_t(p)]; q < sd.prun[size_t(k)][size_t(p) + 1]; ++q) {
        const int4& ru = sd.runs[size_t(r0 + q)];
        if (ru.x > m) break;
        ++nrun;
        steps += (ru.z + 3) / 4;
    }
}

// Work records of every op of a sweep.  A tile (16 rows x 16 columns x nbr
// branches) whose K steps exceed wave_batches load batches is split over up to
// UD_WAVES waves of one workgroup: a lone plan (latency-bound, few tiles) uses
// one batch per wave, a batch of plans (many tiles) four (64-instance batch, FP32
// apply: 0.1028 ms at four, 0.1046 at three, 0.1109 at two, 0.1079 at eight;
// kmx_batch_autotune times 4, 8 and 16 on the batch itself).  Widest tiles first
// (power-of-two wave counts then never straddle a workgroup), longest first.
// Separable ops take their steps from the panels of orientation `side`.
void emit_records(const kmx_sweep& sw, int wave_batches, std::vector<std::vector<int>>& recs,
                  std::vector<kmx_sweep::Op>& ops, int side = 0) {
    const int deg = sw.deg;
    recs.clear();
    ops.clear();
    int sep_k = 0;  // separable op index (the ops of a separable sweep are all UD_SEP, in order)
    for (const kmx_sweep::OpDesc& od : sw.desc) {
        const int T = od.a.T;  // the op's terms
        const int part = od.part;
        const bool fin = od.fin;
        const bool sep = part == UD_SEP;
        const long long nbranch = od.nbranch;
        const int nbr = int(std::min<long long>(nbranch, sw.max_nbr));  // branches per tile (a power of two)
        kmx_sweep::Op op{part, fin, od.a, int(recs.size() / UD_WAVES), 0, nbr, od.terms};
        if (sep) op.a.a_off[0] = sw.side[side].a_base[size_t(sep_k)];
        const bool quad = sep && sw.side[side].quad[size_t(sep_k)] != 0;
        const int d = od.a.d;
        const bool multi = fin || sep;  // one tile sums every term of the op
        const int nterm = multi ? T : 1;
        const int batch = ud_split_batch(nbr) * std::max(1, wave_batches);
        std::vector<HostTile> ot;
        for (long long b = 0; b < nbranch; b += nbr)
            for (int t = 0; t < (multi ? 1 : T); ++t)
                for (int gi = sw.group0[d]; gi < sw.group0[d + 1]; ++gi) {
                    const Group& g = sw.groups[gi];
                    const int n = deg << g.m;
                    for (int m0 = 0; m0 < n; m0 += 16) {
                        // source levels: all (whole operator), >= the rows' lowest (down), < their highest (up)
                        int la = 0, lb = g.m;
                        if (part == UD_DOWN) la = host_level(m0, deg);
                        if (part == UD_UP) lb = host_level(std::min(n, m0 + 16) - 1, deg) - 1;
                        int sper = 0, run0 = 0, nrun = 0;
                        if (sep) {
                            panel_prefix(sw.side[side], sep_k, m0 / 16, g.m, nrun, sper);
                            run0 = sw.side[side].run_base[size_t(sep_k)] +
                                   sw.side[side].prun[size_t(sep_k)][size_t(m0 / 16)];
                            la = run0;
                            lb = nrun;
                        } else {
                            sper = steps_of(la, lb, deg);
                        }
                        const int steps = nterm * sper;
                        const int ctiles = (g.O * g.I + 15) / 16;
                        if (sep && sper >= 1 && sper <= SEP_KMAX && ctiles > 1) {
                            // a short panel: one wave sweeps up to batch / steps column tiles
                            const int per = std::max(1, std::min(ctiles, batch / sper));
                            for (int c0 = 0; c0 < ctiles; c0 += per) {
                                HostTile h{m0, 16 * c0, int(b), t, gi, la, lb, nterm, sper, steps, 1};
                                h.nct = std::min(per, ctiles - c0);
                                ot.push_back(h);
                            }
                            continue;
                        }
                        int gsize = 1;  // waves sharing the tile: at most `batch` steps each
                        while (gsize < UD_WAVES && (steps + gsize - 1) / gsize > batch) gsize *= 2;
                        for (int f0 = 0; f0 < g.O * g.I; f0 += 16)
                            ot.push_back(HostTile{m0, f0, int(b), t, gi, la, lb, nterm, sper, steps, gsize});
                    }
                }
        std::stable_sort(ot.begin(), ot.end(), [](const HostTile& u, const HostTile& v) {
            return u.gsize != v.gsize ? u.gsize > v.gsize : u.steps * u.nct > v.steps * v.nct;
        });
        for (const HostTile& h : ot) {
            const int slot = int(recs.size() % UD_WAVES);
            if (slot % h.gsize) throw std::logic_error("sweep tile straddles a workgroup");
            const Group& g = sw.groups[h.group];
            const int n = deg << g.m;
            int qn = (h.steps + h.gsize - 1) / h.gsize;
            if (quad) qn = (qn + 3) / 4 * 4;  // shares start at a group of four steps
            for (int w = 0; w < h.gsize; ++w) {
                std::vector<int> r(UD_REC, 0);
                r[R_M0] = h.m0;
                r[R_F0] = h.f0;
                r[R_B] = h.b;
                r[R_TOUT] = h.t;
                r[R_O] = g.O;
                r[R_I] = g.I;
                r[R_LA] = h.la;
                r[R_LB] = h.lb;
                r[R_SPER] = h.sper;
                r[R_GSIZE] = h.gsize;
                r[R_GBASE] = slot;
                r[R_MAGIC] = int(g.magic);
                const int q0 = std::min(h.steps, w * qn), q1 = std::min(h.steps, (w + 1) * qn);
                if (sep) {  // steps counted over the tile's runs; R_NCT column tiles
                    r[R_NCT] = h.nct;
                    r[R_STEP0] = h.lb > 0 ? sw.side[side].runs[size_t(h.la)].w / 64 : 0;
                    r[R_SLO] = q0;
                    r[R_SHI] = q1;
                    r[R_TN] = q1 > q0 ? 1 : 0;
                } else if (q1 > q0) {
                    r[R_TLO] = h.t + q0 / h.sper;
                    r[R_SLO] = q0 % h.sper;
                    r[R_TN] = (q1 - 1) / h.sper - q0 / h.sper + 1;
                    r[R_SHI] = (q1 - 1) % h.sper + 1;
                } else {
                    r[R_TLO] = h.t;
                }
                for (int i = 0; i < 16; ++i) {
                    const int row = h.m0 + i;
                    if (row < n) {
                        const int l = host_level(row, deg);
                        r[R_ROWBASE + i] = g.box[l] + int(row - deg * start_of(l)) * g.I;
                        r[R_ROWNL + i] = int(deg * cells_of(l)) * g.I;
                        r[R_ROWLEV + i] = l;
                    } else {
                        r[R_ROWLEV + i] = -1;
                    }
                }
                for (int l = 0; l <= g.m; ++l) r[UD_BOX + l] = g.box[l];
                recs.push_back(std::move(r));
            }
        }
        while (recs.size() % UD_WAVES) recs.push_back(std::vector<int>(UD_REC, 0));  // empty slots
        op.wg1 = int(recs.size() / UD_WAVES);
        ops.push_back(op);
        if (sep) ++sep_k;
    }
}


// Every work record: its tile inside a group of its op, its slots inside the
// arena, its rows and boxes as the group defines them, a separable tile's runs
// inside its panel, its group's levels and the op's storage (host checks, before
// upload).
void check_records(const kmx_sweep& sw, const std::vector<std::vector<int>>& recs,
                   const std::vector<kmx_sweep::Op>& ops, int side = 0) {
    const int deg = sw.deg;
    const long long nb = sw.nb, half = 1LL << (sw.D - 1);
    const std::vector<Group>& groups = sw.groups;
    const std::vector<int>& group0 = sw.group0;
    auto slot_ok = [&](long long off) { return off >= 0 && off + nb <= sw.arena; };
    int sep_k = 0;
    for (const auto& op : ops) {
        const int T = op.a.T;
        // (a separable op's terms are folded into its one combined matrix)
        if (T < 1 || T > KMX_MAX_TERMS || op.terms.empty() || (op.part != UD_SEP && int(op.terms.size()) != T) ||
            (op.part == UD_SEP && T != 1))
            throw std::logic_error("sweep op terms");
        for (int t : op.terms)
            if (t < 0 || t >= sw.T) throw std::logic_error("sweep op terms");
        const bool sep = op.part == UD_SEP;
        const long long nbranch = (op.final || sep) ? 1 : (op.part == UD_FULL ? half : (1LL << op.a.d));
        if (op.nbr < 1 || op.nbr > 4 || nbranch % op.nbr || (op.final && op.nbr != 1))
            throw std::logic_error("sweep branches per tile");
        const SepSide& sd = sw.side[side];
        long long a_lo = 0, a_hi = 0;
        if (sep) {
            if (size_t(sep_k) >= sd.a_base.size() || op.a.a_off[0] != sd.a_base[size_t(sep_k)])
                throw std::logic_error("sweep separable storage offset");
            a_lo = sd.a_base[size_t(sep_k)];
            a_hi = size_t(sep_k) + 1 < sd.a_base.size() ? sd.a_base[size_t(sep_k) + 1] : sd.a_elems;
        }
        for (int wg = op.wg0; wg < op.wg1; ++wg) {
            for (int w = 0; w < UD_WAVES; ++w) {
                const std::vector<int>& r = recs[size_t(wg) * UD_WAVES + w];
                if (r[R_GSIZE] == 0) continue;
                const int gbase = r[R_GBASE], gsize = r[R_GSIZE];
                if (gsize < 1 || gsize > UD_WAVES || gbase < 0 || gbase > w || w >= gbase + gsize ||
                    gbase + gsize > UD_WAVES)
                    throw std::logic_error("sweep record wave group");
                for (int v = gbase; v < gbase + gsize; ++v) {  // the waves of one tile agree on it
                    const std::vector<int>& o = recs[size_t(wg) * UD_WAVES + v];
                    for (int k : {int(R_M0), int(R_F0), int(R_B), int(R_TOUT), int(R_O), int(R_I), int(R_LA),
                                  int(R_LB), int(R_SPER), int(R_GSIZE), int(R_GBASE)})
                        if (o[k] != r[k]) throw std::logic_error("sweep record wave group");
                }
                // the group: same O, I and level-0 box
                const Group* gp = nullptr;
                for (int gi = group0[op.a.d]; gi < group0[op.a.d + 1] && !gp; ++gi)
                    if (groups[gi].O == r[R_O] && groups[gi].I == r[R_I] && groups[gi].box[0] == r[UD_BOX])
                        gp = &groups[gi];
                if (!gp) throw std::logic_error("sweep record group");
                const Group& g = *gp;
                const int n = deg << g.m;
                if (r[R_M0] < 0 || r[R_M0] >= n || r[R_M0] % 16 || r[R_F0] < 0 || r[R_F0] >= g.O * g.I ||
                    r[R_B] < 0 || r[R_B] % op.nbr || r[R_B] + op.nbr > nbranch || r[R_TOUT] < 0 ||
                    (!sep && r[R_TOUT] >= T) || r[R_MAGIC] != int(g.magic))  // (R_STEP0 on separable records)
                    throw std::logic_error("sweep record outside its group");
                for (int i = 0; i < 16; ++i) {
                    const int row = r[R_M0] + i;
                    const int l = row < n ? host_level(row, deg) : -1;
                    if (r[R_ROWLEV + i] != l ||
                        (l >= 0 && (r[R_ROWBASE + i] != g.box[l] + int(row - deg * start_of(l)) * g.I ||
                                    r[R_ROWNL + i] != int(deg * cells_of(l)) * g.I)))
                        throw std::logic_error("sweep record rows");
                }
                for (int l = 0; l <= g.m; ++l)
                    if (r[UD_BOX + l] != g.box[l]) throw std::logic_error("sweep record boxes");
                if (sep) {
                    // one box copy of x in, the op's own slot out
                    if (op.a.x0 || op.final || !slot_ok(op.a.in_base) || !slot_ok(op.a.out_base))
                        throw std::logic_error("sweep separable slots outside the arena");
                    // the runs: the panel's own, levels of the group, positions inside their level,
                    // storage inside the op's, and the wave's steps inside them
                    const int p = r[R_M0] / 16;
                    const int rb = sd.run_base[size_t(sep_k)];
                    const std::vector<int>& pr = sd.prun[size_t(sep_k)];
                    if (size_t(p) + 1 >= pr.size() || r[R_RUN0] != rb + pr[size_t(p)] || r[R_NRUN] < 0 ||
                        r[R_NRUN] > pr[size_t(p) + 1] - pr[size_t(p)])
                        throw std::logic_error("sweep separable record runs");
                    long long steps = 0;
                    const long long st0 = r[R_NRUN] > 0 ? sd.runs[size_t(r[R_RUN0])].w / 64 : 0;
                    if (r[R_STEP0] != st0) throw std::logic_error("sweep separable record first step");
                    for (int q = r[R_RUN0]; q < r[R_RUN0] + r[R_NRUN]; ++q) {
                        const int4& ru = sd.runs[size_t(q)];
                        const long long nl = deg * cells_of(ru.x);
                        const long long nst = (ru.z + 3) / 4;
                        if (ru.x < 0 || ru.x > g.m || ru.x > UD_MAXLEV || ru.y < 0 || ru.y % 4 || ru.z < 1 ||
                            ru.y + ru.z > nl || a_lo + ru.w < a_lo || a_lo + ru.w + 64 * nst > a_hi || ru.w % 64 ||
                            ru.w / 64 != st0 + steps)
                            throw std::logic_error("sweep separable run");
                        for (long long t = 0; t < nst; ++t)  // the codes the kernels read for these steps
                            if (sd.codes[size_t(a_lo / 64 + st0 + steps + t)] !=
                                sep_code(ru.x, int(std::min(4LL, ru.z - 4 * t)), int(ru.y + 4 * t)))
                                throw std::logic_error("sweep separable step code");
                        steps += nst;
                    }
                    if (r[R_SPER] != steps || r[R_TN] != (r[R_SHI] > r[R_SLO] ? 1 : 0) || r[R_SLO] < 0 ||
                        r[R_SHI] < r[R_SLO] || r[R_SHI] > steps)
                        throw std::logic_error("sweep separable record steps");
                    const int nct = r[R_NCT];
                    if (sd.quad[size_t(sep_k)] && (r[R_SLO] % 4 || g.I != 1))
                        throw std::logic_error("sweep quad record");
                    if (nct < 1 || r[R_F0] % 16 || r[R_F0] + 16 * (nct - 1) >= g.O * g.I ||
                        (nct > 1 && (gsize != 1 || r[R_SLO] != 0 || r[R_SHI] != steps || steps < 1 ||
                                     steps > SEP_KMAX || r[R_NRUN] > SEP_KMAX)))
                        throw std::logic_error("sweep separable record column tiles");
                    continue;
                }
                const int sper = steps_of(r[R_LA], r[R_LB], deg);
                if (r[R_LA] < 0 || r[R_LB] > g.m || r[R_SPER] != sper)
                    throw std::logic_error("sweep record outside its group");
                if (r[R_TN] < 0 || (r[R_TN] > 0 && (r[R_TLO] < 0 || r[R_TLO] + r[R_TN] > T || r[R_SLO] < 0 ||
                                                    r[R_SLO] >= sper || r[R_SHI] < 1 || r[R_SHI] > sper ||
                                                    (!op.final && (r[R_TN] != 1 || r[R_TLO] != r[R_TOUT])))))
                    throw std::logic_error("sweep record steps");
                for (int bj = r[R_B]; bj < r[R_B] + op.nbr; ++bj) {  // every branch of the tile
                    for (int t = r[R_TLO]; t < r[R_TLO] + r[R_TN]; ++t)
                        if (!(op.a.x0 && bj == 0) && !slot_ok(op.a.in_base + ((long long)bj * T + t) * nb))
                            throw std::logic_error("sweep input slot outside the arena");
                    for (int t = 0; t < T; ++t) {
                        if (!op.final && t != r[R_TOUT]) continue;
                        if (op.part == UD_UP &&
                            !slot_ok(op.a.add_base + ((long long)(bj | op.a.add_bit) * T + t) * nb))
                            throw std::logic_error("sweep add slot outside the arena");
                        if (!op.final && !slot_ok(op.a.out_base + ((long long)(bj | op.a.out_bit) * T + t) * nb))
                            throw std::logic_error("sweep output slot outside the arena");
                    }
                }
            }
        }
        if (sep) ++sep_k;
    }
}

// max |A - I| over every entry of 1D factor (t, k) (the FP64 copy the plan
// keeps, block-row-major).
double identity_dev(const kmx_plan* p, int t, int k) {
    const int deg = p->deg, nbk = p->nb[k];
    const double* m = p->h_blk64.data() + p->off[t * p->D + k];
    double dev = 0.0;
    for (long long I = 0; I < nbk; ++I)
        for (long long J = 0; J < nbk; ++J) {
            const double* b = m + (I * nbk + J) * deg * deg;
            for (int a = 0; a < deg; ++a)
                for (int c = 0; c < deg; ++c)
                    dev = std::max(dev, std::fabs(b[a * deg + c] - (I == J && a == c ? 1.0 : 0.0)));
        }
    return dev;
}

// Entry (r, c) of factor (t, d), masked by dimension d's 1D connections.
inline double masked_entry(const kmx_plan* p, int t, int d, long long r, long long c) {
    const int deg = p->deg, nbk = p->nb[d];
    const long long I = r / deg, J = c / deg;
    if (!p->bs->mask[size_t(d)][size_t(I * nbk + J)]) return 0.0;
    return p->h_blk64[size_t(p->off[t * p->D + d] + ((I * nbk + J) * deg + r % deg) * deg + c % deg)];
}

// Kept blocks of separable op k: block (I, J) of its combined matrix is kept when
// some term of the op has a connected block there whose largest entry exceeds
// drop_tol times the largest connected entry of that term's matrix (drop_tol = 0:
// every block with a nonzero entry).  Independent of the term scales.
std::vector<unsigned char> sep_pattern(const kmx_plan* p, const kmx_sweep& sw, int k) {
    const kmx_sweep::OpDesc& od = sw.desc[size_t(k)];
    const int d = od.a.d, deg = p->deg, nbk = p->nb[d];
    std::vector<unsigned char> keep(size_t(nbk) * nbk, 0);
    for (int t : od.terms) {
        const double* m = p->h_blk64.data() + p->off[t * p->D + d];
        const std::vector<unsigned char>& mask = p->bs->mask[size_t(d)];
        double big = 0.0;
        std::vector<double> bmax(size_t(nbk) * nbk, 0.0);
        for (long long q = 0; q < (long long)nbk * nbk; ++q) {
            if (!mask[size_t(q)]) continue;
            double v = 0.0;
            for (int e = 0; e < deg * deg; ++e) v = std::max(v, std::fabs(m[q * deg * deg + e]));
            bmax[size_t(q)] = v;
            big = std::max(big, v);
        }
        const double cut = sw.drop_tol * big;
        for (size_t q = 0; q < bmax.size(); ++q)
            if (bmax[q] > cut) keep[q] = 1;
    }
    return keep;
}

// Runs of every op of orientation `s` from the kept-block patterns (orientation 1
// reads them transposed); storage offsets in units of 64 values.
void sep_runs(const kmx_plan* p, kmx_sweep& sw, int s) {
    SepSide& sd = sw.side[s];
    sd.runs.clear();
    sd.codes.clear();
    sd.quad.clear();
    sd.run_base.clear();
    sd.prun.clear();
    sd.a_base.clear();
    const int deg = p->deg;
    long long abase = 0;
    for (size_t k = 0; k < sw.desc.size(); ++k) {
        const int d = sw.desc[k].a.d, nbk = p->nb[d], n = p->n[d];
        const std::vector<unsigned char>& keep = sw.keep[k];
        auto kept = [&](int I, int J) { return keep[s == 0 ? size_t(I) * nbk + J : size_t(J) * nbk + I] != 0; };
        int L = 0;
        while ((deg << L) < n) ++L;
        if ((deg << L) != n) throw std::logic_error("separable op: matrix size is not deg * 2^L");
        const int npan = (n + 15) / 16;
        const bool quad = sep_quad_on() && d == p->D - 1;
        sd.quad.push_back(quad ? 1 : 0);
        sd.run_base.push_back(int(sd.runs.size()));
        sd.a_base.push_back(abase);
        std::vector<int> pr(size_t(npan) + 1, 0);
        long long aoff = 0;
        std::vector<unsigned char> colneed(static_cast<size_t>(nbk), 0);
        for (int pnl = 0; pnl < npan; ++pnl) {
            pr[size_t(pnl)] = int(sd.runs.size()) - sd.run_base.back();
            const long long panel_a0 = aoff;
            const int I0 = (16 * pnl) / deg, I1 = std::min(n - 1, 16 * pnl + 15) / deg;
            for (int J = 0; J < nbk; ++J) {
                bool any = false;
                for (int I = I0; I <= I1 && !any; ++I) any = kept(I, J);
                colneed[size_t(J)] = any ? 1 : 0;
            }
            for (int l = 0; l <= L; ++l) {
                const long long k0 = deg * start_of(l), nl = deg * cells_of(l);
                const long long nst = (nl + 3) / 4;
                long long s0 = 0;
                while (s0 < nst) {
                    auto need = [&](long long st) {
                        const long long a = k0 + 4 * st, b = k0 + std::min(nl, 4 * st + 4) - 1;
                        for (long long J = a / deg; J <= b / deg; ++J)
                            if (colneed[size_t(J)]) return true;
                        return false;
                    };
                    if (!need(s0)) {
                        ++s0;
                        continue;
                    }
                    long long s1 = s0 + 1;
                    while (s1 < nst && need(s1)) ++s1;
                    const long long pos0 = 4 * s0, npos = std::min(nl, 4 * s1) - pos0;
                    if (aoff > (1LL << 30)) throw std::runtime_error("separable op storage beyond 32-bit offsets");
                    sd.runs.push_back(make_int4(l, int(pos0), int(npos), int(aoff)));
                    if (l > 31 || pos0 + npos >= (1LL << 18)) throw std::runtime_error("separable op beyond the step codes");
                    for (long long s = 0; 4 * s < npos; ++s)
                        sd.codes.push_back(sep_code(l, int(std::min(4LL, npos - 4 * s)), int(pos0 + 4 * s)));
                    aoff += 64 * (s1 - s0);
                    s0 = s1;
                }
            }
            if (quad)  // zero steps (no positions) complete the panel's last group of four
                for (; (aoff - panel_a0) % 256; aoff += 64) sd.codes.push_back(sep_code(0, 0, 0));
        }
        pr[size_t(npan)] = int(sd.runs.size()) - sd.run_base.back();
        sd.prun.push_back(std::move(pr));
        if (aoff == 0) sd.codes.push_back(0);    // the dummy step's code
        abase += std::max<long long>(aoff, 64);  // an op without runs still owns 64 values (its tiles' dummy reads)
        if (sd.runs.size() > size_t(1) << 30) throw std::runtime_error("separable op runs beyond 32-bit indices");
    }
    sd.a_elems = abase;
    if ((long long)sd.codes.size() * 64 != abase) throw std::logic_error("separable step codes out of step");
}

// FP64 -> BF16 bit pattern, rounded once to nearest even at 8 significant bits
// (no detour through FP32, whose own rounding could move a tie).
unsigned short host_bf16(double v) {
    if (std::isnan(v)) return 0x7fc0;
    if (v == 0.0 || std::isinf(v)) return (unsigned short)(__builtin_bit_cast(unsigned, float(v)) >> 16);
    int e = 0;
    const double m = std::frexp(v, &e);                     // |m| in [0.5, 1)
    const double r = std::nearbyint(std::ldexp(m, 8));      // 8 significant bits, ties to even
    const float f = float(std::ldexp(r, e - 8));            // exact unless out of range
    const unsigned u = __builtin_bit_cast(unsigned, f);
    if (std::fabs(f) < 1.1754943508222875e-38f) return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    return (unsigned short)(u >> 16);
}

// The combined values (terms summed in order, scales included) of op `only` (or
// every op) of orientation `s` into its compact storage (FP64, FP32 and BF16) and
// the run table; allocates the device buffers on first use.  Waits for the device first
// (queued applies may read them).
void sep_values(const kmx_plan* p, kmx_sweep& sw, int s, int only = -1) {
    SepSide& sd = sw.side[s];
    const int deg = p->deg;
    (void)deg;
    if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
    const bool want32 = p->bs->d_mat32 != nullptr;
    if (!sd.d_codes) {
        // zero padding after the values (SEP_PAD steps)
        const size_t nv = size_t(sd.a_elems) + 64 * SEP_PAD, nc = std::max<size_t>(1, sd.codes.size());
        if (hipMalloc(&sd.d_codes, nc * sizeof(int)) != hipSuccess ||
            hipMalloc(&sd.d_a64, nv * sizeof(double)) != hipSuccess ||
            (want32 && hipMalloc(&sd.d_a32, nv * sizeof(float)) != hipSuccess) ||
            (want32 && hipMalloc(&sd.d_a16, nv * sizeof(unsigned short)) != hipSuccess))
            throw std::bad_alloc();
        if (hipMemset(sd.d_codes, 0, nc * sizeof(int)) != hipSuccess ||
            hipMemset(sd.d_a64, 0, nv * sizeof(double)) != hipSuccess ||
            (want32 && hipMemset(sd.d_a32, 0, nv * sizeof(float)) != hipSuccess) ||
            (want32 && hipMemset(sd.d_a16, 0, nv * sizeof(unsigned short)) != hipSuccess) ||
            hipDeviceSynchronize() != hipSuccess)
            throw std::runtime_error("hipMemset failed");
        if (!sd.codes.empty() &&
            hipMemcpy(sd.d_codes, sd.codes.data(), sd.codes.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
    }
    for (size_t k = 0; k < sw.desc.size(); ++k) {
        if (only >= 0 && int(k) != only) continue;
        const kmx_sweep::OpDesc& od = sw.desc[k];
        const int d = od.a.d, n = p->n[d];
        const long long lo = sd.a_base[k], hi = k + 1 < sd.a_base.size() ? sd.a_base[k + 1] : sd.a_elems;
        std::vector<double> v64(size_t(hi - lo), 0.0);
        const std::vector<int>& pr = sd.prun[k];
        const bool quad = sd.quad[k] != 0;
        for (size_t pnl = 0; pnl + 1 < pr.size(); ++pnl) {
            const long long panel_a0 = pr[pnl] < pr[pnl + 1] ? sd.runs[size_t(sd.run_base[k] + pr[pnl])].w : 0;
            for (int q = pr[pnl]; q < pr[pnl + 1]; ++q) {
                const int4& ru = sd.runs[size_t(sd.run_base[k] + q)];
                const long long k0 = p->deg * start_of(ru.x) + ru.y;
                for (int j = 0; j < ru.z; ++j)
                    for (int li = 0; li < 16; ++li) {
                        const long long r = 16 * (long long)pnl + li;
                        if (r >= n) continue;
                        const long long c = k0 + j;
                        double v = 0.0;
                        for (int t : od.terms)
                            v += p->scale[t] * (s == 0 ? masked_entry(p, t, d, r, c) : masked_entry(p, t, d, c, r));
                        // (row li, position j % 4 of step t): [step][position][row], or in the quad
                        // storage [group t / 4][lane 16 (t % 4) + li][position]
                        const long long t = (ru.w - panel_a0) / 64 + j / 4;
                        const long long at = quad ? panel_a0 + 256 * (t / 4) + 4 * (16 * (t % 4) + li) + j % 4
                                                  : ru.w + 16 * j + li;
                        v64[size_t(at)] = v;
                    }
            }
        }
        if (!v64.empty() && hipMemcpy(sd.d_a64 + lo, v64.data(), v64.size() * sizeof(double), hipMemcpyHostToDevice) !=
                                hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        if (want32 && !v64.empty()) {
            std::vector<float> v32(v64.begin(), v64.end());
            std::vector<unsigned short> v16(v64.size());
            for (size_t i = 0; i < v64.size(); ++i) v16[i] = host_bf16(v64[i]);
            if (hipMemcpy(sd.d_a32 + lo, v32.data(), v32.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(sd.d_a16 + lo, v16.data(), v16.size() * sizeof(unsigned short), hipMemcpyHostToDevice) !=
                    hipSuccess)
                throw std::runtime_error("hipMemcpy failed");
        }
    }
}

void sep_side_free(SepSide& sd) {
    if (sd.d_codes) hipFree(sd.d_codes);
    if (sd.d_a64) hipFree(sd.d_a64);
    if (sd.d_a32) hipFree(sd.d_a32);
    if (sd.d_a16) hipFree(sd.d_a16);
    sd.d_codes = nullptr;
    sd.d_a64 = nullptr;
    sd.d_a32 = nullptr;
    sd.d_a16 = nullptr;
}

// Patterns, runs and values of both orientations, the flop counts, and the plan's
// own launch tables (a batch of one, forward and adjoint).
void sep_build(kmx_plan* p, kmx_sweep& sw) {
    sw.keep.clear();
    sw.kept_blocks = sw.total_blocks = 0;
    for (size_t k = 0; k < sw.desc.size(); ++k) {
        sw.keep.push_back(sep_pattern(p, sw, int(k)));
        for (unsigned char c : sw.keep.back()) sw.kept_blocks += c;
        sw.total_blocks += (long long)sw.keep.back().size();
    }
    for (int s = 0; s < 2; ++s) {
        if (sw.side[s].d_codes) {
            sw.retired.push_back(sw.side[s]);
            sw.side[s].d_codes = nullptr;
            sw.side[s].d_a64 = nullptr;
            sw.side[s].d_a32 = nullptr;
        }
        sep_runs(p, sw, s);
        sep_values(p, sw, s);
    }
    // flops: kept blocks of each group's leading block, times its columns (useful);
    // tiles x steps (issued on the matrix cores)
    const int deg = p->deg;
    sw.flops = sw.mfma_flops = 0;
    for (size_t k = 0; k < sw.desc.size(); ++k) {
        const int d = sw.desc[k].a.d, nbk = p->nb[d];
        for (int gi = sw.group0[d]; gi < sw.group0[d + 1]; ++gi) {
            const Group& g = sw.groups[size_t(gi)];
            const long long cols = (long long)g.O * g.I;
            const int nbm = int((deg << g.m) / deg);
            long long kb = 0;
            for (int I = 0; I < nbm; ++I)
                for (int J = 0; J < nbm; ++J) kb += sw.keep[k][size_t(I) * nbk + J];
            sw.flops += 2LL * kb * deg * deg * cols;
            const int n = deg << g.m;
            for (int m0 = 0; m0 < n; m0 += 16) {
                int nrun, steps;
                panel_prefix(sw.side[0], int(k), m0 / 16, g.m, nrun, steps);
                sw.mfma_flops += 2LL * 16 * 16 * 4 * steps * ((cols + 15) / 16);
            }
        }
    }
    if (sw.self) sw.retired_batches.push_back(sw.self);
    if (sw.selfT) sw.retired_batches.push_back(sw.selfT);
    sw.self = sw.selfT = nullptr;
    sw.self = kmx_sweep_batch_build({p}, sweep_wave_batches(1));
    sw.selfT = kmx_sweep_batch_build({p}, sweep_wave_batches(1), true);
}

}  // namespace

void kmx_sweep_build(kmx_plan* p) {
    kmx_blockset* bs = p->bs;
    if (!bs || !bs->ok || !bs->full_product || p->T > KMX_MAX_TERMS) return;
    const int D = p->D, deg = p->deg, T = p->T;
    std::set<Lev> S;
    for (int k = 0; k < bs->nblk; ++k) S.insert(Lev(bs->blev.begin() + size_t(k) * D, bs->blev.begin() + size_t(k + 1) * D));
    // downward closed: lowering any level keeps the tuple in the set
    for (const Lev& L : S)
        for (int k = 0; k < D; ++k)
            if (L[k] > 0) {
                Lev M = L;
                --M[k];
                if (!S.count(M)) return;
            }
    auto* sw = new kmx_sweep();
    sw->D = D;
    try {
        // ---- boxes: one per level tuple, dimension 1 slowest
        std::map<Lev, long long> boxoff;
        long long nb = 0;
        for (const Lev& L : S) {
            boxoff[L] = nb;
            long long sz = 1;
            for (int k = 0; k < D; ++k) sz *= deg * cells_of(L[k]);
            nb += sz;
        }
        if (nb >= (1LL << 31) - 16) throw std::runtime_error("sweep boxes beyond 32-bit offsets");
        if ((long long)p->N * p->B >= (1LL << 31)) throw std::runtime_error("sweep values beyond 32-bit indices");
        sw->nb = nb;
        for (int k = 0; k < D; ++k)
            if ((long long)p->n[k] * p->n[k] >= (1LL << 31)) throw std::runtime_error("1D matrix beyond 32-bit offsets");
        // ---- groups per op dimension
        std::vector<Group> groups;
        std::vector<int> group0(D + 1);
        std::vector<long long> fl_full(D, 0), fl_down(D, 0), fl_up(D, 0);  // one branch, one term
        for (int d = 0; d < D; ++d) {
            group0[d] = int(groups.size());
            std::map<Lev, int> rmax;
            for (const Lev& L : S) {
                Lev R = L;
                R.erase(R.begin() + d);
                auto it = rmax.find(R);
                if (it == rmax.end()) rmax[R] = L[d];
                else it->second = std::max(it->second, L[d]);
            }
            for (const auto& kv : rmax) {
                Group g{};
                g.m = kv.second;
                if (g.m > UD_MAXLEV) throw std::runtime_error("sweep level beyond the work record");
                if (((long long)deg << g.m) > p->n[d]) throw std::logic_error("sweep group beyond the 1D matrix");
                long long O = 1, I = 1;
                for (int k = 0; k < D - 1; ++k) (k < d ? O : I) *= deg * cells_of(kv.first[k]);
                if (O * I * ((long long)deg << g.m) >= (1LL << 31)) throw std::runtime_error("sweep group too large");
                g.O = int(O);
                g.I = int(I);
                g.magic = 0;
                if (I > 1) {  // exact when (largest column) * (magic * I - 2^32) < 2^32
                    const unsigned long long mg = (1ULL << 32) / (unsigned long long)I + 1;
                    const unsigned long long e = mg * (unsigned long long)I - (1ULL << 32);
                    if (mg < (1ULL << 32) && (unsigned long long)(O * I) * e < (1ULL << 32)) g.magic = unsigned(mg);
                }
                for (int l = 0; l <= g.m; ++l) {
                    Lev L = kv.first;
                    L.insert(L.begin() + d, l);
                    const long long off = boxoff.at(L);  // present: downward closed
                    if (off + O * I * deg * cells_of(l) > nb) throw std::logic_error("sweep box outside the boxes");
                    g.box.push_back(int(off));
                }
                for (int lt = 0; lt <= g.m; ++lt)
                    for (int ls = 0; ls <= g.m; ++ls) {
                        const long long fl = 2LL * deg * cells_of(lt) * deg * cells_of(ls) * O * I;
                        fl_full[d] += fl;
                        (lt <= ls ? fl_down[d] : fl_up[d]) += fl;
                    }
                groups.push_back(g);
            }
        }
        group0[D] = int(groups.size());
        for (const Group& g : groups)  // the device column division is exact on every column
            if (g.magic)
                for (long long f : {0LL, (long long)g.O * g.I - 1, (long long)g.I - 1, (long long)g.I})
                    if (f < (long long)g.O * g.I &&
                        (long long)(((unsigned long long)f * g.magic) >> 32) != f / g.I)
                        throw std::logic_error("sweep column division");
        // ---- arena: slots 1..2^(D-1)-1 of every term (down outputs; later the up outputs,
        // from slot 0 on), then slots 0..2^(D-1)-1 of every term (whole operators).  Slot 0
        // of the down / whole-operator inputs is x, read through the box map.
        const long long half = 1LL << (D - 1);
        const long long zb = 0;
        const long long gb = zb + std::max<long long>(half - 1, 1) * T * nb;
        sw->arena = gb + half * T * nb;
        if (sw->arena > (1LL << 40)) throw std::runtime_error("sweep arena too large");
        // ---- ops and their work records
        std::vector<std::vector<int>> recs;  // one per wave, UD_REC ints
        // branches per tile at most (1, 2 or 4; KMX_SWEEP_BRANCHES overrides): sharing a
        // tile between more branches pays where there are many (d >= 4); on 3D lev 7
        // deg 3 two and one time alike, four is 10 % slower (longer per-wave chains)
        int max_nbr = D >= 4 ? 4 : 2;
        if (const char* e = std::getenv("KMX_SWEEP_BRANCHES")) {
            const int v = std::atoi(e);
            if (v == 1 || v == 2 || v == 4) max_nbr = v;
        }
        auto base_op = [&](int d) {
            UdOp a{};
            a.d = d;
            a.deg = deg;
            a.lda = p->n[d];
            a.T = T;
            a.nb = nb;
            for (int t = 0; t < T; ++t) {
                a.a_off[t] = p->off[t * D + d];
                if (a.a_off[t] < 0 || a.a_off[t] + (long long)p->n[d] * p->n[d] > p->mat_elems)
                    throw std::logic_error("sweep matrix outside the matrices");
                a.scale[t] = 1.0;
            }
            return a;
        };
        std::vector<int> all_terms(T);
        for (int t = 0; t < T; ++t) all_terms[t] = t;
        auto add_op = [&](int part, bool fin, UdOp a, long long nbranch, const std::vector<int>& terms) {
            a.part = part;
            a.final = fin ? 1 : 0;
            sw->desc.push_back(kmx_sweep::OpDesc{part, fin, a, nbranch, terms});
        };
        long long flops = 0;
        // separable terms: each acts in one dimension (a term that acts in none is
        // taken as acting in the first); the terms of dimension d share one whole-
        // operator op on d.  A factor counts as the identity when the caller declared
        // it so or it is within identity_tol of I; FP64 applies need every undeclared
        // one within 4 ulps.  The identity dimensions' connections must hold every
        // present 1D index (mask[I][I] = 1), since the separable form keeps only J = I
        // there.
        const double exact_tol = 4.0 * 2.220446049250313e-16;
        const double near_tol = env_double("KMX_SWEEP_IDENTITY_TOL", 1e-11);
        std::vector<char> ident(size_t(T) * D, 0);
        std::vector<int> active(T, 0);
        bool separable = !p->h_blk64.empty();
        if (const char* e = std::getenv("KMX_SWEEP_SEPARABLE")) separable = separable && std::atoi(e) != 0;
        bool fp64 = true;
        double idev = 0.0;
        for (int t = 0; t < T && separable; ++t) {
            int na = 0;
            for (int k = 0; k < D; ++k) {
                const bool declared = !p->declared_identity.empty() && p->declared_identity[size_t(t) * D + k];
                const double dev = declared ? 0.0 : identity_dev(p, t, k);
                ident[size_t(t) * D + k] = (declared || dev <= near_tol) ? 1 : 0;
                if (!ident[size_t(t) * D + k]) {
                    active[t] = k;
                    ++na;
                } else if (!declared) {
                    idev = std::max(idev, dev);
                    if (dev > exact_tol) fp64 = false;
                }
            }
            separable = na <= 1;
        }
        if (separable) {  // every present 1D index connected to itself in the identity dimensions
            std::vector<char> idim(D, 0);
            for (int t = 0; t < T; ++t)
                for (int k = 0; k < D; ++k)
                    if (ident[size_t(t) * D + k] && k != active[t]) idim[k] = 1;
            for (int e = 0; e < p->N && separable; ++e)
                for (int k = 0; k < D; ++k) {
                    if (!idim[k]) continue;
                    const int I = p->h_idx1d[size_t(e) * D + k];
                    if (!bs->mask[size_t(k)][size_t(I) * p->nb[k] + I]) {
                        separable = false;
                        break;
                    }
                }
        }
        if (separable) {
            sw->separable = true;
            sw->identity = ident;
            sw->fp64 = fp64;
            sw->identity_dev = idev;
            sw->drop_tol = env_double("KMX_SWEEP_DROP_TOL", 1e-15);
            int nsep = 0;
            for (int d = 0; d < D; ++d) {
                std::vector<int> ts;
                for (int t = 0; t < T; ++t)
                    if (active[t] == d) ts.push_back(t);
                if (ts.empty()) continue;
                UdOp a = base_op(d);
                a.T = 1;  // the dimension's terms, folded into one combined operator
                for (int i = 0; i < KMX_MAX_TERMS; ++i) {
                    a.a_off[i] = 0;  // op i's storage offset, set when its runs are laid out
                    a.scale[i] = 1.0;
                }
                a.x0 = 0;
                a.in_base = 0;                            // slot 0: the box copy of x
                a.out_base = (1 + nsep) * (long long)nb;  // slot 1 + k: op k's sum
                ++nsep;
                add_op(UD_SEP, false, a, 1, ts);
            }
        } else if (D == 1) {
            UdOp a = base_op(0);
            a.x0 = 1;
            add_op(UD_FULL, true, a, 1, all_terms);
            flops += fl_full[0] * T;
        } else {
            for (int d = 0; d + 1 < D; ++d) {  // down: slot b -> slot b | 2^d
                UdOp a = base_op(d);
                a.x0 = 1;
                a.in_base = zb - T * nb;
                a.out_base = zb - T * nb;
                a.out_bit = 1 << d;
                add_op(UD_DOWN, false, a, 1LL << d, all_terms);
                flops += fl_down[d] * T * (1LL << d);
            }
            {  // whole operator on dimension D-1: slot b -> whole-operator slot b
                UdOp a = base_op(D - 1);
                a.x0 = 1;
                a.in_base = zb - T * nb;
                a.out_base = gb;
                add_op(UD_FULL, false, a, half, all_terms);
                flops += fl_full[D - 1] * T * half;
            }
            long long src = gb;  // up: slot b + slot b | 2^d -> slot b of the other array
            for (int d = D - 2; d >= 0; --d) {
                const long long dst = src == gb ? zb : gb;
                UdOp a = base_op(d);
                a.in_base = src;
                a.add_base = src;
                a.add_bit = 1 << d;
                a.out_base = dst;
                add_op(UD_UP, d == 0, a, 1LL << d, all_terms);
                flops += fl_up[d] * T * (1LL << d);
                src = dst;
            }
        }
        if (sw->separable) sw->arena = (1 + (long long)sw->desc.size()) * nb;  // x's box copy, one slot per op
        sw->flops = flops;
        sw->mfma_flops = flops;
        sw->deg = deg;
        sw->T = T;
        sw->max_nbr = max_nbr;
        sw->groups = groups;
        sw->group0 = group0;
        // ---- box maps: x -> boxes (read side) and boxes -> y
        std::vector<int> xmap(size_t(nb), -1);
        std::vector<int> ymap(size_t(nb), -1);
        for (int e = 0; e < p->N; ++e) {
            const int blk = bs->eblk[e];
            const int* bi = &bs->binfo[size_t(blk) * (2 * D + 1)];
            Lev L(bs->blev.begin() + size_t(blk) * D, bs->blev.begin() + size_t(blk + 1) * D);
            std::vector<long long> cell(D), stride(D);
            long long ci = bs->ecell[e];
            for (int k = D - 1; k >= 0; --k) {
                if ((1LL << bi[k]) != cells_of(L[k]) || bi[D + k] != start_of(L[k]))
                    throw std::logic_error("sweep block layout");
                cell[k] = ci & ((1LL << bi[k]) - 1);
                ci >>= bi[k];
            }
            long long sd = 1;
            for (int k = D - 1; k >= 0; --k) {
                stride[k] = sd;
                sd *= deg * cells_of(L[k]);
            }
            const long long b0 = boxoff.at(L);
            for (int a = 0; a < p->B; ++a) {
                long long pos = b0;
                int rest = a;
                for (int k = D - 1; k >= 0; --k) {
                    pos += (cell[k] * deg + rest % deg) * stride[k];
                    rest /= deg;
                }
                if (pos < 0 || pos >= nb || xmap[size_t(pos)] != -1) throw std::logic_error("sweep map collision");
                xmap[size_t(pos)] = e * p->B + a;
                ymap[size_t(pos)] = e * p->B + a;
            }
        }
        auto up = [](const auto* h, size_t n, auto*& d) {
            if (hipMalloc(&d, std::max<size_t>(1, n) * sizeof(*h)) != hipSuccess) throw std::bad_alloc();
            if (n && hipMemcpy(d, h, n * sizeof(*h), hipMemcpyHostToDevice) != hipSuccess)
                throw std::runtime_error("hipMemcpy failed");
        };
        up(xmap.data(), xmap.size(), sw->d_xmap);
        up(ymap.data(), ymap.size(), sw->d_ymap);
        if (hipMalloc(&sw->d_arena[0], size_t(sw->arena) * sizeof(double)) != hipSuccess) throw std::bad_alloc();
        if (bs->d_mat32 && hipMalloc(&sw->d_arena[1], size_t(sw->arena) * sizeof(float)) != hipSuccess)
            throw std::bad_alloc();
        p->sw = sw;
        if (sw->separable) {
            // the box copy of x: absent cells stay zero (the gathers write present cells only)
            if (hipMemset(sw->d_arena[0], 0, size_t(nb) * sizeof(double)) != hipSuccess ||
                (sw->d_arena[1] && hipMemset(sw->d_arena[1], 0, size_t(nb) * sizeof(float)) != hipSuccess))
                throw std::runtime_error("hipMemset failed");
            sep_build(p, *sw);
        } else {
            emit_records(*sw, sweep_wave_batches(1), recs, sw->ops);
            check_records(*sw, recs, sw->ops);
            // records interleaved so that lane j reads ints j and 64 + j as one int2
            std::vector<int> flat(recs.size() * UD_REC);
            for (size_t i = 0; i < recs.size(); ++i)
                for (int j = 0; j < UD_REC / 2; ++j) {
                    flat[i * UD_REC + 2 * j] = recs[i][j];
                    flat[i * UD_REC + 2 * j + 1] = recs[i][UD_REC / 2 + j];
                }
            up(flat.data(), flat.size(), sw->d_recs);
        }
    } catch (const std::exception&) {
        p->sw = sw;
        kmx_sweep_free(p);
    }
}

int kmx_sweep_apply(kmx_plan* p, int slot, const double* x, double* y, hipStream_t s) {
    kmx_sweep* sw = p->sw;
    if (sw->separable) {  // x's box copy, every op in one launch, the slots summed into y
        if (!sw->self) {
            set_error("kmx_apply: separable sweep without its launch tables");
            return KMX_E_ARG;
        }
        if (slot == 0 && !sw->fp64) {
            set_error("kmx_apply: this plan's separable sweep takes near-identity factors as the identity "
                      "(FP32 applies only); declare them in kmx_plan_desc.identity or switch the FP64 sweep off");
            return KMX_E_UNSUPPORTED;
        }
        BatchIO io;
        std::memset(&io, 0, sizeof(io));
        io.x[0] = x;
        io.y[0] = y;
        return kmx_sweep_batch_apply(sw->self, slot, io, s);
    }
    if (sw->stale) {
        set_error("kmx_apply: a 1D factor the dimension sweep takes as the identity was replaced "
                  "(kmx_plan_update_term); switch the sweep off (kmx_plan_set_sweep) or recreate the plan");
        return KMX_E_UNSUPPORTED;
    }
    kmx_blockset* bs = p->bs;
    auto run = [&](auto* arena, const auto* mat) {
        using T = std::remove_pointer_t<decltype(arena)>;
        for (const auto& op : sw->ops) {
            UdOp a = op.a;
            if (op.final)  // the ops that scale their terms
                for (int i = 0; i < a.T; ++i) a.scale[i] = p->scale[op.terms[size_t(i)]];
            const dim3 grid(unsigned(op.wg1 - op.wg0)), block(UD_WAVES * 64);
            const int2* rc = reinterpret_cast<const int2*>(sw->d_recs) + size_t(op.wg0) * UD_WAVES * (UD_REC / 2);
            auto pick = [&](auto nbr_c) {
                constexpr int NBR = decltype(nbr_c)::value;
                return op.part == UD_DOWN ? k_ud_op<T, UD_DOWN, false, NBR>
                       : op.part == UD_FULL ? (op.final ? k_ud_op<T, UD_FULL, true, 1> : k_ud_op<T, UD_FULL, false, NBR>)
                                            : (op.final ? k_ud_op<T, UD_UP, true, 1> : k_ud_op<T, UD_UP, false, NBR>);
            };
            auto* kern = op.nbr == 1   ? pick(std::integral_constant<int, 1>{})
                         : op.nbr == 2 ? pick(std::integral_constant<int, 2>{})
                                       : pick(std::integral_constant<int, 4>{});
            hipLaunchKernelGGL(kern, grid, block, 0, s, mat, arena, rc, a, x, sw->d_xmap, y, sw->d_ymap);
        }
    };
    if (slot == 0) {
        run(static_cast<double*>(sw->d_arena[0]), bs->d_mat64);
    } else {
        if (!sw->d_arena[1] || !bs->d_mat32) {
            set_error("kmx_apply: plan has no FP32 blocks (create it with KMX_PLAN_FP32)");
            return KMX_E_ARG;
        }
        run(static_cast<float*>(sw->d_arena[1]), bs->d_mat32);
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_error(std::string("sweep launch: ") + hipGetErrorString(err));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

void kmx_sweep_free(kmx_plan* p) {
    kmx_sweep* sw = p->sw;
    if (!sw) return;
    kmx_sweep_batch_free(sw->self);
    kmx_sweep_batch_free(sw->selfT);
    sep_side_free(sw->side[0]);
    sep_side_free(sw->side[1]);
    for (SepSide& r : sw->retired) sep_side_free(r);
    for (kmx_sweep_batch* b : sw->retired_batches) kmx_sweep_batch_free(b);
    if (sw->d_xmap) hipFree(sw->d_xmap);
    if (sw->d_ymap) hipFree(sw->d_ymap);
    if (sw->d_recs) hipFree(sw->d_recs);
    for (auto& a : sw->d_arena)
        if (a) hipFree(a);
    delete sw;
    p->sw = nullptr;
}

double kmx_sweep_flops(const kmx_plan* p) { return p->sw ? double(p->sw->flops) : 0.0; }

bool kmx_sweep_fp64_ok(const kmx_plan* p) { return p->sw && (!p->sw->separable || p->sw->fp64); }

void kmx_sweep_get_info(const kmx_plan* p, kmx_sweep_info* o) {
    std::memset(o, 0, sizeof(*o));
    const kmx_sweep* sw = p->sw;
    if (!sw) return;
    o->available = 1;
    o->separable = sw->separable ? 1 : 0;
    o->fp64 = kmx_sweep_fp64_ok(p) ? 1 : 0;
    o->identity_dev = sw->identity_dev;
    o->drop_tol = sw->drop_tol;
    o->flops = double(sw->flops);
    o->mfma_flops = double(sw->mfma_flops);
    o->kept_blocks = sw->separable ? sw->kept_blocks : 0;
    o->total_blocks = sw->separable ? sw->total_blocks : 0;
    o->operator_values = sw->separable ? sw->side[0].a_elems : p->mat_elems;
    o->box_values = sw->nb;
}


// ------------------------------------------------------------------ batches
struct kmx_sweep_batch {
    struct Phase {
        int wg0, nwg, maxnbr;  // first workgroup, workgroups, most branches per tile
        bool sep_only;         // every op separable: the separable kernel alone
    };
    std::vector<Phase> phases;
    std::vector<const kmx_plan*> members;
    BoxJob* d_wgjobs = nullptr;  // separable members' box copies of x and slot sums: each workgroup's job
    int njobs = 0, box_wgs = 0;
    BoxJob* d_wgjobs4 = nullptr;  // the same jobs in blocks of 1024 values (k_box_rk3v)
    int box_wgs4 = 0;
    UdBatchOp* d_ops = nullptr;
    int* d_wgop = nullptr;  // op of every workgroup, phases in order
    int* d_recs = nullptr;  // UD_WAVES work records per workgroup
    bool has32 = true;
    bool has16 = true;     // every op separable with BF16 values (BF16 applies on the FP32 arenas)
    bool all_sep = true;   // every member separable (the fused RK3 step)
    long long kdof = 0;    // values of each RK3 stage derivative (every member's box)
    // RK3 workspace (first kmx_sweep_batch_rk3): k1, k2, the box-order state (kdof FP64
    // values each), then the present-cell flags (kdof bytes)
    double* d_k = nullptr;
};

kmx_sweep_batch* kmx_sweep_batch_build(const std::vector<kmx_plan*>& members, int wave_batches, bool transpose) {
    if (members.empty() || members.size() > KMX_BATCH_MAX) throw std::invalid_argument("sweep batch size");
    for (const kmx_plan* p : members)
        if (transpose && (!p || !p->sw || !p->sw->separable))
            throw std::invalid_argument("the adjoint sweep needs the separable form");
    // load batches per wave (0: the batch default, four)
    const int wb_sep = wave_batches > 0 ? wave_batches : sweep_wave_batches(4);
    const int wb_gen = wave_batches > 0 ? wave_batches : sweep_wave_batches(4);
    const int side = transpose ? 1 : 0;
    for (size_t q = 0; q < members.size(); ++q) {
        const kmx_plan* p = members[q];
        if (!p || !p->sw || p->sw->desc.empty()) throw std::invalid_argument("sweep batch member without a sweep");
        for (size_t r = 0; r < q; ++r)
            if (members[r] == p) throw std::invalid_argument("sweep batch member twice (one arena per plan)");
    }
    // every member's records for this batch's split; a member's ops form launch
    // groups: one op each (general sweeps, in order) or all ops together (separable)
    std::vector<std::vector<std::vector<int>>> recs(members.size());
    std::vector<std::vector<kmx_sweep::Op>> mops(members.size());
    std::vector<int> ngroups(members.size());
    int nph = 0;
    for (size_t q = 0; q < members.size(); ++q) {
        const kmx_sweep& sw = *members[q]->sw;
        emit_records(sw, sw.separable ? wb_sep : wb_gen, recs[q], mops[q], side);
        check_records(sw, recs[q], mops[q], side);
        ngroups[q] = sw.separable ? 1 : int(mops[q].size());
        nph = std::max(nph, ngroups[q]);
    }
    auto* sb = new kmx_sweep_batch();
    sb->members.assign(members.begin(), members.end());
    try {
        std::vector<UdBatchOp> ops;
        std::vector<int> wgop;
        std::vector<int> flat;  // records interleaved so that lane j reads ints j and 64 + j as one int2
        for (int ph = 0; ph < nph; ++ph) {  // groups end-aligned: every member's last group in the last phase
            kmx_sweep_batch::Phase P{int(wgop.size()), 0, 1, true};
            for (size_t q = 0; q < members.size(); ++q) {
                kmx_plan* p = members[q];
                const kmx_sweep* sw = p->sw;
                const int gi = ph - (nph - ngroups[q]);
                if (gi < 0) continue;
                const size_t k0 = sw->separable ? 0 : size_t(gi), k1 = sw->separable ? mops[q].size() : k0 + 1;
                for (size_t k = k0; k < k1; ++k) {
                    const kmx_sweep::Op& op = mops[q][k];
                    UdBatchOp o{};
                    o.a = op.a;
                    if (op.final)  // the ops that scale their terms (a separable op's scales are in its storage)
                        for (int i = 0; i < o.a.T; ++i) o.a.scale[i] = p->scale[op.terms[size_t(i)]];
                    const bool sep = op.part == UD_SEP;
                    o.mat[0] = sep ? static_cast<const void*>(sw->side[side].d_a64) : p->bs->d_mat64;
                    o.mat[1] = sep ? static_cast<const void*>(sw->side[side].d_a32) : p->bs->d_mat32;
                    o.mat[2] = sep ? static_cast<const void*>(sw->side[side].d_a16) : nullptr;
                    o.codes = sep ? sw->side[side].d_codes : nullptr;
                    o.quad = sep && sw->side[side].quad[k] ? 1 : 0;
                    if (sep && !sw->side[side].d_a32) sb->has32 = false;
                    if (!sep || !sw->side[side].d_a16 || !sw->d_arena[1]) sb->has16 = false;
                    o.arena[0] = sw->d_arena[0];
                    o.arena[1] = sw->d_arena[1];
                    if (!sw->d_arena[1] || !p->bs->d_mat32) sb->has32 = false;
                    o.xmap = sw->d_xmap;
                    o.ymap = sw->d_ymap;
                    o.inst = int(q);
                    o.kind = ud_kind(op.part, op.final, op.nbr);
                    if (ud_kind_nbr(o.kind) != op.nbr || op.wg1 < op.wg0) throw std::logic_error("sweep batch op");
                    P.maxnbr = std::max(P.maxnbr, op.nbr);
                    if (op.part != UD_SEP || op.nbr != 1) P.sep_only = false;
                    const int oi = int(ops.size());
                    ops.push_back(o);
                    for (int wg = op.wg0; wg < op.wg1; ++wg) {
                        wgop.push_back(oi);
                        for (int w = 0; w < UD_WAVES; ++w) {
                            const std::vector<int>& r = recs[q][size_t(wg) * UD_WAVES + w];
                            for (int j = 0; j < UD_REC / 2; ++j) {
                                flat.push_back(r[j]);
                                flat.push_back(r[UD_REC / 2 + j]);
                            }
                        }
                    }
                    P.nwg += op.wg1 - op.wg0;
                    if (wgop.size() >= size_t(1) << 30) throw std::runtime_error("sweep batch too large");
                }
            }
            if (P.sep_only && P.nwg > 1) {
                // Workgroups in classes of their longest wave's steps (powers of two), longer
                // classes first, emission order (instance by instance) inside a class: the long
                // tiles of every instance start early without scattering each instance's tiles
                // over the whole launch.  (Dealing each op's workgroups to one XCD, balanced by
                // work, was measured 1-7 % slower on the 64-instance batch; runs of
                // sep_xcd_block() neighbouring workgroups per XCD are faster.)
                const size_t rw = size_t(UD_WAVES) * UD_REC;  // ints per workgroup
                auto rec_int = [&](size_t wg, int w, int k) {  // int k of wave w's record
                    const size_t b = wg * rw + size_t(w) * UD_REC;
                    return k < UD_REC / 2 ? flat[b + 2 * size_t(k)] : flat[b + 2 * size_t(k - UD_REC / 2) + 1];
                };
                std::vector<int> cls(size_t(P.nwg), 0);
                for (int g = 0; g < P.nwg; ++g) {
                    long long most = 0;
                    for (int w = 0; w < UD_WAVES; ++w) {
                        const size_t wg = size_t(P.wg0 + g);
                        if (rec_int(wg, w, R_GSIZE) == 0 || rec_int(wg, w, R_TN) == 0) continue;
                        const long long st = (long long)(rec_int(wg, w, R_SHI) - rec_int(wg, w, R_SLO)) *
                                             std::max(1, rec_int(wg, w, R_NCT));
                        most = std::max(most, st);
                    }
                    while (most > 0) {
                        ++cls[size_t(g)];
                        most >>= 1;
                    }
                }
                std::vector<int> idx(size_t(P.nwg));
                for (int g = 0; g < P.nwg; ++g) idx[size_t(g)] = g;
                std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cls[size_t(a)] > cls[size_t(b)]; });
                const int xb = sep_xcd_block();
                if (xb > 1) {
                    // runs of xb workgroups in list order to one XCD (launch position p runs
                    // on XCD p % 8): in each block of 8 xb, run y goes to positions
                    // y, y + 8, ..., so neighbours in the list -- the column tiles of one
                    // panel -- read its operator values through one L2
                    const size_t blk = size_t(8) * size_t(xb);
                    std::vector<int> idx2(idx.size());
                    for (size_t r = 0; r < idx.size(); ++r) {
                        size_t pos = r;
                        const size_t b = r / blk, o = r % blk;
                        if ((b + 1) * blk <= idx.size()) pos = b * blk + (o % size_t(xb)) * 8 + o / size_t(xb);
                        idx2[pos] = idx[r];
                    }
                    idx.swap(idx2);
                }
                std::vector<int> wg2(size_t(P.nwg)), fl2(size_t(P.nwg) * rw);
                for (int g = 0; g < P.nwg; ++g) {
                    const size_t src = size_t(P.wg0 + idx[size_t(g)]);
                    wg2[size_t(g)] = wgop[src];
                    std::copy(flat.begin() + long(src * rw), flat.begin() + long((src + 1) * rw),
                              fl2.begin() + long(size_t(g) * rw));
                }
                std::copy(wg2.begin(), wg2.end(), wgop.begin() + P.wg0);
                std::copy(fl2.begin(), fl2.end(), flat.begin() + long(size_t(P.wg0) * rw));
            }
            sb->phases.push_back(P);
        }
        if (flat.size() != wgop.size() * UD_WAVES * UD_REC) throw std::logic_error("sweep batch records");
        std::vector<BoxJob> jobs, wgjobs;
        long long blk = 0, koff = 0, blk4 = 0;
        std::vector<BoxJob> wgjobs4;
        for (size_t q = 0; q < members.size(); ++q) {
            const kmx_sweep* sw = members[q]->sw;
            if (!sw->separable) {
                sb->all_sep = false;
                continue;
            }
            jobs.push_back(BoxJob{sw->d_xmap, sw->d_ymap, {sw->d_arena[0], sw->d_arena[1]}, sw->nb, koff,
                                  int(sw->desc.size()), int(q), int(blk)});
            koff += (sw->nb + 3) / 4 * 4;  // four-value alignment of the stage workspace
            blk += (sw->nb + 255) / 256;
            if (blk >= (1LL << 31)) throw std::runtime_error("sweep batch too large");
            wgjobs.resize(size_t(blk), jobs.back());
            BoxJob j4 = jobs.back();
            j4.blk0 = int(blk4);
            blk4 += (sw->nb + 1023) / 1024;
            wgjobs4.resize(size_t(blk4), j4);
        }
        sb->njobs = int(jobs.size());
        sb->box_wgs = int(blk);
        sb->box_wgs4 = int(blk4);
        sb->kdof = koff;
        auto up = [](const auto* h, size_t n, auto*& d) {
            if (hipMalloc(&d, std::max<size_t>(1, n) * sizeof(*h)) != hipSuccess) throw std::bad_alloc();
            if (n && hipMemcpy(d, h, n * sizeof(*h), hipMemcpyHostToDevice) != hipSuccess)
                throw std::runtime_error("hipMemcpy failed");
        };
        up(ops.data(), ops.size(), sb->d_ops);
        up(wgop.data(), wgop.size(), sb->d_wgop);
        up(flat.data(), flat.size(), sb->d_recs);
        up(wgjobs.data(), wgjobs.size(), sb->d_wgjobs);
        up(wgjobs4.data(), wgjobs4.size(), sb->d_wgjobs4);
    } catch (...) {
        kmx_sweep_batch_free(sb);
        throw;
    }
    return sb;
}

// A separable phase: one launch of k_sep_batch over its workgroups.
static void launch_sep_phase(const kmx_sweep_batch* sb, const kmx_sweep_batch::Phase& P, int slot, hipStream_t s) {
    const int* wgop = sb->d_wgop + P.wg0;
    const int2* rc = reinterpret_cast<const int2*>(sb->d_recs) + size_t(P.wg0) * UD_WAVES * (UD_REC / 2);
    const dim3 grid(static_cast<unsigned>(P.nwg)), block(UD_WAVES * 64);
    if (slot == 0) hipLaunchKernelGGL((k_sep_batch<double, double>), grid, block, 0, s, sb->d_ops, wgop, rc);
    else if (slot == 1) hipLaunchKernelGGL((k_sep_batch<float, float>), grid, block, 0, s, sb->d_ops, wgop, rc);
    else hipLaunchKernelGGL((k_sep_batch<float, bf16_bits>), grid, block, 0, s, sb->d_ops, wgop, rc);
}

int kmx_sweep_batch_apply(const kmx_sweep_batch* sb, int slot, const BatchIO& io, hipStream_t s) {
    for (const kmx_plan* p : sb->members)
        if (p->sw && p->sw->stale) {
            set_error("kmx_apply: a 1D factor the dimension sweep takes as the identity was replaced "
                      "(kmx_plan_update_term); switch the sweep off and on again, or recreate the plan");
            return KMX_E_UNSUPPORTED;
        }
    if (slot == 1 && !sb->has32) {
        set_error("kmx_batch_apply: a sweep member has no FP32 blocks");
        return KMX_E_ARG;
    }
    if (slot == 2 && !sb->has16) {
        set_error("kmx_apply: BF16 runs on separable sweeps of plans with FP32 blocks only");
        return KMX_E_UNSUPPORTED;
    }
    const dim3 bg(unsigned(sb->box_wgs)), bb(256);
    const bool one = sb->members.size() == 1;
    const IoOne io1{io.x[0], io.y[0]};
    if (sb->njobs > 0) {
        if (one) {
            if (slot == 0) hipLaunchKernelGGL((k_box_gather_batch<double, IoOne>), bg, bb, 0, s, sb->d_wgjobs, io1);
            else hipLaunchKernelGGL((k_box_gather_batch<float, IoOne>), bg, bb, 0, s, sb->d_wgjobs, io1);
        } else {
            if (slot == 0) hipLaunchKernelGGL((k_box_gather_batch<double, BatchIO>), bg, bb, 0, s, sb->d_wgjobs, io);
            else hipLaunchKernelGGL((k_box_gather_batch<float, BatchIO>), bg, bb, 0, s, sb->d_wgjobs, io);
        }
    }
    for (const auto& P : sb->phases) {
        if (P.nwg == 0) continue;
        const int* wgop = sb->d_wgop + P.wg0;
        const int2* rc = reinterpret_cast<const int2*>(sb->d_recs) + size_t(P.wg0) * UD_WAVES * (UD_REC / 2);
        const dim3 grid(unsigned(P.nwg)), block(UD_WAVES * 64);
        if (P.sep_only) {
            launch_sep_phase(sb, P, slot, s);
        } else if (slot == 0) {
            if (P.maxnbr <= 2) hipLaunchKernelGGL((k_ud_batch<double, 2>), grid, block, 0, s, sb->d_ops, wgop, rc, io);
            else hipLaunchKernelGGL((k_ud_batch<double, 4>), grid, block, 0, s, sb->d_ops, wgop, rc, io);
        } else {
            if (P.maxnbr <= 2) hipLaunchKernelGGL((k_ud_batch<float, 2>), grid, block, 0, s, sb->d_ops, wgop, rc, io);
            else hipLaunchKernelGGL((k_ud_batch<float, 4>), grid, block, 0, s, sb->d_ops, wgop, rc, io);
        }
    }
    if (sb->njobs > 0) {
        if (one) {
            if (slot == 0) hipLaunchKernelGGL((k_box_sum_batch<double, IoOne>), bg, bb, 0, s, sb->d_wgjobs, io1);
            else hipLaunchKernelGGL((k_box_sum_batch<float, IoOne>), bg, bb, 0, s, sb->d_wgjobs, io1);
        } else {
            if (slot == 0) hipLaunchKernelGGL((k_box_sum_batch<double, BatchIO>), bg, bb, 0, s, sb->d_wgjobs, io);
            else hipLaunchKernelGGL((k_box_sum_batch<float, BatchIO>), bg, bb, 0, s, sb->d_wgjobs, io);
        }
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_error(std::string("sweep batch launch: ") + hipGetErrorString(err));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

int kmx_sweep_batch_rk3(kmx_sweep_batch* sb, int slot, double* const* fs, const double* dts, int nsteps,
                        hipStream_t s) {
    if (!sb->all_sep || sb->njobs != int(sb->members.size())) {
        set_error("kmx_batch_rk3: every plan of the batch must run a separable sweep in this precision");
        return KMX_E_UNSUPPORTED;
    }
    if (slot == 1 && !sb->has32) {
        set_error("kmx_batch_rk3: a sweep member has no FP32 blocks");
        return KMX_E_ARG;
    }
    if (slot == 2 && !sb->has16) {
        set_error("kmx_batch_rk3: BF16 runs on separable sweeps of plans with FP32 blocks only");
        return KMX_E_UNSUPPORTED;
    }
    for (const kmx_plan* p : sb->members)
        if (p->sw->stale || (slot == 0 && !p->sw->fp64)) {
            set_error("kmx_batch_rk3: a member's sweep cannot serve this precision (replaced identity factor, or "
                      "near-identity factors under FP64)");
            return KMX_E_UNSUPPORTED;
        }
    const dim3 bg(unsigned(sb->box_wgs)), bb(256);
    unsigned char* present = nullptr;
    if (sb->kdof > 0) {
        const bool fresh = !sb->d_k;
        if (fresh && hipMalloc(&sb->d_k, size_t(3 * sb->kdof) * sizeof(double) + size_t(sb->kdof)) != hipSuccess) {
            set_error("kmx_batch_rk3: allocating the stage workspace failed");
            return KMX_E_NOMEM;
        }
        present = reinterpret_cast<unsigned char*>(sb->d_k + 3 * sb->kdof);
        if (fresh) hipLaunchKernelGGL(k_box_present, bg, bb, 0, s, sb->d_wgjobs, present);
    }
    Rk3Args a;
    std::memset(&a, 0, sizeof(a));
    for (size_t q = 0; q < sb->members.size(); ++q) {
        a.f[q] = fs[q];
        a.dt[q] = dts[q];
    }
    a.k1 = sb->d_k;
    if (slot == 0)
        a.k2 = sb->d_k + sb->kdof;
    else
        a.k2 = reinterpret_cast<float*>(sb->d_k) + sb->kdof;  // FP32 stage values
    a.fb = sb->d_k + 2 * sb->kdof;
    a.present = present;
    auto ops = [&]() {
        for (const auto& P : sb->phases) {
            if (P.nwg == 0) continue;
            launch_sep_phase(sb, P, slot, s);
        }
    };
    auto stage0 = [&](auto kern) { hipLaunchKernelGGL(kern, bg, bb, 0, s, sb->d_wgjobs, a, 0); };
    const bool vec = rk3_vec_on();
    const dim3 bg4(unsigned(std::max(1, sb->box_wgs4)));
    auto stage = [&](auto kern, auto kernv, int next) {
        if (vec) hipLaunchKernelGGL(kernv, bg4, bb, 0, s, sb->d_wgjobs4, a, next);
        else hipLaunchKernelGGL(kern, bg, bb, 0, s, sb->d_wgjobs, a, next);
    };
    if (slot == 0) stage0(k_box_rk3<double, 0>);
    else stage0(k_box_rk3<float, 0>);
    for (int st = 0; st < nsteps; ++st) {
        const int next = st + 1 < nsteps ? 1 : 0;
        ops();
        if (slot == 0) stage(k_box_rk3<double, 1>, k_box_rk3v<double, 1>, 0);
        else stage(k_box_rk3<float, 1>, k_box_rk3v<float, 1>, 0);
        ops();
        if (slot == 0) stage(k_box_rk3<double, 2>, k_box_rk3v<double, 2>, 0);
        else stage(k_box_rk3<float, 2>, k_box_rk3v<float, 2>, 0);
        ops();
        if (slot == 0) stage(k_box_rk3<double, 3>, k_box_rk3v<double, 3>, next);
        else stage(k_box_rk3<float, 3>, k_box_rk3v<float, 3>, next);
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_error(std::string("kmx_batch_rk3 launch: ") + hipGetErrorString(err));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

void kmx_sweep_batch_free(kmx_sweep_batch* sb) {
    if (!sb) return;
    if (sb->d_k) hipFree(sb->d_k);
    if (sb->d_ops) hipFree(sb->d_ops);
    if (sb->d_wgop) hipFree(sb->d_wgop);
    if (sb->d_recs) hipFree(sb->d_recs);
    if (sb->d_wgjobs) hipFree(sb->d_wgjobs);
    if (sb->d_wgjobs4) hipFree(sb->d_wgjobs4);
    delete sb;
}

int kmx_sweep_batch_phases(const kmx_sweep_batch* sb) { return sb ? int(sb->phases.size()) : 0; }

bool kmx_sweep_batch_has16(const kmx_sweep_batch* sb) { return sb && sb->has16; }

bool kmx_sweep_bf16_ok(const kmx_plan* p) {
    const kmx_sweep* sw = p ? p->sw : nullptr;
    return sw && sw->separable && sw->self && !sw->stale && sw->d_arena[1] && sw->side[0].d_a16 &&
           sw->self->has16;
}

void kmx_sweep_term_updated(kmx_plan* p, int term, int dim) {
    kmx_sweep* sw = p->sw;
    if (!sw || !sw->separable) return;
    const bool declared = !p->declared_identity.empty() && p->declared_identity[size_t(term) * p->D + dim];
    if (sw->identity[size_t(term) * p->D + dim] && !declared) {
        const double dev = identity_dev(p, term, dim);
        if (dev > env_double("KMX_SWEEP_IDENTITY_TOL", 1e-11)) sw->stale = true;
        else if (dev > 4.0 * 2.220446049250313e-16) sw->fp64 = false;
        sw->identity_dev = std::max(sw->identity_dev, dev);
    }
    for (size_t k = 0; k < sw->desc.size(); ++k) {
        if (sw->desc[k].a.d != dim ||
            std::find(sw->desc[k].terms.begin(), sw->desc[k].terms.end(), term) == sw->desc[k].terms.end())
            continue;
        if (sep_pattern(p, *sw, int(k)) == sw->keep[k]) {  // same kept blocks: new values in place
            sep_values(p, *sw, 0, int(k));
            sep_values(p, *sw, 1, int(k));
        } else {  // the kept blocks moved: new runs, storage and launch tables
            sep_build(p, *sw);
        }
    }
}

int kmx_sweep_scales_updated(kmx_plan* p, int term) {
    kmx_sweep* sw = p->sw;
    if (!sw || !sw->separable) return KMX_OK;
    try {
        for (size_t k = 0; k < sw->desc.size(); ++k)  // the one op whose storage holds the term
            if (std::find(sw->desc[k].terms.begin(), sw->desc[k].terms.end(), term) != sw->desc[k].terms.end()) {
                sep_values(p, *sw, 0, int(k));
                sep_values(p, *sw, 1, int(k));
            }
    } catch (const std::exception& e) {
        set_error(std::string("kmx_plan_set_term_scale: refreshing the separable operator failed: ") + e.what());
        return KMX_E_HIP;
    }
    return KMX_OK;
}

bool kmx_sweep_has_transpose(const kmx_plan* p) { return p->sw && p->sw->selfT; }

int kmx_sweep_apply_transpose(kmx_plan* p, int slot, const double* x, double* y, hipStream_t s) {
    if (slot == 0 && !p->sw->fp64) {
        set_error("kmx_apply_transpose: this plan's separable sweep takes near-identity factors as the identity "
                  "(FP32 applies only)");
        return KMX_E_UNSUPPORTED;
    }
    BatchIO io;
    std::memset(&io, 0, sizeof(io));
    io.x[0] = x;
    io.y[0] = y;
    return kmx_sweep_batch_apply(p->sw->selfT, slot, io, s);
}
