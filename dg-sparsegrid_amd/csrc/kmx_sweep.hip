// kmx_sweep.hip -- the operator apply as d dimension sweeps over the sparse grid.
//
// apply_A.m:100-237 sums, for every target element and every connected source
// element, (A_{t,1}[I_1,J_1] x ... x A_{t,d}[I_d,J_d]) x_j.  When the pair
// list is, between every pair of level groups, exactly the tensor product of
// per-dimension 1D connections (verified at plan time, as connect_nD.m:63-106
// builds it; an all-pairs list trivially) and the set of level tuples is
// downward closed (every sparse or full grid is), the sum factorizes by
// dimension: contract dim d for every source prefix (levels of dims 1..d-1)
// against every target row of dim d, then dim d-1, ..., then dim 1.  Because
// the levels of one dimension are contiguous ranges of its 1D index, each
// contraction is a plain GEMM with the masked 1D matrix:
//
//   stage j, group (p, q)   p = source levels of dims < j, q = target levels of dims > j
//     C[r_j, (p-cells, q-cells)] = sum_{r'_j} Ahat_j[r_j, r'_j] * Z[(p-cells), r'_j, (q-cells)]
//
// with r'_j running over every source level present under prefix p and r_j
// over every target level present above suffix q.  Same terms as the
// pair-by-pair sum, added in another order (7.5x fewer flops on a 3D level-7
// sparse grid with degree 3).  Each GEMM is cut into 16x16 output tiles, one
// wave per tile on the matrix cores (16x16x4 FP64 / FP32); every output is
// written by exactly one wave (no atomics).  Per apply: a gather of x into
// the stage-0 boxes, d stage launches, and a scatter summing the terms.
#include "kmx_internal.h"

#include <algorithm>
#include <map>
#include <set>
#include <stdexcept>
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

// One GEMM of a stage.  A(m, k) = mat[a_off + m + k * lda]; B(k, f) =
// arena[b_off + k * bk + (f % bE0) + (f / bE0) * bs1]; C(m, f) =
// arena[c_off + m * cm + (f % cE0) + ((f / cE0) % cE1) * cs1 + (f / (cE0 * cE1)) * cs2].
struct SwGemm {
    long long a_off, b_off, c_off;
    long long bk, bs1, cs1, cs2;
    int lda, bE0, cm, cE0, cE1;
    int M, N, K;
};

template <typename T, int NC>
__global__ void __launch_bounds__(256) k_sweep_stage(const T* __restrict__ mat, T* __restrict__ arena,
                                                     const SwGemm* __restrict__ gemms, const int4* __restrict__ tiles,
                                                     int ntiles) {
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= ntiles) return;  // whole waves leave together
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int4 tl = tiles[wave];
    const SwGemm g = gemms[tl.x];
    const int m0 = tl.y, f0 = tl.z;
    // operand rows / fibres this lane feeds, clamped in range (fed as zero when outside)
    const int ia = m0 + li;
    const bool aok = ia < g.M;
    const int fb = f0 + li;
    const bool bok = fb < g.N;
    const int fbc = bok ? fb : 0;
    const T* ap = mat + g.a_off + (aok ? ia : 0);
    const T* bp = arena + g.b_off + (fbc % g.bE0) + (long long)(fbc / g.bE0) * g.bs1;
    // K in steps of 4 (lane group lk takes j = 4 s + lk), split into NC
    // independent chains of consecutive steps so that NC loads and NC matrix
    // operations are in flight per wave; the partial tiles are added at the end
    // (fixed order: deterministic).  Steps and columns past K read column / row 0
    // and feed zeros.  Tiles of short K run with NC = 1.
    const int K = g.K;
    const int nsteps = (K + 3) >> 2;
    const int Q = (nsteps + NC - 1) / NC;
    typename SwMfma<T>::V acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = typename SwMfma<T>::V{T(0), T(0), T(0), T(0)};
    const int bk = int(g.bk);
#pragma unroll 2
    for (int u = 0; u < Q; ++u) {
        T av[NC], bv[NC];
        bool ok[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int j = 4 * (c * Q + u) + lk;
            ok[c] = j < K;
            const int jc = ok[c] ? j : 0;
            av[c] = ap[jc * g.lda];
            bv[c] = bp[jc * bk];
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
            acc[c] = SwMfma<T>::run((aok && ok[c]) ? av[c] : T(0), (bok && ok[c]) ? bv[c] : T(0), acc[c]);
    }
#pragma unroll
    for (int c = 1; c < NC; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[0][r] += acc[c][r];
    if (!bok) return;
    const long long cf = (fb % g.cE0) + (long long)((fb / g.cE0) % g.cE1) * g.cs1 +
                         (long long)(fb / (g.cE0 * g.cE1)) * g.cs2;
    T* cp = arena + g.c_off + cf;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = m0 + SwMfma<T>::row(lane, r);
        if (i < g.M) cp[(long long)i * g.cm] = acc[0][r];
    }
}

// stage-0 boxes from x (absent cells: 0)
template <typename T>
__global__ void __launch_bounds__(256) k_sweep_gather(const double* __restrict__ x, const int* __restrict__ map,
                                                      T* __restrict__ out, long long n) {
    for (long long v = blockIdx.x * 256ll + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        const int s = map[v];
        out[v] = s >= 0 ? T(x[s]) : T(0);
    }
}

struct ScatterArgs {
    long long yoff[KMX_MAX_TERMS];
    double scale[KMX_MAX_TERMS];
    int T;
};

// y[v] = sum_t scale_t * Y_t[map[v]]
template <typename T>
__global__ void __launch_bounds__(256) k_sweep_scatter(const T* __restrict__ arena, const long long* __restrict__ map,
                                                       double* __restrict__ y, long long n, ScatterArgs s) {
    for (long long v = blockIdx.x * 256ll + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        const long long m = map[v];
        double acc = 0.0;
        for (int t = 0; t < s.T; ++t) acc += s.scale[t] * double(arena[s.yoff[t] + m]);
        y[v] = acc;
    }
}

unsigned grid_of(long long n) { return unsigned(std::max<long long>(1, std::min<long long>((n + 255) / 256, 65536))); }

}  // namespace
}  // namespace kmx

using namespace kmx;

struct kmx_sweep {
    int D = 0;
    long long arena = 0;  // elements
    long long xn = 0;     // stage-0 box elements (at arena offset 0)
    long long yoff[KMX_MAX_TERMS] = {0};
    std::vector<int> stage_gemm0, stage_tile0;  // per stage: first gemm / tile (D + 1 entries)
    std::vector<int> stage_short;               // per stage: first tile with K < 16 (fewer than 4 steps)
    int* d_xmap = nullptr;
    long long* d_ymap = nullptr;
    SwGemm* d_gemms = nullptr;
    int4* d_tiles = nullptr;
    void* d_arena[2] = {nullptr, nullptr};  // FP64, FP32
    long long flops = 0;                    // useful flops per apply (all terms)
};

namespace {

using Lev = std::vector<int>;

long long cum_cells(int m) { return 1LL << m; }                 // cells in levels 0..m
long long cells_of(int l) { return l == 0 ? 1 : 1LL << (l - 1); }  // cells of level l
long long start_of(int l) { return l == 0 ? 0 : 1LL << (l - 1); }  // first 1D index of level l

struct Builder {
    int D, deg, T;
    std::set<Lev> S;                 // level tuples present
    std::vector<std::set<Lev>> P;    // P[j] = {L[0:j]}
    std::vector<std::set<Lev>> Q;    // Q[j] = {L[j+1:D]}
    long long arena = 0;

    long long fibres(const Lev& l) const {
        long long f = 1;
        for (int v : l) f *= cells_of(v) * deg;
        return f;
    }
    // source levels of dim j present under prefix p (0..max): K rows
    long long Kof(int j, const Lev& p) const {
        int m = -1;
        for (const Lev& L : S)
            if (std::equal(p.begin(), p.end(), L.begin())) m = std::max(m, L[j]);
        if (m < 0) throw std::logic_error("empty prefix");
        return deg * cum_cells(m);
    }
    // target levels of dim j present above suffix q (0..max): M rows
    long long Mof(int j, const Lev& q) const {
        int m = -1;
        for (const Lev& L : S)
            if (std::equal(q.begin(), q.end(), L.begin() + j + 1)) m = std::max(m, L[j]);
        if (m < 0) throw std::logic_error("empty suffix");
        return deg * cum_cells(m);
    }
};

}  // namespace

void kmx_sweep_build(kmx_plan* p) {
    kmx_blockset* bs = p->bs;
    if (!bs || !bs->ok || !bs->full_product || p->T > KMX_MAX_TERMS) return;
    const int D = p->D, deg = p->deg, T = p->T;
    Builder b{D, deg, T, {}, {}, {}, 0};
    for (int k = 0; k < bs->nblk; ++k) b.S.insert(Lev(bs->blev.begin() + size_t(k) * D, bs->blev.begin() + size_t(k + 1) * D));
    // downward closed: lowering any level keeps the tuple in the set
    for (const Lev& L : b.S)
        for (int k = 0; k < D; ++k)
            if (L[k] > 0) {
                Lev M = L;
                --M[k];
                if (!b.S.count(M)) return;
            }
    b.P.resize(D + 1);
    b.Q.resize(D);
    for (const Lev& L : b.S) {
        for (int j = 0; j <= D; ++j) b.P[j].insert(Lev(L.begin(), L.begin() + j));
        for (int j = 0; j < D; ++j) b.Q[j].insert(Lev(L.begin() + j + 1, L.end()));
    }
    auto* sw = new kmx_sweep();
    sw->D = D;
    try {
        // ---- boxes.  O[j] (output of stage j, j = D..0; O[D] = gathered x):
        //   j = D:    key p (levels of dims 0..D-2):  [p fibres] x [r'_{D-1}: K_{D-1}(p)]
        //   0<j<D:    key (ps, q), ps = dims 0..j-2, q = dims j+1..D-1:
        //             [ps fibres] x [r'_{j-1}: K_{j-1}(ps)] x [r_j: M_j(q)] x [q fibres]
        //   j = 0:    key q (dims 1..D-1): [r_0: M_0(q)] x [q fibres]
        long long off = 0;
        std::map<Lev, long long> xbox;  // O[D]
        for (const Lev& pp : b.P[D - 1]) {
            xbox[pp] = off;
            off += b.fibres(pp) * b.Kof(D - 1, pp);
        }
        sw->xn = off;
        // per term: O[j] boxes for j = D-1..0
        std::vector<std::vector<std::map<std::pair<Lev, Lev>, long long>>> obox(
            T, std::vector<std::map<std::pair<Lev, Lev>, long long>>(D));
        for (int t = 0; t < T; ++t) {
            for (int j = D - 1; j >= 1; --j)
                for (const Lev& ps : b.P[j - 1])
                    for (const Lev& q : b.Q[j]) {
                        obox[t][j][{ps, q}] = off;
                        off += b.fibres(ps) * b.Kof(j - 1, ps) * b.Mof(j, q) * b.fibres(q);
                    }
            sw->yoff[t] = off;
            for (const Lev& q : b.Q[0]) {
                obox[t][0][{Lev(), q}] = off;
                off += b.Mof(0, q) * b.fibres(q);
            }
        }
        sw->arena = off;
        // ---- GEMMs and tiles, stage by stage (dim D-1 first)
        std::vector<SwGemm> gemms;
        std::vector<int4> tiles;
        for (int j = D - 1; j >= 0; --j) {
            sw->stage_gemm0.push_back(int(gemms.size()));
            sw->stage_tile0.push_back(int(tiles.size()));
            for (int t = 0; t < T; ++t)
                for (const Lev& pp : b.P[j])
                    for (const Lev& q : b.Q[j]) {
                        SwGemm g{};
                        const long long K = b.Kof(j, pp), M = b.Mof(j, q);
                        const long long pf = b.fibres(pp), qf = b.fibres(q);
                        g.M = int(M);
                        g.K = int(K);
                        if (pf * qf > (1LL << 30)) throw std::runtime_error("stage GEMM too wide");
                        g.N = int(pf * qf);
                        g.a_off = p->off[t * D + j];
                        g.lda = p->n[j];
                        // B: the previous stage's box
                        if (j == D - 1) {
                            g.b_off = xbox.at(pp);
                            g.bk = 1;
                            g.bE0 = 1;
                            g.bs1 = K;
                        } else {
                            // q = (l_{j+1}, qr): box O[j+1] keyed (pp, qr), rows r_{j+1} of level l_{j+1}
                            const int l1 = q[0];
                            const Lev qr(q.begin() + 1, q.end());
                            const long long Mn = b.Mof(j + 1, qr), qrf = b.fibres(qr);
                            g.b_off = obox[t][j + 1].at({pp, qr}) + deg * start_of(l1) * qrf;
                            g.bk = Mn * qrf;
                            g.bE0 = int(cells_of(l1) * deg * qrf);
                            g.bs1 = K * Mn * qrf;
                        }
                        // C: this stage's box
                        g.cm = int(qf);
                        g.cE0 = int(qf);
                        if (j == 0) {
                            g.c_off = obox[t][0].at({Lev(), q});
                            g.cE1 = 1;
                            g.cs1 = 0;
                            g.cs2 = 0;
                        } else {
                            const Lev ps(pp.begin(), pp.end() - 1);
                            const int lp = pp.back();
                            const long long Kp = b.Kof(j - 1, ps);
                            g.c_off = obox[t][j].at({ps, q}) + deg * start_of(lp) * M * qf;
                            g.cE1 = int(cells_of(lp) * deg);
                            g.cs1 = M * qf;
                            g.cs2 = Kp * M * qf;
                        }
                        sw->flops += 2LL * M * K * pf * qf;
                        const int gi = int(gemms.size());
                        gemms.push_back(g);
                        for (long long m0 = 0; m0 < M; m0 += 16)
                            for (long long f0 = 0; f0 < pf * qf; f0 += 16) tiles.push_back(make_int4(gi, int(m0), int(f0), 0));
                    }
            // longest tiles (largest K) first, so that they start in the first wave of the launch
            std::stable_sort(tiles.begin() + sw->stage_tile0.back(), tiles.end(),
                             [&](const int4& u, const int4& v) { return gemms[u.x].K > gemms[v.x].K; });
            int sh = sw->stage_tile0.back();
            while (sh < int(tiles.size()) && gemms[tiles[sh].x].K >= 16) ++sh;
            sw->stage_short.push_back(sh);
        }
        sw->stage_gemm0.push_back(int(gemms.size()));
        sw->stage_tile0.push_back(int(tiles.size()));
        // ---- gather map (stage-0 boxes <- x) and scatter map (y <- term boxes, offsets relative to yoff[t])
        std::vector<int> xmap(size_t(sw->xn), -1);
        std::vector<long long> ymap(size_t(p->N) * p->B, 0);
        for (int e = 0; e < p->N; ++e) {
            const int blk = bs->eblk[e];
            const int* bi = &bs->binfo[size_t(blk) * (2 * D + 1)];
            Lev L(bs->blev.begin() + size_t(blk) * D, bs->blev.begin() + size_t(blk + 1) * D);
            std::vector<long long> cell(D);
            {
                long long ci = bs->ecell[e];
                for (int k = D - 1; k >= 0; --k) {
                    cell[k] = ci & ((1LL << bi[k]) - 1);
                    ci >>= bi[k];
                }
            }
            const Lev pp(L.begin(), L.end() - 1);
            const long long Kx = b.Kof(D - 1, pp);
            const Lev q(L.begin() + 1, L.end());
            const long long qf = b.fibres(q);
            for (int a = 0; a < p->B; ++a) {
                std::vector<int> av(D);
                int rest = a;
                for (int k = D - 1; k >= 0; --k) {
                    av[k] = rest % deg;
                    rest /= deg;
                }
                // x box: p fibre (dims 0..D-2) x r'_{D-1}
                long long pfi = 0;
                for (int k = 0; k < D - 1; ++k) pfi = pfi * (cells_of(L[k]) * deg) + cell[k] * deg + av[k];
                const long long rx = (start_of(L[D - 1]) + cell[D - 1]) * deg + av[D - 1];
                xmap[size_t(xbox.at(pp) + pfi * Kx + rx)] = e * p->B + a;
                // y box: r_0 x q fibre (dims 1..D-1)
                long long qfi = 0;
                for (int k = 1; k < D; ++k) qfi = qfi * (cells_of(L[k]) * deg) + cell[k] * deg + av[k];
                const long long r0 = (start_of(L[0]) + cell[0]) * deg + av[0];
                ymap[size_t(e) * p->B + a] = (obox[0][0].at({Lev(), q}) - sw->yoff[0]) + r0 * qf + qfi;
            }
        }
        // every address a stage can touch stays inside its array (checked here, on the host)
        for (const SwGemm& g : gemms) {
            const long long amax = g.a_off + (g.M - 1) + (long long)(g.K - 1) * g.lda;
            const long long bmax = g.b_off + (long long)(g.K - 1) * g.bk + (std::min(g.N, g.bE0) - 1) +
                                   (long long)((g.N - 1) / g.bE0) * g.bs1;
            const long long cmax = g.c_off + (long long)(g.M - 1) * g.cm + (std::min(g.N, g.cE0) - 1) +
                                   (long long)(std::min((g.N - 1) / g.cE0, g.cE1 - 1)) * g.cs1 +
                                   (long long)((g.N - 1) / (g.cE0 * g.cE1)) * g.cs2;
            if (g.bk * 4LL * g.K >= (1LL << 31) || 4LL * g.lda * g.K >= (1LL << 31))
                throw std::logic_error("sweep strides beyond 32-bit offsets");
            if (g.M < 1 || g.N < 1 || g.K < 1 || g.bE0 < 1 || g.cE0 < 1 || g.cE1 < 1 || g.a_off < 0 || g.b_off < 0 ||
                g.c_off < 0 || amax >= p->mat_elems || bmax >= sw->arena || cmax >= sw->arena)
                throw std::logic_error("sweep GEMM outside its arrays");
        }
        for (int v : xmap)
            if (v < -1 || v >= p->N * p->B) throw std::logic_error("sweep gather map out of range");
        for (long long v : ymap)
            for (int t = 0; t < T; ++t)
                if (v < 0 || sw->yoff[t] + v >= sw->arena) throw std::logic_error("sweep scatter map out of range");
        auto up = [](const auto* h, size_t n, auto*& d) {
            if (hipMalloc(&d, n * sizeof(*h)) != hipSuccess) throw std::bad_alloc();
            if (hipMemcpy(d, h, n * sizeof(*h), hipMemcpyHostToDevice) != hipSuccess)
                throw std::runtime_error("hipMemcpy failed");
        };
        up(xmap.data(), xmap.size(), sw->d_xmap);
        up(ymap.data(), ymap.size(), sw->d_ymap);
        up(gemms.data(), gemms.size(), sw->d_gemms);
        up(tiles.data(), tiles.size(), sw->d_tiles);
        if (hipMalloc(&sw->d_arena[0], size_t(sw->arena) * sizeof(double)) != hipSuccess) throw std::bad_alloc();
        if (bs->d_mat32 && hipMalloc(&sw->d_arena[1], size_t(sw->arena) * sizeof(float)) != hipSuccess)
            throw std::bad_alloc();
        p->sw = sw;
    } catch (const std::exception&) {
        p->sw = sw;
        kmx_sweep_free(p);
    }
}

int kmx_sweep_apply(kmx_plan* p, int slot, const double* x, double* y, hipStream_t s) {
    kmx_sweep* sw = p->sw;
    kmx_blockset* bs = p->bs;
    const bool f64 = slot == 0;
    ScatterArgs sa;
    sa.T = p->T;
    for (int t = 0; t < p->T; ++t) {
        sa.yoff[t] = sw->yoff[t];
        sa.scale[t] = p->scale[t];
    }
    const long long nval = (long long)p->N * p->B;
    auto run = [&](auto* arena, const auto* mat) {
        using T = std::remove_pointer_t<decltype(arena)>;
        hipLaunchKernelGGL(k_sweep_gather<T>, dim3(grid_of(sw->xn)), dim3(256), 0, s, x, sw->d_xmap, arena, sw->xn);
        for (int st = 0; st < sw->D; ++st) {
            const int t0 = sw->stage_tile0[st], ts = sw->stage_short[st], t1 = sw->stage_tile0[st + 1];
            if (ts > t0)  // long K: four chains per wave
                hipLaunchKernelGGL((k_sweep_stage<T, 4>), dim3((ts - t0 + 3) / 4), dim3(256), 0, s, mat, arena,
                                   sw->d_gemms, sw->d_tiles + t0, ts - t0);
            if (t1 > ts)
                hipLaunchKernelGGL((k_sweep_stage<T, 1>), dim3((t1 - ts + 3) / 4), dim3(256), 0, s, mat, arena,
                                   sw->d_gemms, sw->d_tiles + ts, t1 - ts);
        }
        hipLaunchKernelGGL(k_sweep_scatter<T>, dim3(grid_of(nval)), dim3(256), 0, s, arena, sw->d_ymap, y, nval, sa);
    };
    if (f64) {
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
    if (sw->d_xmap) hipFree(sw->d_xmap);
    if (sw->d_ymap) hipFree(sw->d_ymap);
    if (sw->d_gemms) hipFree(sw->d_gemms);
    if (sw->d_tiles) hipFree(sw->d_tiles);
    for (auto& a : sw->d_arena)
        if (a) hipFree(a);
    delete sw;
    p->sw = nullptr;
}

double kmx_sweep_flops(const kmx_plan* p) { return p->sw ? double(p->sw->flops) : 0.0; }
