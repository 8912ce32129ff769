// kmx_kron.hip -- the reference's standalone Kronecker building blocks on gfx950:
//   kmx_kronmult             Y = kron(A1..Ak) X   (kron_multd.m, Kronmult/kronmult{1..6}.m)
//   kmx_gemm_vbatched        variable-size batched GEMM (Kronmult/gemm_vbatched.m)
//   kmx_kronmult_batch_plan  the per-stage GEMM lists of Kronmult/kronmult{k}_batch.m
#include "kmx_internal.h"

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <numeric>

namespace kmx {

// One mode product of the method-1 chain: contract dim `dim` (0 = slowest) of a
// column-major tensor whose dims have current extents ext[] (ext of dim k is
// ncol_k before its contraction and nrow_k after).  One thread per output entry.
struct ModeArgs {
    const double* in;
    double* out;
    const double* A;  // nrow x ncol, column-major
    long long inner;  // product of extents of dims faster than `dim`
    long long outer;  // product of extents of dims slower than `dim` times nvec
    int nrow, ncol;
};

__global__ void __launch_bounds__(256) k_mode_product(ModeArgs m) {
    const long long total = m.inner * m.nrow * m.outer;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const long long s = e % m.inner;
        const long long r = e / m.inner;
        const int a = int(r % m.nrow);
        const long long o = r / m.nrow;
        const double* in = m.in + o * m.ncol * m.inner + s;
        double acc = 0.0;
        for (int b = 0; b < m.ncol; ++b) acc = __builtin_fma(m.A[(long long)b * m.nrow + a], in[b * m.inner], acc);
        m.out[e] = acc;
    }
}

struct GemmDescDev {
    const double* A;
    const double* B;
    double* C;
    int m, n, k, lda, ldb, ldc;
};

// The batch's GEMMs in 64 x 64 tiles of C: workgroup (g, t) takes tiles t, t +
// gridDim.y, ... of GEMM g (row-tile fastest).  k runs in chunks of 16: the chunk's
// 64 x 16 values of op(A) and 16 x 64 of op(B) are staged in LDS, each loaded
// along the matrix's contiguous index, and each thread accumulates a 4 x 4 set of
// C entries (rows tx + 16 r, columns ty + 16 c) over the chunk.  Every entry sums
// its k products in order with FMAs, then C = alpha sum + beta C (C not read when
// beta = 0), as the reference's gemm does entry by entry.
constexpr int GV_T = 64, GV_K = 16;
__global__ void __launch_bounds__(256) k_gemm_vbatched(const GemmDescDev* __restrict__ d, int count, int ta, int tb,
                                                       double alpha, double beta) {
    __shared__ double As[GV_K][GV_T + 1];  // As[l][i] = op(A)(i0 + i, l0 + l)
    __shared__ double Bs[GV_K][GV_T + 1];  // Bs[l][j] = op(B)(l0 + l, j0 + j)
    const int g = blockIdx.x;
    if (g >= count) return;
    const GemmDescDev q = d[g];
    const int mt = (q.m + GV_T - 1) / GV_T, nt = (q.n + GV_T - 1) / GV_T;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    for (int tile = blockIdx.y; tile < mt * nt; tile += gridDim.y) {
        const int i0 = (tile % mt) * GV_T, j0 = (tile / mt) * GV_T;
        double acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
        for (int l0 = 0; l0 < q.k; l0 += GV_K) {
            for (int t = threadIdx.x; t < GV_T * GV_K; t += 256) {
                const int ii = ta ? t / GV_K : t % GV_T, ll = ta ? t % GV_K : t / GV_T;
                const int i = i0 + ii, l = l0 + ll;
                As[ll][ii] = (i < q.m && l < q.k) ? (ta ? q.A[(long long)i * q.lda + l] : q.A[(long long)l * q.lda + i])
                                                  : 0.0;
            }
            for (int t = threadIdx.x; t < GV_T * GV_K; t += 256) {
                const int jj = tb ? t % GV_T : t / GV_K, ll = tb ? t / GV_T : t % GV_K;
                const int j = j0 + jj, l = l0 + ll;
                Bs[ll][jj] = (j < q.n && l < q.k) ? (tb ? q.B[(long long)l * q.ldb + j] : q.B[(long long)j * q.ldb + l])
                                                  : 0.0;
            }
            __syncthreads();
            const int lk = min(GV_K, q.k - l0);
            for (int l = 0; l < lk; ++l) {
                double a[4], b[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) a[r] = As[l][tx + 16 * r];
#pragma unroll
                for (int c = 0; c < 4; ++c) b[c] = Bs[l][ty + 16 * c];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[r][c] = __builtin_fma(a[r], b[c], acc[r][c]);
            }
            __syncthreads();  // the chunk's reads before the next chunk's stores
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + tx + 16 * r, j = j0 + ty + 16 * c;
                if (i < q.m && j < q.n) {
                    double& cv = q.C[(long long)j * q.ldc + i];
                    cv = (beta == 0.0 ? 0.0 : beta * cv) + alpha * acc[r][c];
                }
            }
    }
}

// ---- flops_kron{k} (Kronmult/flops_kron2.m:22-46, flops_kron3..6.m): cost of the
// better method for factors [l..k) and whether method 1 wins (ties -> method 1).
static double flops_best(const int* nr, const int* nc, int l, int k, int* method) {
    if (k - l == 1) {
        if (method) *method = 1;
        return 2.0 * nr[l] * nc[l];  // per vector (only used as a product term)
    }
    if (k - l == 2) {
        const double nrow1 = nr[l], ncol1 = nc[l], nrow2 = nr[l + 1], ncol2 = nc[l + 1];
        const double f1 = 2.0 * nrow2 * ncol2 * ncol1 + 2.0 * nrow2 * ncol1 * nrow1;
        const double f2 = 2.0 * ncol2 * ncol1 * nrow1 + 2.0 * nrow2 * ncol2 * nrow1;
        if (method) *method = (f1 <= f2) ? 1 : 2;
        return std::min(f1, f2);
    }
    double inner = flops_best(nr, nc, l + 1, k, nullptr);
    double prow = 1, pcol = 1;
    for (int i = l + 1; i < k; ++i) {
        prow *= nr[i];
        pcol *= nc[i];
    }
    const double f1 = inner * nc[l] + 2.0 * prow * nc[l] * nr[l];
    const double f2 = 2.0 * pcol * nc[l] * nr[l] + inner * nr[l];
    if (method) *method = (f1 <= f2) ? 1 : 2;
    return std::min(f1, f2);
}

struct Emitter {
    const int* nr;
    const int* nc;
    int K;  // number of factors
    std::vector<std::vector<kmx_gemm_desc>> stages;
    long long region = 0;  // elements per scratch region
    bool sizing = true;

    kmx_gemm_desc mk(int stage, char ta, char tb, int m, int n, int k, int lda, int ldb, int ldc, int bA,
                     long long oA, int bB, long long oB, int bC, long long oC) {
        kmx_gemm_desc g;
        g.stage = stage;
        g.transA = ta;
        g.transB = tb;
        g.m = m;
        g.n = n;
        g.k = k;
        g.lda = lda;
        g.ldb = ldb;
        g.ldc = ldc;
        g.alpha = 1.0;
        g.beta = 0.0;
        g.bufA = bA;
        g.offA = oA;
        g.bufB = bB;
        g.offB = oB;
        g.bufC = bC;
        g.offC = oC;
        return g;
    }
    void push(const kmx_gemm_desc& g) {
        if (!sizing) stages[g.stage - 1].push_back(g);
    }
    // factors [l..K), input (bin, oin), output (bout, oout), nvec vectors,
    // first stage index s0, scratch region depth.
    void emit(int l, int bin, long long oin, int bout, long long oout, long long nvec, int s0, int depth) {
        const int nrow1 = nr[l], ncol1 = nc[l];
        const long long tmpoff = (long long)depth * region;
        if (K - l == 1) {  // kronmult1_batch.m:45-61
            push(mk(s0, 'N', 'N', nrow1, int(nvec), ncol1, nrow1, ncol1, nrow1, KMX_BUF_FACTOR + l, 0, bin, oin,
                    bout, oout));
            return;
        }
        long long prow = 1, pcol = 1;
        for (int i = l + 1; i < K; ++i) {
            prow *= nr[i];
            pcol *= nc[i];
        }
        int method;
        flops_best(nr, nc, l, K, &method);
        const int nstage = K - l;
        if (method == 1) {
            // Ytmp = kronmult_{k-1}(A2..Ak, X): prow x (ncol1*nvec) in scratch
            const long long need = prow * ncol1 * nvec;
            if (sizing) region = std::max(region, need);
            emit(l + 1, bin, oin, KMX_BUF_W0, tmpoff, nvec * ncol1, s0, depth + 1);
            for (long long v = 0; v < nvec; ++v)  // kronmult2_batch.m:71-111, kronmult3_batch (method 1)
                push(mk(s0 + nstage - 1, 'N', 'T', int(prow), nrow1, ncol1, int(prow), nrow1, int(prow), KMX_BUF_W0,
                        tmpoff + v * prow * ncol1, KMX_BUF_FACTOR + l, 0, bout, oout + v * prow * nrow1));
        } else {
            // Ytmp(:, v) = X_v * A1^T : pcol x (nrow1*nvec); then kronmult_{k-1}(A2..Ak, Ytmp)
            const long long need = pcol * nrow1 * nvec;
            if (sizing) region = std::max(region, need);
            for (long long v = 0; v < nvec; ++v)
                push(mk(s0, 'N', 'T', int(pcol), nrow1, ncol1, int(pcol), nrow1, int(pcol), bin,
                        oin + v * pcol * ncol1, KMX_BUF_FACTOR + l, 0, KMX_BUF_W0, tmpoff + v * pcol * nrow1));
            emit(l + 1, KMX_BUF_W0, tmpoff, bout, oout, nvec * nrow1, s0 + 1, depth + 1);
        }
    }
};

}  // namespace kmx

using namespace kmx;

extern "C" int kmx_kronmult_batch_plan(int k, const int* nrow, const int* ncol, int nvec, kmx_gemm_desc* out,
                                       int capacity, int* count, int64_t* work_elems) {
    if (k < 1 || k > 6 || !nrow || !ncol || nvec < 1 || !count) {
        set_error("kmx_kronmult_batch_plan: bad arguments");
        return KMX_E_ARG;
    }
    for (int i = 0; i < k; ++i)
        if (nrow[i] < 1 || ncol[i] < 1) {
            set_error("kmx_kronmult_batch_plan: factor sizes must be positive");
            return KMX_E_SHAPE;
        }
    Emitter em{nrow, ncol, k, {}, 0, true};
    em.stages.assign(k, {});
    em.emit(0, KMX_BUF_X, 0, KMX_BUF_Y, 0, nvec, 1, 0);  // sizing pass
    em.sizing = false;
    em.emit(0, KMX_BUF_X, 0, KMX_BUF_Y, 0, nvec, 1, 0);
    int total = 0;
    for (auto& s : em.stages) total += int(s.size());
    *count = total;
    if (work_elems) *work_elems = em.region * k;
    if (out) {
        if (capacity < total) {
            set_error("kmx_kronmult_batch_plan: capacity too small");
            return KMX_E_ARG;
        }
        int o = 0;
        for (auto& s : em.stages)
            for (auto& g : s) out[o++] = g;
    }
    return KMX_OK;
}

namespace {

// Values of the largest intermediate of the method-1 chain (ext[i] = ncol_i until
// dim i is contracted, then nrow_i; nvec vectors).
long long chain_workspace(int k, const int* nrow, const int* ncol, int nvec) {
    std::vector<long long> e2(ncol, ncol + k);
    long long maxsz = 0;
    for (int d = k - 1; d >= 0; --d) {
        e2[d] = nrow[d];
        long long sz = nvec;
        for (int i = 0; i < k; ++i) sz *= e2[i];
        maxsz = std::max(maxsz, sz);
    }
    return maxsz;
}

// The mode products of kron(A1..Ak) X, last factor first (method 1), through the
// ping-pong buffers buf[0] / buf[1] (buf[1] used only for k > 2).  For k > 6 this
// is also the order of kron_multd.m's recursion (kron_multd.m:64-86: Z = A_k X,
// then kron(A_1..A_{k-1}) on Z's transpose).
void launch_chain(int k, const int* nrow, const int* ncol, const double* const* A, const double* X, int nvec,
                  double* Y, double* const* buf, hipStream_t s) {
    std::vector<long long> ext(ncol, ncol + k);
    const double* cur = X;
    int ping = 0;
    for (int d = k - 1; d >= 0; --d) {
        ModeArgs m;
        m.in = cur;
        m.out = (d == 0) ? Y : buf[ping];
        m.A = A[d];
        m.nrow = nrow[d];
        m.ncol = ncol[d];
        m.inner = 1;
        for (int i = d + 1; i < k; ++i) m.inner *= ext[i];
        m.outer = nvec;
        for (int i = 0; i < d; ++i) m.outer *= ext[i];
        const long long total = m.inner * m.nrow * m.outer;
        const int grid = int(std::min<long long>((total + 255) / 256, 65536));
        if (grid > 0) hipLaunchKernelGGL(k_mode_product, dim3(grid), dim3(256), 0, s, m);
        ext[d] = nrow[d];
        cur = m.out;
        ping ^= 1;
    }
}

bool kron_args_ok(int k, const int* nrow, const int* ncol, const double* const* A, const double* X, int nvec,
                  const double* Y) {
    if (k < 1 || k > KMX_KRON_MAX_FACTORS || !nrow || !ncol || !A || !X || !Y || nvec < 0) return false;
    for (int i = 0; i < k; ++i)
        if (nrow[i] < 1 || ncol[i] < 1 || !A[i]) return false;
    return true;
}

}  // namespace

extern "C" int kmx_kronmult(int k, const int* nrow, const int* ncol, const double* const* A, const double* X,
                            int nvec, double* Y, void* stream) {
    if (!kron_args_ok(k, nrow, ncol, A, X, nvec, Y)) {
        set_error("kmx_kronmult: bad arguments");
        return KMX_E_ARG;
    }
    if (nvec == 0) return KMX_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const long long maxsz = chain_workspace(k, nrow, ncol, nvec);
    double* buf[2] = {nullptr, nullptr};
    if (k > 1) {
        if (hipMalloc(&buf[0], sizeof(double) * maxsz) != hipSuccess ||
            (k > 2 && hipMalloc(&buf[1], sizeof(double) * maxsz) != hipSuccess)) {
            hipFree(buf[0]);
            set_error("kmx_kronmult: workspace allocation failed");
            return KMX_E_NOMEM;
        }
    }
    launch_chain(k, nrow, ncol, A, X, nvec, Y, buf, s);
    hipError_t e = hipGetLastError();
    if (k > 1) {
        hipStreamSynchronize(s);
        hipFree(buf[0]);
        if (buf[1]) hipFree(buf[1]);
    }
    if (e != hipSuccess) {
        set_error(std::string("kmx_kronmult: ") + hipGetErrorString(e));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

// ---- kmx_kronmult without allocation: the chain's two intermediates live in one
// file-scope device array (each GPU holds its own copy for the life of the
// process).  Its device address is looked up once per device, with that device
// current, and only that address reaches the kernels.
static __device__ __align__(256) double scratch[KMX_KRON_SCRATCH_ELEMS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

double* scratch_for_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        g_scratch[dev] = static_cast<double*>(ptr);
    }
    return g_scratch[dev];
}
}  // namespace

extern "C" int kmx_kronmult_noalloc(int k, const int* nrow, const int* ncol, const double* const* A,
                                    const double* X, int nvec, double* Y, void* stream) {
    if (!kron_args_ok(k, nrow, ncol, A, X, nvec, Y)) {
        set_error("kmx_kronmult_noalloc: bad arguments");
        return KMX_E_ARG;
    }
    if (nvec == 0) return KMX_OK;
    const long long maxsz = chain_workspace(k, nrow, ncol, nvec);
    const long long half = KMX_KRON_SCRATCH_ELEMS / 2;
    if (k > 1 && maxsz > half) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "kmx_kronmult_noalloc: an intermediate of %lld values exceeds the scratch limit of %lld "
                      "(KMX_KRON_SCRATCH_ELEMS / 2); use kmx_kronmult",
                      maxsz, half);
        set_error(msg);
        return KMX_E_SHAPE;
    }
    double* base = scratch_for_current_device();
    if (!base) {
        set_error("kmx_kronmult_noalloc: hipGetSymbolAddress of the scratch array failed");
        return KMX_E_HIP;
    }
    double* buf[2] = {base, base + half};
    launch_chain(k, nrow, ncol, A, X, nvec, Y, buf, reinterpret_cast<hipStream_t>(stream));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("kmx_kronmult_noalloc: ") + hipGetErrorString(e));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

extern "C" int kmx_gemm_vbatched(char transA, char transB, const int* m, const int* n, const int* k, double alpha,
                                 const double* const* A, const int* lda, const double* const* B, const int* ldb,
                                 double beta, double* const* C, const int* ldc, int count, void* stream) {
    if (count < 0 || (count > 0 && (!m || !n || !k || !A || !lda || !B || !ldb || !C || !ldc))) {
        set_error("kmx_gemm_vbatched: bad arguments");
        return KMX_E_ARG;
    }
    auto tr = [](char c) { return c == 'T' || c == 't' || c == 'C' || c == 'c'; };
    auto valid = [](char c) { return c == 'N' || c == 'n' || c == 'T' || c == 't' || c == 'C' || c == 'c'; };
    if (!valid(transA) || !valid(transB)) {
        set_error("kmx_gemm_vbatched: trans must be N or T");
        return KMX_E_ARG;
    }
    if (count == 0) return KMX_OK;
    std::vector<GemmDescDev> h(count);
    for (int i = 0; i < count; ++i) {
        if (m[i] < 0 || n[i] < 0 || k[i] < 0) {
            set_error("kmx_gemm_vbatched: negative size");
            return KMX_E_SHAPE;
        }
        const int rowsA = tr(transA) ? k[i] : m[i];
        const int rowsB = tr(transB) ? n[i] : k[i];
        if (lda[i] < std::max(1, rowsA) || ldb[i] < std::max(1, rowsB) || ldc[i] < std::max(1, m[i])) {
            set_error("kmx_gemm_vbatched: leading dimension too small");  // gemm_vbatched.m:17-19 style
            return KMX_E_SHAPE;
        }
        h[i] = GemmDescDev{A[i], B[i], C[i], m[i], n[i], k[i], lda[i], ldb[i], ldc[i]};
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GemmDescDev* d = nullptr;
    if (hipMalloc(&d, sizeof(GemmDescDev) * count) != hipSuccess) {
        set_error("kmx_gemm_vbatched: descriptor allocation failed");
        return KMX_E_NOMEM;
    }
    hipMemcpyAsync(d, h.data(), sizeof(GemmDescDev) * count, hipMemcpyHostToDevice, s);
    long long tiles = 1;  // C tiles of the batch's largest GEMM (grid y; more loop)
    for (int i = 0; i < count; ++i)
        tiles = std::max(tiles, (long long)((m[i] + 63) / 64) * ((n[i] + 63) / 64));
    hipLaunchKernelGGL(k_gemm_vbatched, dim3(unsigned(count), unsigned(std::min<long long>(tiles, 65535))), dim3(256),
                       0, s, d, count, int(tr(transA)), int(tr(transB)), alpha, beta);
    hipError_t e = hipGetLastError();
    hipStreamSynchronize(s);
    hipFree(d);
    if (e != hipSuccess) {
        set_error(std::string("kmx_gemm_vbatched: ") + hipGetErrorString(e));
        return KMX_E_HIP;
    }
    return KMX_OK;
}
