// kmx_realspace.hip -- wavelet space -> real space on gfx950
// (wavelet_to_realspace.m:1-79).
//
// The reference sums, over elements e, kron(Meval_1(:, J_1(e)), ..., Meval_d(:, J_d(e)))
// times the element's deg^d coefficients, where Meval_k(:, J) is the deg-column
// block of dimension k's evaluation matrix at the element's 1D index.  Here the
// elements are grouped by level multi-index: the elements of one group are the
// cells of a dense tensor, and their summed contribution is one rectangular
// Kronecker product of the level's column ranges applied to that tensor (absent
// cells are zeros) -- d mode products per group instead of one kron_multd per
// element.  Same terms, summed in another order.
#include "kmx_internal.h"

#include <algorithm>
#include <map>
#include <stdexcept>
#include <vector>

namespace kmx {
namespace {

struct GatherArgs {
    const double* fw;
    double* X;
    const int* elems;  // cell -> element (row-major over cells, dim 1 slowest), -1 absent
    int cells[KMX_MAX_DIM];
    int D, deg, B;
    long long n;  // entries of X
};

// X[(c_1 deg + a_1), ..., (c_d deg + a_d)] (row-major, dim 1 slowest) from the
// element-major coefficient vector.
__global__ void __launch_bounds__(256) k_gather(GatherArgs g) {
    for (long long v = blockIdx.x * 256ll + threadIdx.x; v < g.n; v += (long long)gridDim.x * 256) {
        long long rest = v;
        int ci = 0, ai = 0, cstride = 1, astride = 1;
        for (int k = g.D - 1; k >= 0; --k) {
            const int ext = g.cells[k] * g.deg;
            const int m = int(rest % ext);
            rest /= ext;
            ci += (m / g.deg) * cstride;
            ai += (m % g.deg) * astride;
            cstride *= g.cells[k];
            astride *= g.deg;
        }
        const int e = g.elems[ci];
        g.X[v] = e >= 0 ? g.fw[(long long)e * g.B + ai] : 0.0;
    }
}

struct ModeArgs {
    const double* in;
    double* out;
    const double* A;  // M x K column-major, leading dimension lda
    long long pre, post;
    int M, K, lda;
    int accumulate;
};

// out[p, i, q] (= or +=) sum_j A(i, j) in[p, j, q]; one thread per output.
__global__ void __launch_bounds__(256) k_mode(ModeArgs m) {
    const long long total = m.pre * m.M * m.post;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const long long q = e % m.post;
        const long long r = e / m.post;
        const int i = int(r % m.M);
        const long long p = r / m.M;
        const double* in = m.in + p * m.K * m.post + q;
        double acc = 0.0;
        for (int j = 0; j < m.K; ++j) acc = __builtin_fma(m.A[i + (long long)j * m.lda], in[j * m.post], acc);
        if (m.accumulate)
            m.out[e] += acc;
        else
            m.out[e] = acc;
    }
}

__global__ void __launch_bounds__(256) k_zero(double* y, long long n) {
    for (long long v = blockIdx.x * 256ll + threadIdx.x; v < n; v += (long long)gridDim.x * 256) y[v] = 0.0;
}

unsigned grid_for(long long n) { return unsigned(std::min<long long>((n + 255) / 256, 65536)); }

int level_of(int I0) {
    int l = 0;
    while ((1 << l) <= I0) ++l;
    return l;
}

}  // namespace
}  // namespace kmx

using namespace kmx;

struct kmx_realspace {
    int D = 0, deg = 0, N = 0, B = 0;
    int npts[KMX_MAX_DIM] = {0}, ncols[KMX_MAX_DIM] = {0};
    long long nout = 0;
    double* d_meval[KMX_MAX_DIM] = {nullptr};
    struct Block {
        int cells[KMX_MAX_DIM];
        int col0[KMX_MAX_DIM];  // first column of the level range in Meval_d
        int eoff;
    };
    std::vector<Block> blocks;
    int* d_elems = nullptr;
    double* d_buf[3] = {nullptr, nullptr, nullptr};
};

static void free_rs(kmx_realspace* r) {
    for (auto& m : r->d_meval)
        if (m) hipFree(m);
    if (r->d_elems) hipFree(r->d_elems);
    for (auto& b : r->d_buf)
        if (b) hipFree(b);
}

extern "C" int kmx_realspace_create(int ndim, int deg, int num_elem, const int* elem_idx1d, int index_base,
                                    const int* npts, const int* ncols, const double* const* meval,
                                    kmx_realspace** out) {
    if (!out || ndim < 1 || ndim > KMX_MAX_DIM || deg < 1 || deg > KMX_MAX_DEG || num_elem < 0 || !npts || !ncols ||
        !meval || (num_elem > 0 && !elem_idx1d) || (index_base != 0 && index_base != 1)) {
        set_error("kmx_realspace_create: bad arguments");
        return KMX_E_ARG;
    }
    *out = nullptr;
    auto* r = new kmx_realspace();
    try {
        r->D = ndim;
        r->deg = deg;
        r->N = num_elem;
        r->B = 1;
        r->nout = 1;
        for (int k = 0; k < ndim; ++k) {
            if (npts[k] < 1 || ncols[k] < deg || ncols[k] % deg || !meval[k])
                throw std::invalid_argument("Meval_d must be npts_d x (deg * cells), npts_d >= 1");
            r->B *= deg;
            r->npts[k] = npts[k];
            r->ncols[k] = ncols[k];
            r->nout *= npts[k];
            const size_t sz = size_t(npts[k]) * ncols[k];
            if (hipMalloc(&r->d_meval[k], sz * sizeof(double)) != hipSuccess) throw std::bad_alloc();
            if (hipMemcpy(r->d_meval[k], meval[k], sz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                throw std::runtime_error("hipMemcpy failed");
        }
        // elements by level multi-index
        std::map<std::vector<int>, int> bid;
        std::vector<std::vector<int>> members;
        std::vector<std::vector<int>> ecell(num_elem);
        for (int e = 0; e < num_elem; ++e) {
            std::vector<int> lev(ndim);
            ecell[e].resize(ndim);
            for (int k = 0; k < ndim; ++k) {
                const int I0 = elem_idx1d[size_t(e) * ndim + k] - index_base;
                if (I0 < 0 || (I0 + 1) * deg > ncols[k]) throw std::invalid_argument("elem_idx1d out of range");
                lev[k] = level_of(I0);
                ecell[e][k] = lev[k] == 0 ? 0 : I0 - (1 << (lev[k] - 1));
            }
            auto it = bid.find(lev);
            if (it == bid.end()) {
                it = bid.emplace(lev, int(members.size())).first;
                members.emplace_back();
            }
            members[it->second].push_back(e);
        }
        std::vector<int> elems;
        long long maxX = 1, maxI = 1;
        r->blocks.resize(members.size());
        for (const auto& kv : bid) {
            auto& b = r->blocks[kv.second];
            long long ncell = 1;
            for (int k = 0; k < ndim; ++k) {
                const int l = kv.first[k];
                b.col0[k] = l == 0 ? 0 : (1 << (l - 1));
                b.cells[k] = std::min(l == 0 ? 1 : (1 << (l - 1)), ncols[k] / deg - b.col0[k]);
                ncell *= b.cells[k];
            }
            if (ncell * r->B > (1LL << 30)) throw std::invalid_argument("level block too large");
            b.eoff = int(elems.size());
            elems.resize(elems.size() + size_t(ncell), -1);
            for (int e : members[kv.second]) {
                long long ci = 0;
                for (int k = 0; k < ndim; ++k) ci = ci * b.cells[k] + ecell[e][k];
                int& slot = elems[size_t(b.eoff) + size_t(ci)];
                if (slot >= 0) throw std::invalid_argument("two elements share one 1D index tuple");
                slot = e;
            }
            maxX = std::max(maxX, ncell * r->B);
            // intermediates of the chain: dims d..k+1 already at npts
            long long ext[KMX_MAX_DIM];
            for (int k = 0; k < ndim; ++k) ext[k] = (long long)b.cells[k] * deg;
            for (int k = ndim - 1; k >= 1; --k) {
                ext[k] = npts[k];
                long long sz = 1;
                for (int m = 0; m < ndim; ++m) sz *= ext[m];
                maxI = std::max(maxI, sz);
            }
        }
        if (!elems.empty()) {
            if (hipMalloc(&r->d_elems, elems.size() * sizeof(int)) != hipSuccess) throw std::bad_alloc();
            if (hipMemcpy(r->d_elems, elems.data(), elems.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
                throw std::runtime_error("hipMemcpy failed");
        }
        if (hipMalloc(&r->d_buf[0], size_t(maxX) * sizeof(double)) != hipSuccess ||
            hipMalloc(&r->d_buf[1], size_t(maxI) * sizeof(double)) != hipSuccess ||
            hipMalloc(&r->d_buf[2], size_t(maxI) * sizeof(double)) != hipSuccess)
            throw std::bad_alloc();
        *out = r;
        return KMX_OK;
    } catch (const std::bad_alloc&) {
        set_error("kmx_realspace_create: device allocation failed");
        free_rs(r);
        delete r;
        return KMX_E_NOMEM;
    } catch (const std::exception& e) {
        set_error(std::string("kmx_realspace_create: ") + e.what());
        free_rs(r);
        delete r;
        return KMX_E_ARG;
    }
}

extern "C" int kmx_realspace_apply(kmx_realspace* r, const double* f_w, double* f_r, void* stream) {
    if (!r || !f_r || (r->N > 0 && !f_w)) {
        set_error("kmx_realspace_apply: bad arguments");
        return KMX_E_ARG;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_zero, dim3(grid_for(r->nout)), dim3(256), 0, s, f_r, r->nout);
    const int D = r->D, deg = r->deg;
    for (const auto& b : r->blocks) {
        GatherArgs g;
        g.fw = f_w;
        g.X = r->d_buf[0];
        g.elems = r->d_elems + b.eoff;
        g.D = D;
        g.deg = deg;
        g.B = r->B;
        g.n = r->B;
        for (int k = 0; k < D; ++k) {
            g.cells[k] = b.cells[k];
            g.n *= b.cells[k];
        }
        hipLaunchKernelGGL(k_gather, dim3(grid_for(g.n)), dim3(256), 0, s, g);
        long long ext[KMX_MAX_DIM];
        for (int k = 0; k < D; ++k) ext[k] = (long long)b.cells[k] * deg;
        const double* cur = r->d_buf[0];
        int ping = 1;
        for (int k = D - 1; k >= 0; --k) {  // last dimension first (kron_multd's method-1 order)
            ModeArgs m;
            m.in = cur;
            m.out = k == 0 ? f_r : r->d_buf[ping];
            m.A = r->d_meval[k] + (long long)b.col0[k] * deg * r->npts[k];
            m.lda = r->npts[k];
            m.M = r->npts[k];
            m.K = int(ext[k]);
            m.pre = 1;
            m.post = 1;
            for (int q = 0; q < k; ++q) m.pre *= ext[q];
            for (int q = k + 1; q < D; ++q) m.post *= ext[q];
            m.accumulate = k == 0;
            const long long total = m.pre * m.M * m.post;
            hipLaunchKernelGGL(k_mode, dim3(grid_for(total)), dim3(256), 0, s, m);
            ext[k] = r->npts[k];
            cur = m.out;
            ping = 3 - ping;
        }
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_error(std::string("kmx_realspace_apply: ") + hipGetErrorString(err));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

extern "C" int kmx_realspace_destroy(kmx_realspace* r) {
    if (!r) return KMX_OK;
    free_rs(r);
    delete r;
    return KMX_OK;
}
