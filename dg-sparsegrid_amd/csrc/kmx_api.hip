// kmx_api.hip -- C ABI of libkmx: resident operator plans and their entry points.
// Host-side work here is set-up only (layout conversion, pair-list ordering);
// every apply is a gfx950 kernel (kmx_apply.hip).
#include "kmx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace kmx {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace kmx

using namespace kmx;

namespace {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NoMem : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T* upload(const T* h, size_t n) {
    if (n == 0) return nullptr;
    T* d = nullptr;
    if (hipMalloc(&d, sizeof(T) * n) != hipSuccess) throw NoMem("device allocation failed");
    hip_check(hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}

template <typename T>
void dfree(T*& p) {
    if (p) hipFree(p);
    p = nullptr;
}

void free_plan(kmx_plan* p) {
    dfree(p->d_idx1d);
    dfree(p->d_row_ptr);
    dfree(p->d_cols);
    dfree(p->d_row_order);
    dfree(p->d_blk64);
    dfree(p->d_blk32);
    dfree(p->d_col_ptr);
    dfree(p->d_rows);
    dfree(p->d_col_order);
    dfree(p->d_blkT64);
    dfree(p->d_blkT32);
    dfree(p->d_x_host_stage);
    dfree(p->d_y_host_stage);
    kmx_sweep_free(p);
    kmx_blocks_free(p);
}

// Block-row-major copy of one column-major n x n matrix (see KArgs in
// kmx_internal.h); `transpose` stores the blocks of M^T instead.
void block_matrix(const double* M, int n, int deg, bool transpose, double* out) {
    const int nb = n / deg;
    for (int I = 0; I < nb; ++I)
        for (int J = 0; J < nb; ++J)
            for (int a = 0; a < deg; ++a)
                for (int b = 0; b < deg; ++b) {
                    // forward: (row I*deg+a, col J*deg+b) of M
                    // transpose: (row I*deg+a, col J*deg+b) of M^T = M(J*deg+b, I*deg+a)
                    const double v = transpose ? M[size_t(I * deg + a) * n + (J * deg + b)]
                                               : M[size_t(J * deg + b) * n + (I * deg + a)];
                    out[((size_t(I) * nb + J) * deg + a) * deg + b] = v;
                }
}

// rows ordered by descending work (pair count), ties by row id
std::vector<int> lpt_order(const std::vector<long long>& rp, int N) {
    std::vector<int> ord(N);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
    return ord;
}

int code_of(const std::exception& e) {
    if (dynamic_cast<const HipError*>(&e)) return KMX_E_HIP;
    if (dynamic_cast<const NoMem*>(&e)) return KMX_E_NOMEM;
    return KMX_E_ARG;
}

}  // namespace

extern "C" {

const char* kmx_last_error(void) { return kmx::g_err.c_str(); }
const char* kmx_version(void) { return "kmx 0.1 gfx950"; }

int kmx_sync(void* stream) {
    hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        set_error(std::string("kmx_sync: ") + hipGetErrorString(e));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

int kmx_plan_create(const kmx_plan_desc* desc, kmx_plan** out) {
    if (!desc || !out) {
        set_error("kmx_plan_create: null argument");
        return KMX_E_ARG;
    }
    *out = nullptr;
    kmx_plan* p = new kmx_plan();
    try {
        const int D = desc->ndim, deg = desc->deg, N = desc->num_elem, T = desc->num_terms;
        if (D < 1 || D > KMX_MAX_DIM) throw ArgError("ndim must be 1..6");
        if (deg < 1 || deg > KMX_MAX_DEG) throw ArgError("deg must be 1..8");
        if (T < 1 || T > KMX_MAX_TERMS) throw ArgError("num_terms must be 1..16");
        if (N < 0) throw ArgError("num_elem < 0");
        if (!desc->mat_rows || !desc->mats || (N > 0 && !desc->elem_idx1d)) throw ArgError("null array");
        if (desc->index_base != 0 && desc->index_base != 1) throw ArgError("index_base must be 0 or 1");
        if (desc->row_ptr && !desc->cols && N > 0) throw ArgError("row_ptr given without cols");
        p->D = D;
        p->deg = deg;
        p->N = N;
        p->T = T;
        p->flags = desc->flags;
        int B = 1;
        for (int k = 0; k < D; ++k) B *= deg;
        p->B = B;
        for (int t = 0; t < KMX_MAX_TERMS; ++t) p->scale[t] = 1.0;
        hip_check(hipGetDevice(&p->device), "hipGetDevice");

        long long moff = 0;
        for (int k = 0; k < D; ++k) {
            const int n = desc->mat_rows[k];
            if (n < deg || n % deg) throw std::runtime_error("mat_rows[d] must be a positive multiple of deg");
            p->n[k] = n;
            p->nb[k] = n / deg;
        }
        for (int t = 0; t < T; ++t)
            for (int k = 0; k < D; ++k) {
                if (!desc->mats[t * D + k]) throw ArgError("null matrix pointer");
                p->off[t * D + k] = moff;
                moff += (long long)p->n[k] * p->n[k];
            }
        p->mat_elems = moff;

        // element table, 0-based, bounds-checked against the matrices
        std::vector<int> idx(size_t(N) * D);
        for (long long e = 0; e < (long long)N * D; ++e) {
            const int v = desc->elem_idx1d[e] - desc->index_base;
            const int k = int(e % D);
            if (v < 0 || v >= p->nb[k]) throw ArgError("elem_idx1d out of range for the 1D matrix size");
            idx[e] = v;
        }

        // pair list
        std::vector<long long> rp(size_t(N) + 1);
        std::vector<int> cols;
        if (desc->row_ptr) {
            if (desc->row_ptr[0] != 0) throw ArgError("row_ptr[0] must be 0");
            for (int i = 0; i < N; ++i) {
                if (desc->row_ptr[i + 1] < desc->row_ptr[i]) throw ArgError("row_ptr must be non-decreasing");
                rp[i] = desc->row_ptr[i];
            }
            rp[N] = desc->row_ptr[N];
            cols.resize(size_t(rp[N]));
            for (long long q = 0; q < rp[N]; ++q) {
                const int j = desc->cols[q] - desc->index_base;
                if (j < 0 || j >= N) throw ArgError("cols entry out of range");
                cols[q] = j;
            }
        } else {
            p->all_pairs = true;
            for (int i = 0; i <= N; ++i) rp[i] = (long long)i * N;
        }
        p->P = rp[N];

        // blocked coefficient matrices
        p->h_blk64.assign(size_t(moff), 0.0);
        for (int t = 0; t < T; ++t)
            for (int k = 0; k < D; ++k)
                block_matrix(desc->mats[t * D + k], p->n[k], deg, false, p->h_blk64.data() + p->off[t * D + k]);
        p->d_blk64 = upload(p->h_blk64.data(), size_t(moff));
        const bool want32 = (desc->flags & KMX_PLAN_FP32) != 0;
        if (want32) {
            std::vector<float> f(p->h_blk64.begin(), p->h_blk64.end());
            p->d_blk32 = upload(f.data(), f.size());
        }
        p->d_idx1d = upload(idx.data(), idx.size());
        p->h_idx1d = idx;
        // factors declared the identity: checked against the matrices given
        if (desc->identity) {
            p->declared_identity.assign(desc->identity, desc->identity + size_t(T) * D);
            for (int t = 0; t < T; ++t)
                for (int k = 0; k < D; ++k) {
                    if (!p->declared_identity[size_t(t) * D + k]) continue;
                    const double* M = desc->mats[t * D + k];
                    const long long n = p->n[k];
                    double dev = 0.0;
                    for (long long c = 0; c < n; ++c)
                        for (long long r = 0; r < n; ++r) dev = std::max(dev, std::fabs(M[c * n + r] - (r == c ? 1.0 : 0.0)));
                    if (dev > 1e-8) {
                        char msg[160];
                        std::snprintf(msg, sizeof msg, "term %d dim %d is declared the identity but max |A - I| = %.3g",
                                      t, k, dev);
                        throw ArgError(msg);
                    }
                }
        }
        std::vector<int> ord = lpt_order(rp, N);
        p->d_row_order = upload(ord.data(), ord.size());
        if (!p->all_pairs) {
            p->d_row_ptr = upload(rp.data(), rp.size());
            p->d_cols = upload(cols.data(), cols.size());
        }

        // adjoint: transposed pair list (CSC) and transposed blocks
        if (desc->flags & KMX_PLAN_ADJOINT) {
            std::vector<long long> cp(size_t(N) + 1, 0);
            std::vector<int> rows(size_t(p->P));
            if (p->all_pairs) {
                for (int j = 0; j <= N; ++j) cp[j] = (long long)j * N;
                for (long long q = 0; q < p->P; ++q) rows[q] = int(q % N);
            } else {
                for (long long q = 0; q < p->P; ++q) cp[size_t(cols[q]) + 1]++;
                for (int j = 0; j < N; ++j) cp[j + 1] += cp[j];
                std::vector<long long> fill(cp.begin(), cp.end() - 1);
                for (int i = 0; i < N; ++i)
                    for (long long q = rp[i]; q < rp[i + 1]; ++q) rows[fill[cols[q]]++] = i;
            }
            p->h_blkT64.assign(size_t(moff), 0.0);
            for (int t = 0; t < T; ++t)
                for (int k = 0; k < D; ++k)
                    block_matrix(desc->mats[t * D + k], p->n[k], deg, true, p->h_blkT64.data() + p->off[t * D + k]);
            p->d_blkT64 = upload(p->h_blkT64.data(), size_t(moff));
            if (want32) {
                std::vector<float> f(p->h_blkT64.begin(), p->h_blkT64.end());
                p->d_blkT32 = upload(f.data(), f.size());
            }
            std::vector<int> ordT = lpt_order(cp, N);
            p->d_col_order = upload(ordT.data(), ordT.size());
            if (!p->all_pairs) {
                p->d_col_ptr = upload(cp.data(), cp.size());
                p->d_rows = upload(rows.data(), rows.size());
            }
        }
        if (desc->flags & KMX_PLAN_LEVEL_BLOCKS) {
            std::vector<const double*> mats(desc->mats, desc->mats + size_t(T) * D);
            kmx_blocks_build(p, idx, rp, cols, mats.data());
        }
        kmx_plan_apply_cached_variants(p);
        *out = p;
        return KMX_OK;
    } catch (const std::exception& e) {
        set_error(std::string("kmx_plan_create: ") + e.what());
        int c = code_of(e);
        if (std::string(e.what()).find("must be a positive multiple") != std::string::npos) c = KMX_E_SHAPE;
        free_plan(p);
        delete p;
        return c;
    }
}

int kmx_plan_update_term(kmx_plan* p, int term, int dim, const double* mat, void* stream) {
    if (!p || !mat || term < 0 || term >= p->T || dim < 0 || dim >= p->D) {
        set_error("kmx_plan_update_term: bad arguments");
        return KMX_E_ARG;
    }
    try {
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        const int n = p->n[dim];
        const long long off = p->off[term * p->D + dim];
        const size_t cnt = size_t(n) * n;
        block_matrix(mat, n, p->deg, false, p->h_blk64.data() + off);
        hip_check(hipMemcpyAsync(p->d_blk64 + off, p->h_blk64.data() + off, cnt * sizeof(double),
                                 hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        std::vector<float> f;
        if (p->d_blk32) {
            f.assign(p->h_blk64.begin() + off, p->h_blk64.begin() + off + cnt);
            hip_check(hipMemcpyAsync(p->d_blk32 + off, f.data(), cnt * sizeof(float), hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync");
        }
        std::vector<float> fT;
        if (p->d_blkT64) {
            block_matrix(mat, n, p->deg, true, p->h_blkT64.data() + off);
            hip_check(hipMemcpyAsync(p->d_blkT64 + off, p->h_blkT64.data() + off, cnt * sizeof(double),
                                     hipMemcpyHostToDevice, s), "hipMemcpyAsync");
            if (p->d_blkT32) {
                fT.assign(p->h_blkT64.begin() + off, p->h_blkT64.begin() + off + cnt);
                hip_check(hipMemcpyAsync(p->d_blkT32 + off, fT.data(), cnt * sizeof(float), hipMemcpyHostToDevice, s),
                          "hipMemcpyAsync");
            }
        }
        // the float staging vectors die here: wait for the copies that read them
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        kmx_blocks_update_term(p, term, dim, mat, s);
        kmx_sweep_term_updated(p, term, dim);
        ++p->version;
        return KMX_OK;
    } catch (const std::exception& e) {
        set_error(std::string("kmx_plan_update_term: ") + e.what());
        return code_of(e);
    }
}

int kmx_plan_set_term_scale(kmx_plan* p, int term, double alpha) {
    if (!p || term < 0 || term >= p->T) {
        set_error("kmx_plan_set_term_scale: bad arguments");
        return KMX_E_ARG;
    }
    p->scale[term] = alpha;
    ++p->version;
    return kmx_sweep_scales_updated(p, term);
}

int kmx_plan_destroy(kmx_plan* p) {
    if (!p) return KMX_OK;
    free_plan(p);
    delete p;
    return KMX_OK;
}

int kmx_plan_info(const kmx_plan* p, int64_t* num_pairs, int* block, int64_t* dof) {
    if (!p) {
        set_error("kmx_plan_info: null plan");
        return KMX_E_ARG;
    }
    if (num_pairs) *num_pairs = p->P;
    if (block) *block = p->B;
    if (dof) *dof = (int64_t)p->N * p->B;
    return KMX_OK;
}

int kmx_apply(kmx_plan* p, const double* x, double* y, int prec, void* stream) {
    if (!p || (p->N > 0 && (!x || !y))) {
        set_error("kmx_apply: null argument");
        return KMX_E_ARG;
    }
    if (x == y && p->N > 0) {
        set_error("kmx_apply: x and y must not alias");
        return KMX_E_ARG;
    }
    return kmx_launch_apply(p, false, x, y, prec, reinterpret_cast<hipStream_t>(stream));
}

int kmx_apply_transpose(kmx_plan* p, const double* x, double* y, int prec, void* stream) {
    if (!p || (p->N > 0 && (!x || !y))) {
        set_error("kmx_apply_transpose: null argument");
        return KMX_E_ARG;
    }
    const int slot = prec == KMX_PREC_FP64 ? 0 : 1;
    const bool on_sweep = prec != KMX_PREC_BF16 && kmx_sweep_has_transpose(p) && p->var[slot].sweep;
    if (!p->d_blkT64 && !on_sweep) {
        set_error("kmx_apply_transpose: plan was created without KMX_PLAN_ADJOINT (and runs no separable sweep)");
        return KMX_E_ARG;
    }
    if (x == y && p->N > 0) {
        set_error("kmx_apply_transpose: x and y must not alias");
        return KMX_E_ARG;
    }
    return kmx_launch_apply(p, true, x, y, prec, reinterpret_cast<hipStream_t>(stream));
}

int kmx_apply_host(kmx_plan* p, const double* xh, double* yh, int prec) {
    if (!p || (p->N > 0 && (!xh || !yh))) {
        set_error("kmx_apply_host: null argument");
        return KMX_E_ARG;
    }
    try {
        const size_t n = size_t(p->N) * p->B;
        if (n == 0) return KMX_OK;
        if (!p->d_x_host_stage) {
            if (hipMalloc(&p->d_x_host_stage, n * sizeof(double)) != hipSuccess ||
                hipMalloc(&p->d_y_host_stage, n * sizeof(double)) != hipSuccess)
                throw NoMem("staging allocation failed");
        }
        hip_check(hipMemcpy(p->d_x_host_stage, xh, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        int rc = kmx_launch_apply(p, false, p->d_x_host_stage, p->d_y_host_stage, prec, nullptr);
        if (rc != KMX_OK) return rc;
        hip_check(hipMemcpy(yh, p->d_y_host_stage, n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        return KMX_OK;
    } catch (const std::exception& e) {
        set_error(std::string("kmx_apply_host: ") + e.what());
        return code_of(e);
    }
}

}  // extern "C"
