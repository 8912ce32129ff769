// kmx_coeff.hip -- the 1D operator of one dimension of one term, built on gfx950:
// sd_term_coeff_matrix.m:1-27 (the chain of pterm matrices) of
// pterm_coeff_matrix.m:1-26 (each pterm's RHS matrix L moved to the RHS with its
// LHS mass matrix M: M \ L) of coeff_matrix.m:1-374 (the DG volume and flux
// terms of a MASS / DIV / GRAD pterm on 2^lev cells, rotated to the multiwavelet
// basis by the forward transform of OperatorTwoScale_wavelet2.m).
//
// Everything stays in the real-space (cell) basis until the end: there M is
// block diagonal (one deg x deg block per cell) and L block banded (the cell and
// its neighbours, wrapped for periodic boundaries), so M \ L is one small solve
// per cell, the chain is a product of block-banded matrices, and the rotation
// W = F R F^T runs once, through the transform's row supports (each row of F is
// nonzero on one contiguous range of columns).  The reference rotates first and
// solves, (F M F^T) \ (F L F^T) = F^-T (M \ L) F^T, which equals F (M \ L) F^T for
// an orthogonal F; the transform built from the two-scale filter data is
// orthogonal to ||F F^T - I|| (about 2e-14 up to deg 3, 6e-14 at deg 4, 2.3e-13 at
// deg 5 and 6), and the two results differ by about that much besides roundoff.
#include "kmx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace kmx {
namespace {

// ---------------------------------------------------------------- host math
// lgwt.m:16-59: n Gauss-Legendre nodes and weights on [a, b], ascending.
void lgwt(int n, double a, double b, std::vector<double>& x, std::vector<double>& w) {
    const int N = n - 1, N1 = N + 1, N2 = N + 2;
    std::vector<double> y(N1), y0(N1, 2.0), Lp(N1);
    for (int i = 0; i < N1; ++i) {
        const double xu = N1 == 1 ? -1.0 : -1.0 + 2.0 * i / (N1 - 1);
        y[i] = std::cos((2.0 * i + 1) * M_PI / (2 * N + 2)) + (0.27 / N1) * std::sin(M_PI * xu * N / N2);
    }
    std::vector<double> L(size_t(N1) * N2);
    for (int it = 0; it < 200; ++it) {
        double dmax = 0.0;
        for (int i = 0; i < N1; ++i) dmax = std::max(dmax, std::fabs(y[i] - y0[i]));
        if (!(dmax > 2.220446049250313e-16)) break;
        for (int i = 0; i < N1; ++i) {
            L[size_t(i) * N2 + 0] = 1.0;
            L[size_t(i) * N2 + 1] = y[i];
            for (int k = 2; k <= N1; ++k)
                L[size_t(i) * N2 + k] =
                    ((2 * k - 1) * y[i] * L[size_t(i) * N2 + k - 1] - (k - 1) * L[size_t(i) * N2 + k - 2]) / k;
            Lp[i] = N1 * (L[size_t(i) * N2 + N1 - 1] - y[i] * L[size_t(i) * N2 + N2 - 1]) / (1 - y[i] * y[i]);
            y0[i] = y[i];
            y[i] = y0[i] - L[size_t(i) * N2 + N2 - 1] / Lp[i];
        }
    }
    x.assign(N1, 0.0);
    w.assign(N1, 0.0);
    for (int i = 0; i < N1; ++i) {  // reversed: ascending nodes
        x[N1 - 1 - i] = (a * (1 - y[i]) + b * (1 + y[i])) / 2;
        w[N1 - 1 - i] = (b - a) / ((1 - y[i] * y[i]) * Lp[i] * Lp[i]);
    }
}

// lin_legendre.m / lin_dlegendre.m: orthonormal Legendre values and derivatives
// (k of them) at x in [-1, 1].
void legendre(double x, int k, double* v, double* dv) {
    std::vector<double> P(std::max(k, 2)), dP(std::max(k, 2));
    P[0] = 1.0;
    dP[0] = 0.0;
    if (k > 1) {
        P[1] = x;
        dP[1] = 1.0;
    }
    for (int n = 2; n < k; ++n) {
        P[n] = ((2 * n - 1) * x * P[n - 1] - (n - 1) * P[n - 2]) / n;
        dP[n] = n * P[n - 1] + x * dP[n - 1];
    }
    for (int n = 0; n < k; ++n) {
        const double s = std::sqrt(2.0 * n + 1.0);
        if (v) v[n] = P[n] * s;
        if (dv) dv[n] = dP[n] * s;
    }
}

// default_quad_number.m
int quad_number(int deg) { return std::max(deg + 1, 10); }

// OperatorTwoScale_wavelet2.m:1-144 as row supports: row r of the deg * 2^lev
// square transform is nonzero on columns [c0[r], c0[r] + len[r]) with values
// val[off[r] ...].  H0 / G0: the deg x deg two-scale filters (row-major), entries
// below 1e4 eps taken as zero as the reference does.
struct Fmwt {
    int n = 0;
    std::vector<int> c0, len;
    std::vector<long long> off;
    std::vector<double> val;
};
Fmwt build_fmwt(int deg, int lev, const double* H0in, const double* G0in) {
    Fmwt F;
    const int n = deg << lev;
    F.n = n;
    if (lev == 0) {
        for (int r = 0; r < n; ++r) {
            F.c0.push_back(0);
            F.len.push_back(n);
            F.off.push_back((long long)F.val.size());
            for (int c = 0; c < n; ++c) F.val.push_back(r == c ? 1.0 : 0.0);
        }
        return F;
    }
    const double tol = 1e4 * 2.220446049250313e-16;
    std::vector<double> H0(size_t(deg) * deg), G0(size_t(deg) * deg), H1(H0.size()), G1(G0.size());
    for (int i = 0; i < deg; ++i)
        for (int j = 0; j < deg; ++j) {
            double h = H0in[i * deg + j], g = G0in[i * deg + j];
            if (std::fabs(h) < tol) h = 0.0;
            if (std::fabs(g) < tol) g = 0.0;
            const double sg = ((i + j) % 2) ? -1.0 : 1.0;
            H0[i * deg + j] = h;
            G0[i * deg + j] = g;
            H1[i * deg + j] = sg * h;
            G1[i * deg + j] = ((deg % 2) ? -1.0 : 1.0) * sg * g;
        }
    std::vector<double> Hmat(size_t(deg) * n, 0.0), Gmat(size_t(deg) * n, 0.0);  // row-major deg x n
    for (int i = 0; i < deg; ++i) Hmat[size_t(i) * n + i] = 1.0;
    std::vector<std::vector<double>> blocks(size_t(2 * lev + 3));
    std::vector<int> bcols(size_t(2 * lev + 3), 0);
    auto mul = [&](const std::vector<double>& A, const std::vector<double>& Htmp, int half, std::vector<double>& out,
                   int c0) {  // out[:, c0:c0+half] = A @ Htmp (deg x half)
        for (int i = 0; i < deg; ++i)
            for (int c = 0; c < half; ++c) {
                double s = 0.0;
                for (int k = 0; k < deg; ++k) s += A[size_t(i) * deg + k] * Htmp[size_t(k) * half + c];
                out[size_t(i) * n + c0 + c] = s;
            }
    };
    for (int j = lev - 1; j >= 0; --j) {
        const int isize = n >> j, half = isize / 2;
        std::vector<double> Htmp(size_t(deg) * half);
        for (int i = 0; i < deg; ++i)
            for (int c = 0; c < half; ++c) Htmp[size_t(i) * half + c] = Hmat[size_t(i) * n + c];
        mul(G0, Htmp, half, Gmat, 0);
        mul(G1, Htmp, half, Gmat, half);
        mul(H0, Htmp, half, Hmat, 0);
        mul(H1, Htmp, half, Hmat, half);
        auto take = [&](const std::vector<double>& M, int cols) {
            std::vector<double> b(size_t(deg) * cols);
            for (int i = 0; i < deg; ++i)
                for (int c = 0; c < cols; ++c) b[size_t(i) * cols + c] = M[size_t(i) * n + c];
            return b;
        };
        blocks[size_t(2 * j + 2)] = take(Gmat, isize);
        bcols[size_t(2 * j + 2)] = isize;
        const int hc = j == 0 ? n : isize;
        blocks[size_t(2 * j + 1)] = take(Hmat, hc);
        bcols[size_t(2 * j + 1)] = hc;
    }
    auto push_row = [&](int c0, int l, const double* v) {
        F.c0.push_back(c0);
        F.len.push_back(l);
        F.off.push_back((long long)F.val.size());
        F.val.insert(F.val.end(), v, v + l);
    };
    for (int i = 0; i < deg; ++i) push_row(0, n, &blocks[1][size_t(i) * n]);  // the coarsest scaling functions
    for (int j = 0; j < lev; ++j) {
        const int cells = 1 << j, isize = n >> j;
        const std::vector<double>& Fm = blocks[size_t(2 + 2 * j)];
        for (int c = 0; c < cells; ++c)
            for (int i = 0; i < deg; ++i) push_row(c * isize, isize, &Fm[size_t(i) * isize]);
    }
    return F;
}

// ---------------------------------------------------------------- device
// Block-banded real-space matrix on N cells: block row i holds the deg x deg
// blocks (i, i + o) for o = -b .. b (cell index mod N when periodic, else absent
// past the ends), row-major inside a block: B[((i * (2b + 1) + (o + b)) * deg + a) * deg + c].
struct Band {
    double* d = nullptr;
    int b = 0;
};

__device__ __forceinline__ long long band_at(int N, int deg, int b, int i, int o, int a, int c) {
    return ((long long)(i * (2 * b + 1) + (o + b)) * deg + a) * deg + c;
}

// One pterm's real-space L (coeff_matrix.m:105-323) as a band of width 1, and the
// LHS mass blocks, one thread per (cell, a, c).
struct CellArgs {
    int N, deg, nq, type, ibcl, ibcr;  // type 0 mass, 1 div, 2 grad; BCs 'D' 'N' 'P' as chars
    double h, LF;
    const double* pq;   // nq x deg: orthonormal Legendre at the nodes / sqrt(h)
    const double* dpq;  // nq x deg: derivatives * 2 / h / sqrt(h)
    const double* qw;   // nq
    const double* pL;   // deg: values at -1 / sqrt(h)
    const double* pR;   // deg: values at +1 / sqrt(h)
    const double* gq;   // N x nq: g dV at the nodes
    const double* gf;   // N + 1: g dV at the faces
    const double* lq;   // N x nq: LHS mass g * moment dV at the nodes
    double* L;          // band b = 1 (div / mass), before any transpose
    double* M;          // N x deg x deg
};
__global__ void __launch_bounds__(256) k_cell_blocks(CellArgs a) {
    const int deg = a.deg;
    const long long tot = (long long)a.N * deg * deg;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int i = int(e / (deg * deg)), r = int(e % (deg * deg)), ra = r / deg, rc = r % deg;
        const double jac = a.h / 2;
        double mass = 0.0, vol = 0.0, lhs = 0.0;
        for (int q = 0; q < a.nq; ++q) {
            const double pw = a.pq[q * deg + rc] * a.qw[q] * jac;
            mass += a.pq[q * deg + ra] * a.gq[(long long)i * a.nq + q] * pw;
            vol -= a.dpq[q * deg + ra] * a.gq[(long long)i * a.nq + q] * pw;
            lhs += a.pq[q * deg + ra] * a.lq[(long long)i * a.nq + q] * pw;
        }
        a.M[e] = lhs;
        double blk[3] = {0.0, a.type == 0 ? mass : vol, 0.0};  // columns of cells i-1, i, i+1
        if (a.type != 0) {
            const int N = a.N;
            const double FCL = a.gf[i], FCR = a.gf[i + 1];
            const double pLa = a.pL[ra], pLc = a.pL[rc], pRa = a.pR[ra], pRc = a.pR[rc];
            double B0 = -pLa * pRc * FCL / 2 + a.LF * fabs(FCL) / 2 * (-pLa * pRc);
            double B1 = -pLa * pLc * FCL / 2 + a.LF * fabs(FCL) / 2 * (pLa * pLc);
            double B2 = pRa * pRc * FCR / 2 + a.LF * fabs(FCR) / 2 * (pRa * pRc);
            double B3 = pRa * pLc * FCR / 2 + a.LF * fabs(FCR) / 2 * (-pRa * pLc);
            if (a.ibcl == 'D' && i == 0) B0 = B1 = 0.0;
            if (a.ibcr == 'D' && i == N - 1) B2 = B3 = 0.0;
            if (a.ibcl == 'N' && i == 0) {
                B0 = 0.0;
                B1 = -pLa * pLc * FCL;
            }
            if (a.ibcr == 'N' && i == N - 1) {
                B2 = pRa * pRc * FCR;
                B3 = 0.0;
            }
            const bool periodic = a.ibcl == 'P' || a.ibcr == 'P';
            bool use0 = i > 0, use3 = i < N - 1;
            if (periodic) use0 = use3 = true;
            if (N == 1) {  // one cell: every flux block lands on it
                blk[1] += (use0 ? B0 : 0.0) + B1 + B2 + (use3 ? B3 : 0.0);
            } else {
                blk[0] += use0 ? B0 : 0.0;
                blk[1] += B1 + B2;
                blk[2] += use3 ? B3 : 0.0;
            }
        }
        for (int o = -1; o <= 1; ++o) a.L[band_at(a.N, deg, 1, i, o, ra, rc)] = blk[o + 1];
    }
}

// GRAD: the matrix is -(DIV)^T (coeff_matrix.m:323): block (i, i + o) of the
// result is -(block (i + o, i) of the input)^T.
__global__ void __launch_bounds__(256) k_band_neg_transpose(const double* __restrict__ in, double* __restrict__ out,
                                                             int N, int deg, int b, int periodic) {
    const long long tot = (long long)N * (2 * b + 1) * deg * deg;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int c = int(e % deg), a = int((e / deg) % deg);
        const long long io = e / (deg * deg);
        const int o = int(io % (2 * b + 1)) - b, i = int(io / (2 * b + 1));
        int j = i + o;
        double v = 0.0;
        if (periodic) j = ((j % N) + N) % N;
        if (j >= 0 && j < N) v = -in[band_at(N, deg, b, j, -o, c, a)];
        out[e] = v;
    }
}

// R = blockdiag(M)^-1 L, per cell: Gaussian elimination with partial pivoting of
// M_i, applied to the cell's 2b + 1 blocks (one thread per cell).
__global__ void __launch_bounds__(64) k_band_mass_solve(const double* __restrict__ M, double* __restrict__ L, int N,
                                                        int deg, int b, int* __restrict__ singular) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= N) return;
    double A[KMX_MAX_DEG * KMX_MAX_DEG];
    int piv[KMX_MAX_DEG];
    for (int e = 0; e < deg * deg; ++e) A[e] = M[(long long)i * deg * deg + e];
    for (int k = 0; k < deg; ++k) {
        int p = k;
        for (int r = k + 1; r < deg; ++r)
            if (fabs(A[r * deg + k]) > fabs(A[p * deg + k])) p = r;
        piv[k] = p;
        if (p != k)
            for (int c = 0; c < deg; ++c) {
                const double t = A[k * deg + c];
                A[k * deg + c] = A[p * deg + c];
                A[p * deg + c] = t;
            }
        if (A[k * deg + k] == 0.0) {
            *singular = 1;
            return;
        }
        for (int r = k + 1; r < deg; ++r) {
            const double f = A[r * deg + k] / A[k * deg + k];
            A[r * deg + k] = f;
            for (int c = k + 1; c < deg; ++c) A[r * deg + c] -= f * A[k * deg + c];
        }
    }
    for (int o = -b; o <= b; ++o) {
        double* X = L + band_at(N, deg, b, i, o, 0, 0);
        for (int c = 0; c < deg; ++c) {
            double col[KMX_MAX_DEG];
            for (int r = 0; r < deg; ++r) col[r] = X[r * deg + c];
            for (int k = 0; k < deg; ++k) {
                const double t = col[k];
                col[k] = col[piv[k]];
                col[piv[k]] = t;
            }
            for (int r = 1; r < deg; ++r)
                for (int k = 0; k < r; ++k) col[r] -= A[r * deg + k] * col[k];
            for (int r = deg - 1; r >= 0; --r) {
                for (int k = r + 1; k < deg; ++k) col[r] -= A[r * deg + k] * col[k];
                col[r] /= A[r * deg + r];
            }
            for (int r = 0; r < deg; ++r) X[r * deg + c] = col[r];
        }
    }
}

// C = A B for bands A (width ba) and B (width bb): C has width ba + bb; block
// (i, i + o) = sum over o1 of A(i, i + o1) B(i + o1, i + o).  One thread per value.
__global__ void __launch_bounds__(256) k_band_mul(const double* __restrict__ A, int ba, const double* __restrict__ B,
                                                  int bb, double* __restrict__ C, int N, int deg, int periodic) {
    const int bc = ba + bb;
    const long long tot = (long long)N * (2 * bc + 1) * deg * deg;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int c = int(e % deg), r = int((e / deg) % deg);
        const long long io = e / (deg * deg);
        const int o = int(io % (2 * bc + 1)) - bc, i = int(io / (2 * bc + 1));
        double s = 0.0;
        for (int o1 = -ba; o1 <= ba; ++o1) {
            const int o2 = o - o1;
            if (o2 < -bb || o2 > bb) continue;
            int k = i + o1;
            if (periodic) k = ((k % N) + N) % N;
            if (k < 0 || k >= N) continue;
            if (!periodic && (i + o < 0 || i + o >= N)) continue;
            for (int t = 0; t < deg; ++t)
                s += A[band_at(N, deg, ba, i, o1, r, t)] * B[band_at(N, deg, bb, k, o2, t, c)];
        }
        C[e] = s;
    }
}

// T = F R, n x n column-major: T[r, j] = sum over k in row r's support with
// cell(k) within b of cell(j) of F[r, k] R[k, j].  One thread per entry, r fastest.
__global__ void __launch_bounds__(256) k_fmwt_left(const int* __restrict__ c0, const int* __restrict__ len,
                                                   const long long* __restrict__ off, const double* __restrict__ fv,
                                                   const double* __restrict__ R, int b, int N, int deg, int periodic,
                                                   double* __restrict__ T) {
    const int n = N * deg;
    const long long tot = (long long)n * n;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int r = int(e % n), j = int(e / n);
        const int cj = j / deg, jc = j % deg;
        const int k0 = c0[r], k1 = c0[r] + len[r];
        const double* f = fv + off[r] - k0;
        double s = 0.0;
        for (int o = -b; o <= b; ++o) {  // source cells ck with ck + (-o) = cj, i.e. R block (ck, cj), offset cj - ck
            int ck = cj - o;
            if (periodic) ck = ((ck % N) + N) % N;
            if (ck < 0 || ck >= N) continue;
            const int ka = std::max(k0, ck * deg), kb = std::min(k1, ck * deg + deg);
            for (int k = ka; k < kb; ++k) s += f[k] * R[band_at(N, deg, b, ck, o, k - ck * deg, jc)];
        }
        T[e] = s;
    }
}

// W = T F^T, n x n column-major: W[r, s] = sum over k in row s's support of
// T[r, k] F[s, k].  One thread per entry, r fastest (F's row s read uniformly).
__global__ void __launch_bounds__(256) k_fmwt_right(const int* __restrict__ c0, const int* __restrict__ len,
                                                    const long long* __restrict__ off, const double* __restrict__ fv,
                                                    const double* __restrict__ T, int n, double* __restrict__ W) {
    const long long tot = (long long)n * n;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
        const int r = int(e % n), s = int(e / n);
        const int k0 = c0[s], k1 = k0 + len[s];
        const double* f = fv + off[s] - k0;
        double acc = 0.0;
        for (int k = k0; k < k1; ++k) acc += T[(long long)k * n + r] * f[k];
        W[e] = acc;
    }
}

template <typename T>
T* dalloc(size_t n, std::vector<void*>& owned) {
    T* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) throw std::bad_alloc();
    owned.push_back(p);
    return p;
}
template <typename T>
T* dup(const std::vector<T>& h, std::vector<void*>& owned, hipStream_t s) {
    T* d = dalloc<T>(h.size(), owned);
    if (!h.empty() && hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess)
        throw std::runtime_error("hipMemcpyAsync failed");
    return d;
}
unsigned grid_for(long long n) { return unsigned(std::min<long long>((n + 255) / 256, 65536)); }

}  // namespace
}  // namespace kmx

using namespace kmx;

extern "C" int kmx_quad_points(int deg, int lev, double xmin, double xmax, int* nq, double* x) {
    if (deg < 1 || deg > KMX_MAX_DEG || lev < 0 || lev > 24 || !(xmax > xmin) || !nq) {
        set_error("kmx_quad_points: bad arguments");
        return KMX_E_ARG;
    }
    const int q = quad_number(deg);
    *nq = q;
    if (!x) return KMX_OK;
    std::vector<double> qx, qw;
    lgwt(q, -1.0, 1.0, qx, qw);
    const long long N = 1LL << lev;
    const double h = (xmax - xmin) / double(N);
    for (long long i = 0; i < N; ++i)
        for (int k = 0; k < q; ++k) x[i * q + k] = ((qx[k] + 1) / 2 + double(i)) * h + xmin;
    return KMX_OK;
}

extern "C" int kmx_fmwt_matrix(int deg, int lev, const double* H0, const double* G0, double* F) {
    if (deg < 1 || deg > KMX_MAX_DEG || lev < 0 || lev > 14 || !H0 || !G0 || !F) {
        set_error("kmx_fmwt_matrix: bad arguments");
        return KMX_E_ARG;
    }
    const Fmwt f = build_fmwt(deg, lev, H0, G0);
    const long long n = f.n;
    std::fill(F, F + n * n, 0.0);
    for (long long r = 0; r < n; ++r)
        for (int k = 0; k < f.len[size_t(r)]; ++k) F[(long long)(f.c0[size_t(r)] + k) * n + r] = f.val[size_t(f.off[size_t(r)] + k)];
    return KMX_OK;
}

extern "C" int kmx_sd_term_matrix(int deg, int lev, double xmin, double xmax, const double* H0, const double* G0,
                                  int npterms, const kmx_pterm_desc* pterms, double* mat, void* stream) {
    if (deg < 1 || deg > KMX_MAX_DEG || lev < 0 || lev > 14 || !(xmax > xmin) || !H0 || !G0 || npterms < 1 ||
        !pterms || !mat) {
        set_error("kmx_sd_term_matrix: bad arguments");
        return KMX_E_ARG;
    }
    for (int p = 0; p < npterms; ++p) {
        const kmx_pterm_desc& t = pterms[p];
        auto bc_ok = [](char c) { return c == 'D' || c == 'N' || c == 'P'; };
        if (t.type < KMX_PTERM_MASS || t.type > KMX_PTERM_GRAD || !t.g_quad || !t.lhs_quad ||
            (t.type != KMX_PTERM_MASS && (!t.g_face || !bc_ok(t.ibcl) || !bc_ok(t.ibcr)))) {
            set_error("kmx_sd_term_matrix: bad pterm description");
            return KMX_E_ARG;
        }
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::vector<void*> owned;
    int rc = KMX_OK;
    try {
        const int N = 1 << lev, n = deg * N, nq = quad_number(deg);
        const double h = (xmax - xmin) / N;
        std::vector<double> qx, qw;
        lgwt(nq, -1.0, 1.0, qx, qw);
        std::vector<double> pq(size_t(nq) * deg), dpq(size_t(nq) * deg), pL(deg), pR(deg);
        for (int q = 0; q < nq; ++q) {
            legendre(qx[q], deg, &pq[size_t(q) * deg], &dpq[size_t(q) * deg]);
            for (int k = 0; k < deg; ++k) {
                pq[size_t(q) * deg + k] /= std::sqrt(h);
                dpq[size_t(q) * deg + k] *= 2.0 / h / std::sqrt(h);
            }
        }
        legendre(-1.0, deg, pL.data(), nullptr);
        legendre(1.0, deg, pR.data(), nullptr);
        for (int k = 0; k < deg; ++k) {
            pL[k] /= std::sqrt(h);
            pR[k] /= std::sqrt(h);
        }
        double* d_pq = dup(pq, owned, s);
        double* d_dpq = dup(dpq, owned, s);
        double* d_qw = dup(qw, owned, s);
        double* d_pL = dup(pL, owned, s);
        double* d_pR = dup(pR, owned, s);
        int* d_sing = dalloc<int>(1, owned);
        if (hipMemsetAsync(d_sing, 0, sizeof(int), s) != hipSuccess) throw std::runtime_error("hipMemsetAsync failed");
        // per pterm: R_p = M_p^-1 L_p (band 1), then the chain R = R_1 ... R_P (band P)
        bool periodic = false;
        double* chain = nullptr;
        int bchain = 0;
        const size_t blk = size_t(N) * deg * deg;
        for (int p = 0; p < npterms; ++p) {
            const kmx_pterm_desc& t = pterms[p];
            const bool per = t.type != KMX_PTERM_MASS && (t.ibcl == 'P' || t.ibcr == 'P');
            periodic = periodic || per;
            std::vector<double> gq(t.g_quad, t.g_quad + size_t(N) * nq), lq(t.lhs_quad, t.lhs_quad + size_t(N) * nq);
            std::vector<double> gf(size_t(N) + 1, 0.0);
            if (t.g_face) gf.assign(t.g_face, t.g_face + N + 1);
            CellArgs a{};
            a.N = N;
            a.deg = deg;
            a.nq = nq;
            a.type = t.type == KMX_PTERM_MASS ? 0 : 1;
            a.ibcl = t.ibcl;
            a.ibcr = t.ibcr;
            a.h = h;
            a.LF = t.LF;
            a.pq = d_pq;
            a.dpq = d_dpq;
            a.qw = d_qw;
            a.pL = d_pL;
            a.pR = d_pR;
            a.gq = dup(gq, owned, s);
            a.gf = dup(gf, owned, s);
            a.lq = dup(lq, owned, s);
            a.L = dalloc<double>(blk * 3, owned);
            a.M = dalloc<double>(blk, owned);
            hipLaunchKernelGGL(k_cell_blocks, dim3(grid_for((long long)blk)), dim3(256), 0, s, a);
            double* R = a.L;
            if (t.type == KMX_PTERM_GRAD) {
                R = dalloc<double>(blk * 3, owned);
                hipLaunchKernelGGL(k_band_neg_transpose, dim3(grid_for((long long)blk * 3)), dim3(256), 0, s, a.L, R,
                                   N, deg, 1, per ? 1 : 0);
            }
            hipLaunchKernelGGL(k_band_mass_solve, dim3((N + 63) / 64), dim3(64), 0, s, a.M, R, N, deg, 1, d_sing);
            if (!chain) {
                chain = R;
                bchain = 1;
            } else {
                const int bc = bchain + 1;
                double* C = dalloc<double>(blk * (2 * bc + 1), owned);
                hipLaunchKernelGGL(k_band_mul, dim3(grid_for((long long)blk * (2 * bc + 1))), dim3(256), 0, s, chain,
                                   bchain, R, 1, C, N, deg, periodic ? 1 : 0);
                chain = C;
                bchain = bc;
            }
        }
        // (a band wider than a periodic mesh names some cells under several offsets;
        // every consumer sums over the offsets, so each product term is counted once)
        // rotation W = F R F^T
        const Fmwt F = build_fmwt(deg, lev, H0, G0);
        int* d_c0 = dup(F.c0, owned, s);
        int* d_len = dup(F.len, owned, s);
        long long* d_off = dup(F.off, owned, s);
        double* d_fv = dup(F.val, owned, s);
        double* T = dalloc<double>(size_t(n) * n, owned);
        hipLaunchKernelGGL(k_fmwt_left, dim3(grid_for((long long)n * n)), dim3(256), 0, s, d_c0, d_len, d_off, d_fv,
                           chain, bchain, N, deg, periodic ? 1 : 0, T);
        hipLaunchKernelGGL(k_fmwt_right, dim3(grid_for((long long)n * n)), dim3(256), 0, s, d_c0, d_len, d_off, d_fv,
                           T, n, mat);
        int sing = 0;
        if (hipMemcpyAsync(&sing, d_sing, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            throw std::runtime_error("kernel or copy failed");
        if (hipGetLastError() != hipSuccess) throw std::runtime_error("launch failed");
        if (sing) {
            set_error("kmx_sd_term_matrix: a cell's LHS mass block is singular");
            rc = KMX_E_ARG;
        }
    } catch (const std::bad_alloc&) {
        set_error("kmx_sd_term_matrix: device allocation failed");
        rc = KMX_E_NOMEM;
    } catch (const std::exception& e) {
        set_error(std::string("kmx_sd_term_matrix: ") + e.what());
        rc = KMX_E_HIP;
    }
    hipStreamSynchronize(s);
    for (void* p : owned) hipFree(p);
    return rc;
}
