// kmx_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU, IEEE-FP64 restatement of ASGarD-MATLAB's operator-application path
// (project-asgard/DG-SparseGrid).  It is the checker for the HIP library and
// nothing else: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it.  The product never links it.
//
// Every function cites the reference file:line it follows.  Conventions are
// MATLAB's: 1-based element ids and 1D indices, column-major dense matrices,
// element blocks with dim 1 slowest and dim d fastest (krond order).
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   perm_* / index_leq_max      gold/generated-inputs/permutations/*
//   hash tables                 gold/generated-inputs/element_table/*
//   connect_1D / connect_nD     gold/generated-inputs/connectivity/*
//   kronmult / kron_multd       closed form kron(A1..Ad)*X (numpy.kron), tiny
//   apply_A                     matrix-free == assembled (krond path),
//                               2D full-grid closed form, gold RK3 trajectories
//   gemm_vbatched               per-entry closed form (numpy matmul)
//   kronmult*_batch lists       replaying the lists reproduces kronmult
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Rows = std::vector<std::vector<int>>;
thread_local std::string g_err;

// ---------------------------------------------------------------- PERMUTE/
// PERMUTE/perm_eq_d.m:1-80 (tuples with sum == SumEq, i-th entry <= LevMax(i);
// the last index runs in the outer loop, ascending unless last_index_decreasing)
Rows perm_eq_d(int idim, const std::vector<int>& levmax, int sumeq, bool lid) {
    Rows out;
    if (idim == 1) {
        if (levmax[0] >= sumeq) out.push_back({sumeq});
        return out;
    }
    int top = std::min(levmax[idim - 1], sumeq);
    for (int s = 0; s <= top; ++s) {
        int ilast = lid ? (top - s) : s;
        std::vector<int> lm(levmax.begin(), levmax.begin() + (idim - 1));
        Rows sub = perm_eq_d(idim - 1, lm, sumeq - ilast, lid);
        for (auto& r : sub) {
            r.push_back(ilast);
            out.push_back(std::move(r));
        }
    }
    return out;
}

// PERMUTE/perm_leq_d.m:1-46 (all sums SumMax..0, descending by default)
Rows perm_leq_d(int idim, const std::vector<int>& levmax, int summax, bool increasing) {
    Rows out;
    for (int k = 0; k <= summax; ++k) {
        int s = increasing ? k : (summax - k);
        Rows part = perm_eq_d(idim, levmax, s, false);
        out.insert(out.end(), part.begin(), part.end());
    }
    return out;
}

// PERMUTE/perm_eq.m:1-48
Rows perm_eq(int idim, int n, bool lid) {
    Rows out;
    if (idim == 1) {
        out.push_back({n});
        return out;
    }
    for (int i = 0; i <= n; ++i) {
        Rows sub = lid ? perm_eq(idim - 1, i, lid) : perm_eq(idim - 1, n - i, lid);
        int last = lid ? (n - i) : i;
        for (auto& r : sub) {
            r.push_back(last);
            out.push_back(std::move(r));
        }
    }
    return out;
}

// PERMUTE/perm_leq.m:1-44
Rows perm_leq(int idim, int n, bool order_by_n) {
    Rows out;
    if (order_by_n) {
        for (int i = 0; i <= n; ++i) {
            Rows part = perm_eq(idim, i, false);
            out.insert(out.end(), part.begin(), part.end());
        }
        return out;
    }
    if (idim == 1) {
        for (int i = 0; i <= n; ++i) out.push_back({i});
        return out;
    }
    for (int i = 0; i <= n; ++i) {
        Rows sub = perm_leq(idim - 1, n - i, order_by_n);
        for (auto& r : sub) {
            r.push_back(i);
            out.push_back(std::move(r));
        }
    }
    return out;
}

// PERMUTE/perm_max.m:1-63
Rows perm_max(int idim, const std::vector<int>& levmax, bool lid) {
    Rows out;
    int n = levmax[idim - 1];
    if (idim == 1) {
        for (int i = 0; i <= n; ++i) out.push_back({lid ? n - i : i});
        return out;
    }
    Rows ivec = perm_max(idim - 1, levmax, lid);
    for (int i = 0; i <= n; ++i)
        for (const auto& r0 : ivec) {
            auto r = r0;
            r.push_back(lid ? n - i : i);
            out.push_back(std::move(r));
        }
    return out;
}

// PERMUTE/index_leq_max.m:1-47 (1-based positions into each Nmat{i})
Rows index_leq_max(int ndim, const std::vector<std::vector<int>>& nmat, int levsum, int levmax) {
    Rows out;
    if (ndim == 1) {
        const auto& v = nmat[0];
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] <= levsum && v[i] <= levmax) out.push_back({int(i) + 1});
        return out;
    }
    const auto& v = nmat[ndim - 1];
    for (size_t i = 0; i < v.size(); ++i) {
        if (!(v[i] <= levmax)) continue;
        Rows sub = index_leq_max(ndim - 1, nmat, levsum - v[i], levmax);
        for (auto& r : sub) {
            r.push_back(int(i) + 1);
            out.push_back(std::move(r));
        }
    }
    return out;
}

// PERMUTE/index_max.m:1-40
Rows index_max(int ndim, const std::vector<std::vector<int>>& nmat, int levmax) {
    Rows out;
    if (ndim == 1) {
        const auto& v = nmat[0];
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] <= levmax) out.push_back({int(i) + 1});
        return out;
    }
    Rows sub = index_max(ndim - 1, nmat, levmax);
    const auto& v = nmat[ndim - 1];
    for (size_t i = 0; i < v.size(); ++i) {
        if (!(v[i] <= levmax)) continue;
        for (const auto& r0 : sub) {
            auto r = r0;
            r.push_back(int(i) + 1);
            out.push_back(std::move(r));
        }
    }
    return out;
}

// ------------------------------------------------------------ element table
// lev_cell_to_1D_index.m:6-9
inline int lev_cell_to_1d(int lev, int cell) { return lev == 0 ? 1 : (1 << (lev - 1)) + cell + 1; }

// lev_cell_to_1D_index_set.m:1-37 (dim 1 fastest)
Rows lev_cell_set(const std::vector<int>& levels) {
    int dim = int(levels.size());
    std::vector<int> isz(dim);
    long total = 1;
    for (int d = 0; d < dim; ++d) {
        isz[d] = 1 << std::max(0, levels[d] - 1);
        total *= isz[d];
    }
    Rows out;
    out.reserve(total);
    std::vector<int> c(dim, 0);
    for (long n = 0; n < total; ++n) {
        out.push_back(c);
        for (int d = 0; d < dim; ++d) {
            if (++c[d] < isz[d]) break;
            c[d] = 0;
        }
    }
    return out;
}

struct Elements {
    std::vector<int> lev, cell, idx1d;  // N x d, row-major
    std::vector<uint64_t> id;           // md_lev_pos_to_idx ids (1-based)
    int n = 0;
};

// md_lev_pos_to_idx.m:1-28
uint64_t md_lev_pos_to_idx(const std::vector<int>& lev, const std::vector<int>& pos, int max_lev) {
    uint64_t idx = 1, stride = 1;
    for (size_t d = 0; d < lev.size(); ++d) {
        if (lev[d] > max_lev) throw std::runtime_error("md_lev_pos_to_idx: lev > max_lev");
        idx += uint64_t(lev_cell_to_1d(lev[d], pos[d]) - 1) * stride;
        stride *= (uint64_t(1) << max_lev);
    }
    return idx;
}

void append_cases(const Rows& ptable, int max_lev, Elements& e) {
    for (const auto& levels : ptable) {
        Rows cells = lev_cell_set(levels);
        for (const auto& c : cells) {
            for (size_t d = 0; d < levels.size(); ++d) {
                e.lev.push_back(levels[d]);
                e.cell.push_back(c[d]);
                e.idx1d.push_back(lev_cell_to_1d(levels[d], c[d]));
            }
            e.id.push_back(max_lev >= 0 ? md_lev_pos_to_idx(levels, c, max_lev) : 0);
            e.n++;
        }
    }
}

// hash_table_sparse_nD.m:1-118 (new hash: descending level-sum order)
Elements hash_table_sparse_nd(const std::vector<int>& lev_vec, int max_lev, bool sg) {
    int nd = int(lev_vec.size());
    Rows ptable = sg ? perm_leq_d(nd, lev_vec, *std::max_element(lev_vec.begin(), lev_vec.end()), false)
                     : perm_max(nd, lev_vec, false);
    Elements e;
    append_cases(ptable, max_lev, e);
    return e;
}

// hash_table_nD.m:1-142 (old hash: perm_leq order, uniform level)
Elements hash_table_nd(int nd, int lev, bool sg) {
    Rows ptable = sg ? perm_leq(nd, lev, false) : perm_max(nd, std::vector<int>(nd, lev), false);
    Elements e;
    append_cases(ptable, -1, e);
    return e;
}

// connect_1D.m:1-51 (dense 0/1, 2^Lev square, 1-based I,J stored 0-based)
std::vector<uint8_t> connect_1d(int Lev) {
    int n = 1 << Lev;
    std::vector<uint8_t> con(size_t(n) * n, 0);
    auto set = [&](int I, int J) { con[size_t(I - 1) * n + (J - 1)] = 1; con[size_t(J - 1) * n + (I - 1)] = 1; };
    for (int Lx = 0; Lx <= Lev; ++Lx) {
        int ncx = 1 << std::max(0, Lx - 1);
        for (int Px = 0; Px < ncx; ++Px) {
            int I = lev_cell_to_1d(Lx, Px);
            for (int q = std::max(Px - 1, 0); q <= std::min(Px + 1, ncx - 1); ++q) set(I, lev_cell_to_1d(Lx, q));
            if (Px == 0) set(I, lev_cell_to_1d(Lx, ncx - 1));
            else if (Px == ncx - 1) set(I, lev_cell_to_1d(Lx, 0));
            for (int Ly = Lx + 1; Ly <= Lev; ++Ly) {
                int nL = Ly - Lx;
                int ncy = 1 << (Ly - 1);
                int lo, hi;
                if (Lx > 0) {
                    lo = std::max((1 << nL) * Px - 1, 0);
                    hi = std::min((1 << nL) * Px + (1 << nL), ncy - 1);
                } else {
                    lo = std::max((1 << (nL - 1)) * Px - 1, 0);
                    hi = std::min((1 << (nL - 1)) * Px + (1 << (nL - 1)), ncy - 1);
                }
                set(I, lev_cell_to_1d(Ly, 0));
                for (int q = lo; q <= hi; ++q) set(I, lev_cell_to_1d(Ly, q));
                set(I, lev_cell_to_1d(Ly, ncy - 1));
            }
        }
    }
    return con;
}

// connect_nD.m:1-129 over the old hash's inverse table; result rows sorted
void connect_nd(int ndim, const Elements& e, int levsum, int levmax, bool sg,
                std::vector<int64_t>& row_ptr, std::vector<int>& cols) {
    int Lev = std::max(levsum, levmax);
    int n1 = 1 << Lev;
    std::vector<uint8_t> c1 = connect_1d(Lev);
    // Key1dMesh (connect_nD.m:45-53): 1D index -> (lev, cell)
    std::vector<int> klev, kcell;
    for (int Lx = 0; Lx <= Lev; ++Lx)
        for (int Px = 0; Px < (1 << std::max(0, Lx - 1)); ++Px) { klev.push_back(Lx); kcell.push_back(Px); }
    std::map<std::vector<int>, int> hash;  // key [lev.., cell..] -> 1-based element
    for (int i = 0; i < e.n; ++i) {
        std::vector<int> key(2 * ndim);
        for (int d = 0; d < ndim; ++d) { key[d] = e.lev[i * ndim + d]; key[ndim + d] = e.cell[i * ndim + d]; }
        hash[key] = i + 1;
    }
    row_ptr.assign(1, 0);
    cols.clear();
    std::vector<std::vector<int>> levarr(ndim), cellarr(ndim);
    for (int ii = 0; ii < e.n; ++ii) {
        for (int d = 0; d < ndim; ++d) {
            int I = e.idx1d[ii * ndim + d];
            levarr[d].clear();
            cellarr[d].clear();
            for (int J = 1; J <= n1; ++J)
                if (c1[size_t(I - 1) * n1 + (J - 1)]) { levarr[d].push_back(klev[J - 1]); cellarr[d].push_back(kcell[J - 1]); }
        }
        Rows res = sg ? index_leq_max(ndim, levarr, levsum, levmax) : index_max(ndim, levarr, levmax);
        std::vector<int> idxJ;
        idxJ.reserve(res.size());
        std::vector<int> key(2 * ndim);
        for (const auto& r : res) {
            for (int d = 0; d < ndim; ++d) { key[d] = levarr[d][r[d] - 1]; key[ndim + d] = cellarr[d][r[d] - 1]; }
            auto it = hash.find(key);
            if (it == hash.end()) throw std::runtime_error("connect_nD: key not in hash table");
            idxJ.push_back(it->second);
        }
        std::sort(idxJ.begin(), idxJ.end());  // connect_nD.m:101-103
        cols.insert(cols.end(), idxJ.begin(), idxJ.end());
        row_ptr.push_back(int64_t(cols.size()));
    }
}

// -------------------------------------------------------------- Kronmult/
// Column-major dense matrix view.
struct Mat {
    int r, c;
    const double* p;
    double operator()(int i, int j) const { return p[size_t(j) * r + i]; }
};

// C(m x n) = opA(m x k) * opB(k x n), column-major, sums over k ascending.
void matmul(bool ta, bool tb, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
            double* C, int ldc, double alpha, double beta) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) {
                double a = ta ? A[size_t(i) * lda + l] : A[size_t(l) * lda + i];
                double b = tb ? B[size_t(l) * ldb + j] : B[size_t(j) * ldb + l];
                s += a * b;
            }
            double& c = C[size_t(j) * ldc + i];
            c = (beta == 0.0 ? 0.0 : beta * c) + alpha * s;
        }
}

// Kronmult/kronmult1.m:17  Y = A1 * reshape(X, ncol1, [])
std::vector<double> kronmult1(const Mat& A1, const std::vector<double>& X) {
    int nv = int(X.size() / A1.c);
    std::vector<double> Y(size_t(A1.r) * nv);
    matmul(false, false, A1.r, nv, A1.c, A1.p, A1.r, X.data(), A1.c, Y.data(), A1.r, 1.0, 0.0);
    return Y;
}

std::vector<double> kronmult_k(const std::vector<Mat>& A, size_t off, const std::vector<double>& X);

// Kronmult/kronmult2.m:27-39 (method by flop count), :75-140 method 1, :141-160 method 2
std::vector<double> kronmult2(const Mat& A1, const Mat& A2, const std::vector<double>& X) {
    int nrowX = A1.c * A2.c;
    int nvec = int(X.size() / nrowX);
    double f1 = 2.0 * (double(A2.r) * A1.c) * (A2.c + A1.r);
    double f2 = 2.0 * (double(A2.r) + A1.c) * (double(A2.c) * A1.r);
    int nrowY = A1.r * A2.r;
    std::vector<double> Y(size_t(nrowY) * nvec);
    if (f1 <= f2) {
        std::vector<double> T = kronmult1(A2, X);  // (nrow2) x (ncol1*nvec)
        int n1 = A2.r;
        for (int v = 0; v < nvec; ++v)
            matmul(false, true, n1, A1.r, A1.c, T.data() + size_t(v) * n1 * A1.c, n1, A1.p, A1.r,
                   Y.data() + size_t(v) * nrowY, n1, 1.0, 0.0);
    } else {
        // Ytmp(:, block v) = reshape(X_v, ncol2, ncol1) * A1^T ; Y = A2 * Ytmp
        std::vector<double> T(size_t(A2.c) * A1.r * nvec);
        for (int v = 0; v < nvec; ++v)
            matmul(false, true, A2.c, A1.r, A1.c, X.data() + size_t(v) * nrowX, A2.c, A1.p, A1.r,
                   T.data() + size_t(v) * A2.c * A1.r, A2.c, 1.0, 0.0);
        matmul(false, false, A2.r, A1.r * nvec, A2.c, A2.p, A2.r, T.data(), A2.c, Y.data(), A2.r, 1.0, 0.0);
    }
    return Y;
}

// Kronmult/kronmult3..6.m method 1 (always_use_method1 = 1, kronmult3.m:5,73-89;
// kronmult6.m:62-82): Ytmp = kronmult_{k-1}(A2..Ak, X); Y_v = Ytmp_v * A1^T.
std::vector<double> kronmult_k(const std::vector<Mat>& A, size_t off, const std::vector<double>& X) {
    size_t k = A.size() - off;
    if (k == 1) return kronmult1(A[off], X);
    if (k == 2) return kronmult2(A[off], A[off + 1], X);
    long ncolX = 1, nrowY = 1;
    for (size_t i = off; i < A.size(); ++i) { ncolX *= A[i].c; nrowY *= A[i].r; }
    int nvec = int(X.size() / ncolX);
    std::vector<double> T = kronmult_k(A, off + 1, X);
    const Mat& A1 = A[off];
    long nrowT = long(T.size()) / nvec;
    int msize = int(nrowT / A1.c);
    std::vector<double> Y(size_t(nrowY) * nvec);
    for (int v = 0; v < nvec; ++v)
        matmul(false, true, msize, A1.r, A1.c, T.data() + size_t(v) * nrowT, msize, A1.p, A1.r,
               Y.data() + size_t(v) * nrowY, msize, 1.0, 0.0);
    return Y;
}

// kron_multd.m:1-87 (dispatch 1..6, recursion through transposes beyond 6)
std::vector<double> kron_multd(const std::vector<Mat>& A, const std::vector<double>& X) {
    size_t nk = A.size();
    if (nk <= 6) return kronmult_k(A, 0, X);
    long isX = 1, isY = 1;
    for (const auto& a : A) { isX *= a.c; isY *= a.r; }
    long nvec = long(X.size()) / isX;
    std::vector<double> Y(size_t(isY) * nvec);
    const Mat& An = A.back();
    std::vector<Mat> head(A.begin(), A.end() - 1);
    for (long v = 0; v < nvec; ++v) {
        std::vector<double> Xk(X.begin() + v * isX, X.begin() + (v + 1) * isX);
        int nc = An.c;
        int rest = int(isX / nc);
        std::vector<double> Z(size_t(An.r) * rest);
        matmul(false, false, An.r, rest, nc, An.p, An.r, Xk.data(), nc, Z.data(), An.r, 1.0, 0.0);
        // transpose(Z): rest x An.r
        std::vector<double> Zt(Z.size());
        for (int i = 0; i < An.r; ++i)
            for (int j = 0; j < rest; ++j) Zt[size_t(i) * rest + j] = Z[size_t(j) * An.r + i];
        std::vector<double> Yt = kron_multd(head, Zt);  // (isY/An.r) x An.r
        int hr = int(isY / An.r);
        for (int i = 0; i < hr; ++i)
            for (int j = 0; j < An.r; ++j) Y[size_t(v) * isY + size_t(i) * An.r + j] = Yt[size_t(j) * hr + i];
    }
    return Y;
}

// krond.m:7-17  Y = kron(A1, ..., An)
std::vector<double> krond(const std::vector<Mat>& A, int& rows, int& colsOut) {
    std::vector<double> Y(A[0].p, A[0].p + size_t(A[0].r) * A[0].c);
    int r = A[0].r, c = A[0].c;
    for (size_t k = 1; k < A.size(); ++k) {
        const Mat& B = A[k];
        int R = r * B.r, C = c * B.c;
        std::vector<double> Z(size_t(R) * C);
        for (int j1 = 0; j1 < c; ++j1)
            for (int i1 = 0; i1 < r; ++i1) {
                double a = Y[size_t(j1) * r + i1];
                for (int j2 = 0; j2 < B.c; ++j2)
                    for (int i2 = 0; i2 < B.r; ++i2)
                        Z[size_t(j1 * B.c + j2) * R + (i1 * B.r + i2)] = a * B(i2, j2);
            }
        Y.swap(Z);
        r = R;
        c = C;
    }
    rows = r;
    colsOut = c;
    return Y;
}

// ------------------------------------------------------------- apply_A.m
// In-place method-1 chain on one square block (deg^d values, dim d fastest):
// contract dim d first and dim 1 last, each output summing over b ascending --
// the operation order of kronmult{d} method 1 for square deg x deg factors.
void kron_block(int ndim, int deg, const double* const* blk /*ndim deg x deg column-major*/,
                double* v, double* w) {
    long B = 1;
    for (int k = 0; k < ndim; ++k) B *= deg;
    long stride = 1;  // stride of dim k in the block: dim ndim-1 has stride 1
    for (int k = ndim - 1; k >= 0; --k) {
        const double* A = blk[k];
        long outer = B / (stride * deg);
        for (long o = 0; o < outer; ++o)
            for (long s = 0; s < stride; ++s) {
                const double* in = v + o * stride * deg + s;
                double* out = w + o * stride * deg + s;
                for (int a = 0; a < deg; ++a) {
                    double acc = 0.0;
                    for (int b = 0; b < deg; ++b) acc += A[size_t(b) * deg + a] * in[b * stride];
                    out[a * stride] = acc;
                }
            }
        std::swap(v, w);
        stride *= deg;
    }
    if (ndim % 2 == 1) std::memcpy(w, v, sizeof(double) * B);  // result back in caller's v
}

struct ApplyArgs {
    int ndim, deg, N, T;
    const int* n;                // per-dim 1D matrix size (deg * 2^L_d)
    const double* const* mats;   // T*ndim column-major n_d x n_d, index t*ndim + d
    const int* idx1d;            // N x ndim, 1-based
    const int64_t* row_ptr;      // N+1 or null (all pairs)
    const int* cols;             // 1-based element ids
};

// apply_A.m:100-237, explicit branch: loop elem -> j -> t, ftmpA(i1:i2) += Y.
void apply_rows(const ApplyArgs& a, const double* x, double* y, long r0, long r1, bool use_recursion) {
    int d = a.ndim, deg = a.deg;
    long B = 1;
    for (int k = 0; k < d; ++k) B *= deg;
    std::vector<double> blkbuf(size_t(d) * deg * deg), v(B), w(B);
    std::vector<const double*> blk(d);
    for (int k = 0; k < d; ++k) blk[k] = blkbuf.data() + size_t(k) * deg * deg;
    for (long e = r0; e < r1; ++e) {
        double* ye = y + e * B;
        std::fill(ye, ye + B, 0.0);  // ftmp = zeros (apply_A.m:11)
        long nc = a.row_ptr ? long(a.row_ptr[e + 1] - a.row_ptr[e]) : a.N;
        for (long jj = 0; jj < nc; ++jj) {
            long j = a.row_ptr ? long(a.cols[a.row_ptr[e] + jj]) - 1 : jj;
            const double* xj = x + j * B;
            for (int t = 0; t < a.T; ++t) {
                // kron_mat_list{d} = mat(Index_I{d}, Index_J{d})  (apply_A.m:163-168)
                for (int k = 0; k < d; ++k) {
                    int I = a.idx1d[e * d + k] - 1, J = a.idx1d[j * d + k] - 1;
                    const double* M = a.mats[t * d + k];
                    int nk = a.n[k];
                    double* dst = blkbuf.data() + size_t(k) * deg * deg;
                    for (int cb = 0; cb < deg; ++cb)
                        for (int rb = 0; rb < deg; ++rb)
                            dst[size_t(cb) * deg + rb] = M[size_t(J * deg + cb) * nk + (I * deg + rb)];
                }
                if (use_recursion) {
                    std::vector<Mat> ml;
                    for (int k = 0; k < d; ++k) ml.push_back(Mat{deg, deg, blk[k]});
                    std::vector<double> X(xj, xj + B);
                    std::vector<double> Y = kron_multd(ml, X);
                    for (long q = 0; q < B; ++q) ye[q] += Y[q];
                } else {
                    std::copy(xj, xj + B, v.begin());
                    kron_block(d, deg, blk.data(), v.data(), w.data());
                    for (long q = 0; q < B; ++q) ye[q] += v[q];
                }
            }
        }
    }
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

#define ORC_TRY(...)                        \
    try {                                   \
        __VA_ARGS__                         \
    } catch (const std::exception& ex) {    \
        g_err = ex.what();                  \
        return -1;                          \
    }

static long emit(const Rows& r, int ncol, int* out) {
    if (out)
        for (size_t i = 0; i < r.size(); ++i)
            for (int c = 0; c < ncol; ++c) out[i * ncol + c] = r[i][c];
    return long(r.size());
}

long orc_perm_eq_d(int idim, const int* levmax, int sumeq, int lid, int* out) {
    ORC_TRY(return emit(perm_eq_d(idim, std::vector<int>(levmax, levmax + idim), sumeq, lid != 0), idim, out);)
}
long orc_perm_leq_d(int idim, const int* levmax, int summax, int increasing, int* out) {
    ORC_TRY(return emit(perm_leq_d(idim, std::vector<int>(levmax, levmax + idim), summax, increasing != 0), idim, out);)
}
long orc_perm_eq(int idim, int n, int lid, int* out) { ORC_TRY(return emit(perm_eq(idim, n, lid != 0), idim, out);) }
long orc_perm_leq(int idim, int n, int order_by_n, int* out) {
    ORC_TRY(return emit(perm_leq(idim, n, order_by_n != 0), idim, out);)
}
long orc_perm_max(int idim, const int* levmax, int lid, int* out) {
    ORC_TRY(return emit(perm_max(idim, std::vector<int>(levmax, levmax + idim), lid != 0), idim, out);)
}
// nmat: concatenated lists, lens[ndim]
long orc_index_leq_max(int ndim, const int* vals, const int* lens, int levsum, int levmax, int* out) {
    ORC_TRY({
        std::vector<std::vector<int>> nm(ndim);
        for (int d = 0, o = 0; d < ndim; o += lens[d], ++d) nm[d].assign(vals + o, vals + o + lens[d]);
        return emit(index_leq_max(ndim, nm, levsum, levmax), ndim, out);
    })
}

int orc_lev_cell_to_1d(int lev, int cell) { return lev_cell_to_1d(lev, cell); }

// New hash.  Returns N; with non-null outputs fills lev/cell/idx1d (N x nd) and ids.
long orc_hash_table_sparse_nd(int nd, const int* lev_vec, int max_lev, int sg, int* lev, int* cell, int* idx1d,
                              uint64_t* ids) {
    ORC_TRY({
        Elements e = hash_table_sparse_nd(std::vector<int>(lev_vec, lev_vec + nd), max_lev, sg != 0);
        if (lev) std::copy(e.lev.begin(), e.lev.end(), lev);
        if (cell) std::copy(e.cell.begin(), e.cell.end(), cell);
        if (idx1d) std::copy(e.idx1d.begin(), e.idx1d.end(), idx1d);
        if (ids) std::copy(e.id.begin(), e.id.end(), ids);
        return long(e.n);
    })
}

// Old hash (uniform level).
long orc_hash_table_nd(int nd, int lev, int sg, int* levo, int* cell, int* idx1d) {
    ORC_TRY({
        Elements e = hash_table_nd(nd, lev, sg != 0);
        if (levo) std::copy(e.lev.begin(), e.lev.end(), levo);
        if (cell) std::copy(e.cell.begin(), e.cell.end(), cell);
        if (idx1d) std::copy(e.idx1d.begin(), e.idx1d.end(), idx1d);
        return long(e.n);
    })
}

int orc_connect_1d(int lev, uint8_t* out) {
    ORC_TRY({
        auto c = connect_1d(lev);
        std::copy(c.begin(), c.end(), out);
        return 0;
    })
}

// connect_nD over the old hash of (nd, lev, grid).  Two calls: cols == null
// returns P (and fills row_ptr), then fill cols.
long orc_connect_nd(int nd, int lev, int sg_table, int levsum, int levmax, int sg_filter, int64_t* row_ptr,
                    int* cols) {
    ORC_TRY({
        Elements e = hash_table_nd(nd, lev, sg_table != 0);
        std::vector<int64_t> rp;
        std::vector<int> cc;
        connect_nd(nd, e, levsum, levmax, sg_filter != 0, rp, cc);
        if (row_ptr) std::copy(rp.begin(), rp.end(), row_ptr);
        if (cols) std::copy(cc.begin(), cc.end(), cols);
        return long(cc.size());
    })
}

// Y = kron(A1..Ak) X via the reference recursion (kron_multd).  A[k] column-major nrow[k] x ncol[k].
int orc_kron_multd(int k, const int* nrow, const int* ncol, const double* const* A, const double* X, long nvec,
                   double* Y) {
    ORC_TRY({
        std::vector<Mat> ml;
        long isX = 1;
        for (int i = 0; i < k; ++i) { ml.push_back(Mat{nrow[i], ncol[i], A[i]}); isX *= ncol[i]; }
        std::vector<double> Xv(X, X + isX * nvec);
        std::vector<double> Yv = kron_multd(ml, Xv);
        std::copy(Yv.begin(), Yv.end(), Y);
        return 0;
    })
}

int orc_krond(int k, const int* nrow, const int* ncol, const double* const* A, double* Y) {
    ORC_TRY({
        std::vector<Mat> ml;
        for (int i = 0; i < k; ++i) ml.push_back(Mat{nrow[i], ncol[i], A[i]});
        int r, c;
        std::vector<double> Yv = krond(ml, r, c);
        std::copy(Yv.begin(), Yv.end(), Y);
        return 0;
    })
}

// apply_A explicit branch.  mode 0: in-place method-1 chain; mode 1: through the
// kron_multd recursion (slow, for cross-checking mode 0).  nthreads > 1 splits
// element rows over OpenMP threads (each row is still summed in reference order).
int orc_apply_A(int ndim, int deg, int N, int T, const int* n, const double* const* mats, const int* idx1d,
                const int64_t* row_ptr, const int* cols, const double* x, double* y, int mode, int nthreads) {
    ApplyArgs a{ndim, deg, N, T, n, mats, idx1d, row_ptr, cols};
    std::string err;
#ifdef _OPENMP
    if (nthreads > 1) {
#pragma omp parallel num_threads(nthreads)
        {
            try {
#pragma omp for schedule(dynamic, 4)
                for (long e = 0; e < N; ++e) apply_rows(a, x, y, e, e + 1, mode == 1);
            } catch (const std::exception& ex) {
#pragma omp critical
                err = ex.what();
            }
        }
    } else
#endif
    {
        (void)nthreads;
        try {
            apply_rows(a, x, y, 0, N, mode == 1);
        } catch (const std::exception& ex) {
            err = ex.what();
        }
    }
    if (!err.empty()) {
        g_err = err;
        return -1;
    }
    return 0;
}

// apply_A.m:170-180 build_A branch: dense A (column-major, N*B square),
// A(global_row, global_col) += krond(kron_mat_list).
int orc_build_A(int ndim, int deg, int N, int T, const int* n, const double* const* mats, const int* idx1d,
                const int64_t* row_ptr, const int* cols, double* A) {
    ORC_TRY({
        long B = 1;
        for (int k = 0; k < ndim; ++k) B *= deg;
        long dof = B * N;
        std::fill(A, A + dof * dof, 0.0);
        std::vector<double> blkbuf(size_t(ndim) * deg * deg);
        for (long e = 0; e < N; ++e) {
            long nc = row_ptr ? long(row_ptr[e + 1] - row_ptr[e]) : N;
            for (long jj = 0; jj < nc; ++jj) {
                long j = row_ptr ? long(cols[row_ptr[e] + jj]) - 1 : jj;
                for (int t = 0; t < T; ++t) {
                    std::vector<Mat> ml;
                    for (int k = 0; k < ndim; ++k) {
                        int I = idx1d[e * ndim + k] - 1, J = idx1d[j * ndim + k] - 1;
                        const double* M = mats[t * ndim + k];
                        double* dst = blkbuf.data() + size_t(k) * deg * deg;
                        for (int cb = 0; cb < deg; ++cb)
                            for (int rb = 0; rb < deg; ++rb)
                                dst[size_t(cb) * deg + rb] = M[size_t(J * deg + cb) * n[k] + (I * deg + rb)];
                        ml.push_back(Mat{deg, deg, dst});
                    }
                    int r, c;
                    std::vector<double> view = krond(ml, r, c);
                    for (long cc = 0; cc < B; ++cc)
                        for (long rr = 0; rr < B; ++rr) A[size_t(j * B + cc) * dof + (e * B + rr)] += view[size_t(cc) * B + rr];
                }
            }
        }
        return 0;
    })
}

// Kronmult/gemm.m:1-29 and gemm_vbatched.m:1-38 (with k taken from klist; the
// reference's kk = nlist(ibatch) at gemm_vbatched.m:25 is a typo that only
// coincides with klist for square problems).  Column-major, leading dims given.
int orc_gemm_vbatched(char ta, char tb, const int* m, const int* n, const int* k, double alpha,
                      const double* const* A, const int* lda, const double* const* B, const int* ldb, double beta,
                      double* const* C, const int* ldc, int count) {
    ORC_TRY({
        bool tA = (ta == 'T' || ta == 't' || ta == 'C' || ta == 'c');
        bool tB = (tb == 'T' || tb == 't' || tb == 'C' || tb == 'c');
        for (int i = 0; i < count; ++i) matmul(tA, tB, m[i], n[i], k[i], A[i], lda[i], B[i], ldb[i], C[i], ldc[i], alpha, beta);
        return 0;
    })
}

// wavelet_to_realspace.m:44-79: for each element, the deg-column blocks of
// every dimension's Meval (npts_d x ncols_d, column-major) at the element's 1D
// index form a rectangular Kronecker product applied to the element's
// coefficients (kron_multd); the results are summed.
int orc_wavelet_to_realspace(int nd, int deg, int N, const int* idx1d, const int* npts, const int* ncols,
                             const double* const* meval, const double* fw, double* fr) {
    ORC_TRY({
        long B = 1, npt = 1;
        for (int d = 0; d < nd; ++d) {
            B *= deg;
            npt *= npts[d];
        }
        std::fill(fr, fr + npt, 0.0);
        std::vector<std::vector<double>> cols(nd);
        for (int i = 0; i < N; ++i) {
            std::vector<Mat> ml;
            for (int d = 0; d < nd; ++d) {
                const int ID = idx1d[size_t(i) * nd + d];  // 1-based
                if (ID < 1 || ID * deg > ncols[d]) throw std::runtime_error("wavelet_to_realspace: index out of range");
                cols[d].assign(meval[d] + size_t(ID - 1) * deg * npts[d], meval[d] + size_t(ID) * deg * npts[d]);
                ml.push_back(Mat{npts[d], deg, cols[d].data()});
            }
            std::vector<double> X(fw + size_t(i) * B, fw + size_t(i + 1) * B);
            std::vector<double> Y = kron_multd(ml, X);
            for (long p = 0; p < npt; ++p) fr[p] += Y[p];
        }
        return 0;
    })
}

int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
