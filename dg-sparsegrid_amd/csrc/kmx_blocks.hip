// kmx_blocks.hip -- the level-block structure of a plan, which the dimension
// sweep (kmx_sweep.hip) is built on.
//
// apply_A.m:100-237 sums, over connected element pairs (i, j) and terms t,
// (A_{t,1}[I_1,J_1] x ... x A_{t,d}[I_d,J_d]) x_j into y_i.  Group the
// elements by their level multi-index (l_1..l_d): the elements of one group
// are cells of a d-dimensional tensor (2^max(l_k-1,0) cells along dim k, the
// 1D indices of a level being contiguous, lev_cell_to_1D_index.m:6).  When
// the pairs between a row group L and a column group L' are exactly the cells
// whose per-dimension 1D indices are connected (a tensor product of 1D masks,
// which is how connect_nD.m:63-106 builds them from connect_1D and what an
// all-pairs list gives trivially), the sum over those pairs is a sum of mode
// products of the masked 1D sub-matrices on dense cell tensors, and the sum over
// all block pairs factors into the dimension sweep.  kmx_blocks_build verifies
// that structure (ok), records whether unconnected block pairs also have an
// empty mask product (full_product), and uploads the masked 1D matrices the
// general sweep reads; any failure leaves the plan on the pair-by-pair kernels.
#include "kmx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <unordered_map>

namespace kmx {
namespace {

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

}  // namespace
}  // namespace kmx

using namespace kmx;

// --------------------------------------------------------------- structure
static void build_impl(kmx_plan* p, const std::vector<int>& idx, const std::vector<long long>& rp,
                       const std::vector<int>& cols, const double* const* mats) {
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
    // masked matrices
    std::vector<double> m64(size_t(p->mat_elems));
    for (int t = 0; t < T; ++t)
        for (int k = 0; k < D; ++k) masked_matrix(mats[t * D + k], p->n[k], deg, bs->mask[k], m64.data() + p->off[t * D + k]);
    bs->d_mat64 = up(m64.data(), m64.size());
    if (p->flags & KMX_PLAN_FP32) {
        std::vector<float> m32(m64.begin(), m64.end());
        bs->d_mat32 = up(m32.data(), m32.size());
    }
    bs->ok = true;
}

// Detect the level-block structure; any failure (structure or resources)
// leaves the plan on the pair-by-pair path with the reason in bs->why.
void kmx_blocks_build(kmx_plan* p, const std::vector<int>& idx, const std::vector<long long>& rp,
                      const std::vector<int>& cols, const double* const* mats) {
    p->bs = new kmx_blockset();
    try {
        build_impl(p, idx, rp, cols, mats);
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
    dfree(bs->d_mat64);
    dfree(bs->d_mat32);
    delete bs;
    p->bs = nullptr;
}
