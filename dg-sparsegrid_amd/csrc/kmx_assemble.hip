// kmx_assemble.hip -- the assembled operator of apply_A.m's build_A branch
// (apply_A.m:170-180 dense A, 150-155 and 246-248 the sparse triplets), on the
// device: for every connected pair (row element e, column element j) in the plan's
// pair order (rows in element order, a row's pairs in connectivity order), the
// deg^d x deg^d block sum_t scale_t krond(A_{t,1}(I_1, J_1), ..., A_{t,d}(I_d, J_d))
// (krond.m:7-17: dim 1 slowest), column-major -- the values A_s3 collects
// (view(:)), in its order.  Row / column of block entry (r, c) of pair (e, j):
// e * B + r and j * B + c (0-based, B = deg^d), as A_s2 / A_s1 hold them.
// Each entry multiplies its d factors left to right and adds the terms in order,
// without fused multiply-adds, as krond's successive kron products and
// apply_A.m's `A = A + view` do.
#include "kmx_internal.h"

#include <cstring>

namespace kmx {
namespace {

constexpr int AS_THREADS = 256;

// one workgroup per row element; its pairs x B^2 entries, 256 at a time
__global__ void __launch_bounds__(AS_THREADS) k_assemble(KArgs<double> a, int D, int deg, int B, double* out) {
    const int e = blockIdx.x;
    const long long q0 = a.row_ptr ? a.row_ptr[e] : (long long)e * a.N;
    const long long nq = a.row_ptr ? a.row_ptr[e + 1] - q0 : a.N;
    const long long BB = (long long)B * B;
    const int* ie = a.idx1d + (long long)e * D;
    for (long long w = threadIdx.x; w < nq * BB; w += AS_THREADS) {
        const long long q = w / BB;
        const int ent = int(w - q * BB);
        const int c = ent / B, r = ent - c * B;  // column-major inside the block
        const int j = a.row_ptr ? a.cols[q0 + q] : int(q);
        const int* ij = a.idx1d + (long long)j * D;
        // digits of r and c, dim 1 slowest
        int rd[KMX_MAX_DIM], cd[KMX_MAX_DIM];
        int rr = r, cc = c;
        for (int k = D - 1; k >= 0; --k) {
            rd[k] = rr % deg;
            rr /= deg;
            cd[k] = cc % deg;
            cc /= deg;
        }
        double s = 0.0;
        for (int t = 0; t < a.nterms; ++t) {
            double p = 1.0;
            for (int k = 0; k < D; ++k) {
                // block (I, J) of the (t, k) matrix, row-major [a][b]
                const double v = a.blk[a.off[t * D + k] + ((long long)ie[k] * a.nb[k] + ij[k]) * deg * deg +
                                       rd[k] * deg + cd[k]];
                p = k == 0 ? v : __dmul_rn(p, v);
            }
            s = __dadd_rn(s, a.scale[t] == 1.0 ? p : __dmul_rn(a.scale[t], p));
        }
        out[(q0 + q) * BB + ent] = s;
    }
}

}  // namespace
}  // namespace kmx

using namespace kmx;

extern "C" int kmx_plan_assemble(kmx_plan* p, double* blocks, void* stream) {
    if (!p || (p->P > 0 && !blocks)) {
        set_error("kmx_plan_assemble: null argument");
        return KMX_E_ARG;
    }
    if (p->N == 0 || p->P == 0) return KMX_OK;
    KArgs<double> a;
    std::memset(&a, 0, sizeof(a));
    a.blk = p->d_blk64;
    for (int i = 0; i < p->T * p->D; ++i) a.off[i] = p->off[i];
    for (int k = 0; k < p->D; ++k) a.nb[k] = p->nb[k];
    for (int t = 0; t < p->T; ++t) a.scale[t] = p->scale[t];
    a.idx1d = p->d_idx1d;
    a.row_ptr = p->all_pairs ? nullptr : p->d_row_ptr;
    a.cols = p->all_pairs ? nullptr : p->d_cols;
    a.N = p->N;
    a.nterms = p->T;
    const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_assemble, dim3(unsigned(p->N)), dim3(AS_THREADS), 0, s, a, p->D, p->deg, p->B, blocks);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("kmx_plan_assemble: ") + hipGetErrorString(e));
        return KMX_E_HIP;
    }
    return KMX_OK;
}
