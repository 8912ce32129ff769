// kmx_apply_f32.hip -- FP32 instantiations of the operator-apply kernels
// (the mixed-precision path, kmx_kernels.h) and the grouped FP32 kernel.
#include "kmx_kernels.h"

namespace kmx {

LaunchFn pick_f32(int D, int deg, int nt, int lds) { return pick_impl<float>(D, deg, nt, lds); }
LaunchFn pick_split_f32(int D, int deg, int nt, bool lds) { return pick_split_impl<float>(D, deg, nt, lds); }

hipError_t launch_generic_f32(const KArgs<float>& a, int D, int deg, int nrows, size_t smem, hipStream_t s) {
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_apply_generic<float>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_apply_generic<float>, dim3(nrows), dim3(256), smem, s, a, D, deg);
    return hipGetLastError();
}

hipError_t launch_batch_f32(const KArgs<float>* args, const int2* items, int nitems, const int* shape,
                            const BatchIO& io, int grid, size_t smem, hipStream_t s) {
    hipLaunchKernelGGL(k_apply_batch<float>, dim3(grid), dim3(256), smem, s, args, items, nitems, shape, io);
    return hipGetLastError();
}

int batch_occupancy_f32(size_t smem) {
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k_apply_batch<float>), 256,
                                                     smem) != hipSuccess)
        occ = 1;
    return occ;
}

// -------------------------------------------------------------- grouped FP32
// Pairs of a row are pre-sorted so that pairs sharing (J_1..J_{d-1}) are
// adjacent ("prefix groups").  A lane takes one group: it contracts dim d of
// every member's x_j per term and sums them (W_t), then applies dims d-1..1
// once per group and term.  Same sum as the pair-by-pair chain, fewer flops.
template <int D, int DEG, int NT, int TMAX>
__global__ void __launch_bounds__(NT) k_apply_grouped(KArgs<float> a) {
    constexpr int B = Pow<DEG, D>::v;
    constexpr int NW = NT / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    const int row = a.row_order[blockIdx.x];
    const int tid = threadIdx.x;
    RowCtx<float, D, DEG, true> c;
    c.a = &a;
#pragma unroll
    for (int k = 0; k < D; ++k) c.I[k] = a.idx1d[(long long)row * D + k];
    float* strip = reinterpret_cast<float*>(smem);
    {
        int so = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            c.soff[k] = so;
            so += a.nb[k] * DEG * DEG;
        }
        c.sterm = so;
        for (int t = 0; t < a.nterms; ++t) {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float* src = a.blk + a.off[t * D + k] + (long long)c.I[k] * a.nb[k] * DEG * DEG;
                float* dst = strip + t * so + c.soff[k];
                const int len = a.nb[k] * DEG * DEG;
                for (int e = tid; e < len; e += NT) dst[e] = src[e];
            }
        }
        __syncthreads();
        c.strip = strip;
    }
    float acc[B];
#pragma unroll
    for (int e = 0; e < B; ++e) acc[e] = 0.f;

    const long long gb = a.grp_row_ptr[row], ge = a.grp_row_ptr[row + 1] - 1;  // groups gb..ge-1
    for (long long g = gb + tid; g < ge; g += NT) {
        const int p0 = a.grp_ptr[g], p1 = a.grp_ptr[g + 1];
        int J[D];
        {
            const int j0 = a.cols[p0];
#pragma unroll
            for (int k = 0; k < D - 1; ++k) J[k] = a.idx1d[(long long)j0 * D + k];
        }
        float W[TMAX][B];
#pragma unroll
        for (int t = 0; t < TMAX; ++t)
#pragma unroll
            for (int e = 0; e < B; ++e) W[t][e] = 0.f;
        for (int p = p0; p < p1; ++p) {
            const int j = a.cols[p];
            const int Jd = a.idx1d[(long long)j * D + (D - 1)];
            float xv[B];
            const double* xj = a.x + (long long)j * B;
#pragma unroll
            for (int e = 0; e < B; ++e) xv[e] = float(xj[e]);
#pragma unroll
            for (int t = 0; t < TMAX; ++t) {
                if (t < a.nterms) {
                    float m[DEG * DEG];
                    c.load_block(t, D - 1, Jd, m);
                    float v[B];
#pragma unroll
                    for (int e = 0; e < B; ++e) v[e] = xv[e];
                    contract_dim<float, D, DEG, D - 1>(v, m);
#pragma unroll
                    for (int e = 0; e < B; ++e) W[t][e] += v[e];
                }
            }
        }
#pragma unroll
        for (int t = 0; t < TMAX; ++t) {
            if (t < a.nterms) {
                Chain<float, D, DEG, true, D - 2>::run(W[t], c, t, J);
                const float s = a.scale[t];
#pragma unroll
                for (int e = 0; e < B; ++e) acc[e] = madd(s, W[t][e], acc[e]);
            }
        }
    }
    __syncthreads();
    float* red = reinterpret_cast<float*>(smem);
    const int lane = tid & 63, w = tid >> 6;
    wave_reduce_store<float, B>(acc, red + w * B, lane);
    __syncthreads();
    for (int e = tid; e < B; e += NT) {
        float s = red[e];
#pragma unroll
        for (int ww = 1; ww < NW; ++ww) s += red[ww * B + e];
        a.y[(long long)row * B + e] = double(s);
    }
}

// grouped FP32 kernels are built for the configurations that pay off
template <int D, int DEG, int TMAX, int NT = 256>
hipError_t launch_grouped(const void* args, int nrows, size_t smem, hipStream_t s) {
    auto kern = k_apply_grouped<D, DEG, NT, TMAX>;
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(nrows), dim3(NT), smem, s, *reinterpret_cast<const KArgs<float>*>(args));
    return hipGetLastError();
}

LaunchFn pick_grouped(int D, int deg, int T, int nt) {
    if (T > 4) return nullptr;
    if (nt == 512) {
        if (D == 3 && deg == 3) return T <= 3 ? &launch_grouped<3, 3, 3, 512> : &launch_grouped<3, 3, 4, 512>;
        if (D == 3 && deg == 2) return &launch_grouped<3, 2, 4, 512>;
        if (D == 2 && deg == 3) return &launch_grouped<2, 3, 4, 512>;
    }
    if (D == 2 && deg == 2) return &launch_grouped<2, 2, 4>;
    if (D == 2 && deg == 3) return &launch_grouped<2, 3, 4>;
    if (D == 2 && deg == 4) return &launch_grouped<2, 4, 4>;
    if (D == 3 && deg == 2) return &launch_grouped<3, 2, 4>;
    if (D == 3 && deg == 3) return T <= 3 ? &launch_grouped<3, 3, 3> : &launch_grouped<3, 3, 4>;
    return nullptr;
}

}  // namespace kmx
