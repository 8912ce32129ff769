// kmx_apply_f32.hip -- FP32 instantiations of the operator-apply kernels
// (the mixed-precision path, kmx_kernels.h).
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

}  // namespace kmx
