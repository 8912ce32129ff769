// kmx_internal.h -- device-side plan layout shared by the gfx950 kernels and
// the host code of libkmx.  Not part of the public ABI (include/kmx.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/kmx.h"

#define KMX_MAX_DIM 6
#define KMX_MAX_TERMS 16
#define KMX_MAX_DEG 8

namespace kmx {

void set_error(const std::string& msg);

// Coefficient blocks of one precision.  For term t and dim k the n_k x n_k
// 1D matrix is stored block-row-major: block (I, J) (0-based 1D indices) at
// off[t*D+k] + (I*nb_k + J)*deg^2, each deg x deg block row-major [a][b] with
// a the row inside I and b the column inside J.  For a fixed row element the
// strip of blocks (I, :) is therefore contiguous.
template <typename T>
struct KArgs {
    const T* blk;
    long long off[KMX_MAX_TERMS * KMX_MAX_DIM];
    int nb[KMX_MAX_DIM];
    T scale[KMX_MAX_TERMS];
    const int* idx1d;      // N x D, 0-based
    const long long* row_ptr;  // N+1, null => all pairs
    const int* cols;       // 0-based element ids
    const int* row_order;  // work order of rows (heaviest first)
    const double* x;
    double* y;
    int N;
    int nterms;
    int strip_elems;       // sum over (t,k) of nb_k*deg^2 (LDS strip size in T elements)
    // grouped variant only
    const int* grp_ptr;    // per row: [row_ptr..] groups (CSR over pairs sorted by prefix)
    const long long* grp_row_ptr;  // N+1 offsets into grp_ptr
};

}  // namespace kmx

// Launch variant of the pair-by-pair kernels: workgroup size (0 = default
// 256) and LDS staging of the row's coefficient strips (-1 = when they fit).
struct kmx_variant {
    int nt = 0;
    int lds = -1;
    int kind = 0;  // 0: one lane per pair, 1: pair split over deg lanes
};

struct kmx_plan {
    int D = 0, deg = 0, N = 0, T = 0, B = 0;
    long long P = 0;
    bool all_pairs = false;
    unsigned flags = 0;
    int n[KMX_MAX_DIM] = {0};
    int nb[KMX_MAX_DIM] = {0};
    long long off[KMX_MAX_TERMS * KMX_MAX_DIM] = {0};
    long long mat_elems = 0;  // total elements of all blocked matrices
    double scale[KMX_MAX_TERMS];
    int device = 0;
    kmx_variant var[2];  // [0] FP64, [1] FP32 (pair-by-pair kernels)

    // forward operator
    int* d_idx1d = nullptr;
    long long* d_row_ptr = nullptr;
    int* d_cols = nullptr;
    int* d_row_order = nullptr;
    double* d_blk64 = nullptr;
    float* d_blk32 = nullptr;
    // grouped (prefix) layout of the pair list, FP32 grouped variant
    int* d_gcols = nullptr;           // pairs re-sorted by (J_1..J_{d-1}, J_d) per row
    int* d_grp_ptr = nullptr;         // group start offsets (absolute into d_gcols), per row + sentinel
    long long* d_grp_row_ptr = nullptr;  // N+1 offsets into d_grp_ptr
    long long G = 0;                  // total groups

    // adjoint
    long long* d_col_ptr = nullptr;
    int* d_rows = nullptr;
    int* d_col_order = nullptr;
    double* d_blkT64 = nullptr;
    float* d_blkT32 = nullptr;

    // host copies for kmx_plan_update_term
    std::vector<double> h_blk64;
    std::vector<double> h_blkT64;
    double* d_x_host_stage = nullptr;  // kmx_apply_host staging
    double* d_y_host_stage = nullptr;
};

// Launchers (kmx_apply.hip)
int kmx_launch_apply(kmx_plan* p, bool transpose, const double* x, double* y, int prec, hipStream_t s);
void kmx_plan_apply_cached_variants(kmx_plan* p);
