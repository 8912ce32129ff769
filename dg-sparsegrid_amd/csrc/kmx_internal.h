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
};

// x / y device pointers of the instances of a batch, by slot (a kernel argument).
#define KMX_BATCH_MAX 128
struct BatchIO {
    const double* x[KMX_BATCH_MAX];
    double* y[KMX_BATCH_MAX];
};

}  // namespace kmx

// Launch variant of the pair-by-pair kernels: workgroup size (0 = default
// 256) and LDS staging of the row's coefficient strips (-1 = when they fit).
struct kmx_variant {
    int nt = 0;
    int lds = -1;
    int kind = 0;          // 0: one lane per pair, 1: pair split over deg lanes
    int sweep = 0;         // 1: the dimension-sweep path (when the plan has one)
};

// Level-block structure of a plan: elements grouped by their level
// multi-index (one dense cell tensor per group), per-dimension 1D connection
// masks, and the masked 1D matrices.  ok is set only when every connected
// pair of blocks holds exactly the tensor product of the per-dimension masks,
// so a factorized (dimension-sweep) sum equals the pair-by-pair sum term for term.
struct kmx_blockset {
    bool ok = false;
    std::string why;
    int nblk = 0;
    std::vector<int> binfo;   // per block: D log2 cell counts, D first 1D indices, cell-table offset
    std::vector<int> elems;   // per block: element of each cell (row-major, dim 1 slowest), -1 if absent
    std::vector<int> eblk, ecell;
    std::vector<int> blev;    // per block: D levels
    bool full_product = false;  // also every unconnected block pair has an empty mask product
    std::vector<std::vector<unsigned char>> mask;  // per dim: nb x nb, [I * nb + J]
    double* d_mat64 = nullptr;  // masked 1D matrices, column-major, offsets off[t*D+k]
    float* d_mat32 = nullptr;
};

struct kmx_sweep;  // dimension-sweep path (kmx_sweep.hip)

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

    // adjoint
    long long* d_col_ptr = nullptr;
    int* d_rows = nullptr;
    int* d_col_order = nullptr;
    double* d_blkT64 = nullptr;
    float* d_blkT32 = nullptr;

    // host copies for kmx_plan_update_term and the sweep set-up
    std::vector<int> h_idx1d;  // N x D, 0-based
    std::vector<double> h_blk64;
    std::vector<double> h_blkT64;
    double* d_x_host_stage = nullptr;  // kmx_apply_host staging
    double* d_y_host_stage = nullptr;

    // factors the caller declared the identity (kmx_plan_desc.identity; empty: none)
    std::vector<unsigned char> declared_identity;
    // bumped by every change of a term matrix or scale (batches refresh their copies)
    unsigned long long version = 0;

    kmx_blockset* bs = nullptr;  // level-block structure (null unless KMX_PLAN_LEVEL_BLOCKS)
    kmx_sweep* sw = nullptr;     // dimension-sweep path (null when the structure does not allow it)
};

// Launchers (kmx_apply.hip)
int kmx_launch_apply(kmx_plan* p, bool transpose, const double* x, double* y, int prec, hipStream_t s);
void kmx_plan_apply_cached_variants(kmx_plan* p);

// Level-block structure (kmx_blocks.hip); rp / cols: the 0-based CSR pair list
// (ignored for all-pairs plans)
void kmx_blocks_build(kmx_plan* p, const std::vector<int>& idx, const std::vector<long long>& rp,
                      const std::vector<int>& cols, const double* const* mats);
void kmx_blocks_update_term(kmx_plan* p, int term, int dim, const double* mat, hipStream_t s);
void kmx_blocks_free(kmx_plan* p);

// Dimension-sweep path (kmx_sweep.hip).  slot: 0 FP64, 1 FP32, 2 BF16 (separable
// sweeps only: BF16 values of the combined operators on the FP32 arena, BF16
// matrix-core products with FP32 accumulation).
void kmx_sweep_build(kmx_plan* p);
int kmx_sweep_apply(kmx_plan* p, int slot, const double* x, double* y, hipStream_t s);
void kmx_sweep_free(kmx_plan* p);
double kmx_sweep_flops(const kmx_plan* p);
// whether FP64 applies may use the plan's sweep (a separable sweep whose undeclared
// identity factors are exact to 4 ulps, or a general sweep)
bool kmx_sweep_fp64_ok(const kmx_plan* p);
// a separable sweep with BF16 values (BF16 applies, forward and, with its tables, adjoint)
bool kmx_sweep_bf16_ok(const kmx_plan* p);
void kmx_sweep_get_info(const kmx_plan* p, kmx_sweep_info* out);
// after kmx_plan_update_term(term, dim) (the FP64 block copy already updated)
void kmx_sweep_term_updated(kmx_plan* p, int term, int dim);
// after kmx_plan_set_term_scale: the plan's own launch tables take the new scales
int kmx_sweep_scales_updated(kmx_plan* p, int term);
// The sweeps of several plans merged into one launch per phase (kmx_sweep.hip):
// a plan's ops form launch groups (one op each for a general sweep, in order; all
// ops together for a separable one), phase k runs group k - (phases - groups of
// the plan) of every plan, so each plan's groups keep their order and every
// plan's last group falls in the last phase; separable plans add one launch
// before (box copies of x) and one after (their slots summed into y).  members:
// distinct plans, each with a sweep path; io slot q is members[q].  Term scales
// are taken at build time; wave_batches <= 0: the batch default (four load
// batches per wave); transpose: the separable members' transposed combined
// matrices (the adjoint apply).  Build throws std::exception on failure.
struct kmx_sweep_batch;
kmx_sweep_batch* kmx_sweep_batch_build(const std::vector<kmx_plan*>& members, int wave_batches = 0,
                                       bool transpose = false);
// the adjoint apply on a separable sweep (transposed combined matrices, same launches)
bool kmx_sweep_has_transpose(const kmx_plan* p);
int kmx_sweep_apply_transpose(kmx_plan* p, int slot, const double* x, double* y, hipStream_t s);
int kmx_sweep_batch_apply(const kmx_sweep_batch* sb, int slot, const kmx::BatchIO& io, hipStream_t s);
// nsteps fused RK3 steps of a batch whose members are all separable (fs / dts by
// member); allocates the stage workspace on its first call only
int kmx_sweep_batch_rk3(kmx_sweep_batch* sb, int slot, double* const* fs, const double* dts, int nsteps,
                        hipStream_t s);
void kmx_sweep_batch_free(kmx_sweep_batch* sb);
int kmx_sweep_batch_phases(const kmx_sweep_batch* sb);
bool kmx_sweep_batch_has16(const kmx_sweep_batch* sb);  // every op separable with BF16 values
