/*
 * kmx.h -- C ABI of the MI355X (gfx950) operator-application library.
 *
 * Drop-in for ASGarD-MATLAB's (project-asgard/DG-SparseGrid) operator
 * application: apply_A.m's explicit branch, kron_multd.m / Kronmult/kronmult
 * {1..6}.m and Kronmult/gemm_vbatched.m.  A thin MATLAB MEX gateway, Python
 * ctypes or any C caller binds these symbols; no C++ or HIP types appear in
 * the signatures (streams are passed as void*, i.e. a hipStream_t or NULL for
 * the default stream).
 *
 * Conventions follow the reference: 1D coefficient matrices are dense,
 * column-major, FP64 (pde.terms{t}.terms_1D{d}.mat); the state vector f is
 * element-major with deg^d values per element, dim 1 slowest and dim d fastest
 * inside a block (krond order, apply_A.m:139,185); element 1D indices are
 * A_data.element_local_index_D (global_matrix.m:44-70); the pair list is the
 * connectivity cell array of connect_nD.m flattened to CSR, or "all pairs"
 * (apply_A.m:104-108,126-130).
 *
 * Every entry returns 0 (KMX_OK) or a negative KMX_E_* code; kmx_last_error()
 * gives the calling thread's message.  Nothing throws across this boundary.
 */
#ifndef KMX_H
#define KMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (replace the reference's error()/assert, kron_multd.m:7-12,24-28) */
#define KMX_OK 0
#define KMX_E_ARG (-1)         /* bad argument (null pointer, out-of-range value) */
#define KMX_E_SHAPE (-2)       /* inconsistent sizes (kron_multd.m:24-28, kronmult2.m:17-22) */
#define KMX_E_HIP (-3)         /* a HIP runtime call failed */
#define KMX_E_NOMEM (-4)       /* device allocation failed */
#define KMX_E_UNSUPPORTED (-5) /* valid request this build does not handle */

/* ---- precision of the contractions (x and y are always FP64 in memory) */
#define KMX_PREC_FP64 0       /* plain FP64, the reference's arithmetic (parity <= 1e-13 rel.) */
#define KMX_PREC_FP32 1       /* FP32 operands and accumulation, same pair-by-pair algorithm */
#define KMX_PREC_BF16 3       /* BF16 operands on the matrix cores (16x16x16 BF16 products),
                                 FP32 accumulation: the separable dimension sweep of a
                                 KMX_PLAN_LEVEL_BLOCKS | KMX_PLAN_FP32 plan, its combined
                                 operators rounded once to BF16, x rounded as it is read
                                 (else KMX_E_UNSUPPORTED).  ~2e-3 relative: a precision study */

/* ---- plan flags */
#define KMX_PLAN_FP32 1u    /* also keep FP32 copies of the coefficient blocks */
#define KMX_PLAN_ADJOINT 2u /* also keep the transposed pair list / blocks (kmx_apply_transpose) */
#define KMX_PLAN_LEVEL_BLOCKS 4u /* also detect the level-block structure and build the
                                    dimension sweep on it (kmx_plan_set_sweep); without it every
                                    apply runs pair by pair, in the reference's order of
                                    operations (apply_A.m:100-237) */

typedef struct kmx_plan kmx_plan;

/* Operator description.  All pointers are HOST pointers; kmx_plan_create copies
 * (and re-lays out) everything to the device, so the caller may free them after
 * the call.  Replaces the arguments apply_A(pde,opts,A_data,f,deg,...) reads
 * (apply_A.m:1): pde.terms{t}.terms_1D{d}.mat, A_data.element_local_index_D,
 * pde.connectivity / opts.use_connectivity, deg. */
typedef struct {
    int ndim;                  /* d, 1..6                                         */
    int deg;                   /* block edge (opts.deg), 1..8                     */
    int num_elem;              /* N = numel(A_data.element_global_row_index)      */
    int num_terms;             /* T terms applied (after the imex filter, :23-30) */
    const int* mat_rows;       /* [ndim] n_d = deg*2^L_d (square 1D matrices)     */
    const double* const* mats; /* [T*ndim] column-major n_d x n_d; mats[t*ndim+d] */
    const int* elem_idx1d;     /* [N*ndim] 1D index of element e in dim d, [e*ndim+d] */
    const int64_t* row_ptr;    /* [N+1] CSR offsets (0-based); NULL => all N^2 pairs */
    const int* cols;           /* [row_ptr[N]] connected element ids per row       */
    int index_base;            /* 1: elem_idx1d and cols are 1-based (MATLAB); 0: C */
    unsigned flags;            /* KMX_PLAN_* */
    /* [T*ndim] or NULL: identity[t*ndim+d] = 1 declares factor (t, d) the identity
     * -- a MASS pterm with g = 1 whose M\M (pterm_coeff_matrix.m:20) is I up to
     * roundoff.  The pair-by-pair path still applies the matrix as given; the
     * separable dimension sweep takes it as exactly I, for FP64 applies too.
     * kmx_plan_create rejects a declared factor with max |A - I| > 1e-8. */
    const unsigned char* identity;
} kmx_plan_desc;

/* Build a resident plan (coefficient blocks, element table, pair list stay on
 * the current HIP device until kmx_plan_destroy). */
int kmx_plan_create(const kmx_plan_desc* desc, kmx_plan** plan);

/* Replace one 1D matrix (term t, dim d; 0-based) -- the time-dependent
 * coefficient refresh of get_coeff_mats(...,TD=1) (get_coeff_mats.m:29-53).
 * Host pointer, column-major n_d x n_d.  Stream-ordered on `stream`. */
int kmx_plan_update_term(kmx_plan* plan, int term, int dim, const double* mat, void* stream);

/* Scale term t's contribution by alpha (default 1).  Used for per-term
 * coefficient scalings (batched instances, sensitivities).  The new scale is used
 * by every later apply of the plan, and by every batch holding the plan from its
 * next kmx_batch_apply on (which then refreshes the batch's copies, synchronously).
 * On a plan whose dimension sweep has the separable form it waits for the device
 * and rewrites the sweep's operator storage in place.  Not under graph capture. */
int kmx_plan_set_term_scale(kmx_plan* plan, int term, double alpha);

int kmx_plan_destroy(kmx_plan* plan);

/* Time the kernel variants (workgroup size 128/256/512, LDS staging of the
 * row's coefficient strips on/off) on this plan for FP64 and FP32 and keep the
 * fastest; plans of the same (d, deg, terms, pair density) created later in
 * the process start from that choice.  Synchronous on `stream`; not for use
 * under graph capture.  No counterpart in the reference (MATLAB picks nothing;
 * its flops_kron{k} method choice is the only cost model it has). */
int kmx_plan_autotune(kmx_plan* plan, void* stream);

/* The variant a plan currently uses for prec (KMX_PREC_FP64 / KMX_PREC_FP32): threads per
 * workgroup (negative: pair split over deg lanes) and the LDS staging of the row's
 * coefficient strips (0 none, 1 every dim, 2 every dim but dim 1). */
int kmx_plan_get_variant(const kmx_plan* plan, int prec, int* threads, int* lds_staged);

/* Dimension-sweep path of a KMX_PLAN_LEVEL_BLOCKS plan.  Elements sharing a
 * level multi-index form a dense cell tensor; when every pair of such level
 * groups is the tensor product of the 1D connections (connect_nD.m:63-106
 * builds them that way, an all-pairs list trivially; connected or not) and the
 * level tuples form a downward-closed set (sparse and full grids), the apply
 * runs as d sweeps, one GEMM per (source prefix, target suffix) of level
 * groups per dimension, on the matrix cores -- the same terms in another
 * order.  set: on = 1 uses it for prec (KMX_E_UNSUPPORTED when the plan has
 * none), 0 goes back; kmx_plan_autotune also times it.  BF16 runs on the
 * separable form only and always (on = 1 is accepted when it is available,
 * on = 0 refused).  get: whether the plan has it for prec, whether prec uses
 * it, and its flops per apply.  No counterpart in the reference. */
int kmx_plan_set_sweep(kmx_plan* plan, int prec, int on);
int kmx_plan_get_sweep(const kmx_plan* plan, int prec, int* available, int* active, double* flops);

/* The separable form of the sweep (every term acts in one dimension; its other
 * factors are the identity).  A factor counts as the identity when declared so
 * (kmx_plan_desc.identity) or within 1e-11 of I entrywise; FP64 applies may use
 * the form only when every undeclared such factor is I to 4 ulps (fp64 = 1),
 * FP32 applies always (the difference, identity_dev times the other factors, is
 * far below FP32 rounding; the reference's own MASS factors M\M deviate by
 * 1e-14 .. 3e-13).  Each dimension's terms are summed, scales included, into one
 * combined operator kept block-sparse: a deg x deg block is dropped when its
 * largest entry is at most drop_tol (1e-15) times the largest entry of every
 * term's matrix that has it -- in the wavelet basis the blocks between cells of
 * disjoint support are zero up to the transform's roundoff (1e-20 .. 1e-17 of
 * the largest entry), so what is dropped is roundoff.  No counterpart in the
 * reference. */
typedef struct {
    int available;          /* the plan has a dimension-sweep path             */
    int separable;          /* ... in the separable form                        */
    int fp64;               /* FP64 applies may use it                          */
    double identity_dev;    /* largest max |A - I| of the undeclared identity factors */
    double drop_tol;        /* block drop tolerance (separable form)            */
    double flops;           /* useful flops per apply (kept blocks)             */
    double mfma_flops;      /* matrix-core flops issued per apply               */
    int64_t kept_blocks;    /* kept / all deg x deg blocks of the combined operators */
    int64_t total_blocks;
    int64_t operator_values; /* values of operator storage an apply reads        */
    int64_t box_values;     /* values of one intermediate (every level box)     */
} kmx_sweep_info;
int kmx_plan_sweep_info(const kmx_plan* plan, kmx_sweep_info* info);

/* num_pairs = P, block = deg^d, dof = N*deg^d. */
int kmx_plan_info(const kmx_plan* plan, int64_t* num_pairs, int* block, int64_t* dof);

/* y = A x  (apply_A.m:100-237, explicit branch; y is overwritten, matching the
 * zero-initialised ftmp of apply_A.m:11).  x, y: DEVICE pointers, N*deg^d FP64.
 * prec: KMX_PREC_*.  Asynchronous on `stream`. */
int kmx_apply(kmx_plan* plan, const double* x, double* y, int prec, void* stream);

/* y = A^T x over the same pair list and terms (adjoint; no counterpart in the
 * reference, which has no gradient code).  Needs KMX_PLAN_ADJOINT, or a plan
 * whose dimension sweep has the separable form and is on for prec (the sweep
 * then runs with the transposed combined 1D matrices). */
int kmx_apply_transpose(kmx_plan* plan, const double* x, double* y, int prec, void* stream);

/* The assembled operator of apply_A.m's build_A branch (apply_A.m:150-155 and
 * 170-180, sparse triplets at 246-248; krond.m:7-17), on the device:
 * blocks[q * B*B + c * B + r] = sum_t scale_t krond(A_{t,1}(I_1, J_1), ...,
 * A_{t,d}(I_d, J_d)) at entry (r, c) of pair q's block, B = deg^d -- the values
 * A_s3 collects, in its order.  Pairs in the plan's order: rows in element order,
 * a row's pairs as given at plan creation (all-pairs plans: q = e * N + j); the
 * entry's global row / column are e * B + r and j * B + c (0-based).  Each entry
 * multiplies its factors dimension by dimension and adds the terms in order, as
 * krond and A = A + view do.  blocks: DEVICE pointer to num_pairs * B^2 doubles.
 * Asynchronous on `stream`. */
int kmx_plan_assemble(kmx_plan* plan, double* blocks, void* stream);

/* Host-pointer convenience for callers without device memory (a MEX gateway):
 * copies x in, applies, copies y out, synchronises. */
int kmx_apply_host(kmx_plan* plan, const double* x_host, double* y_host, int prec);

/* ---- batched instances: one persistent launch over several independent plans
 * (many PDE instances / adapted grids of uneven size, levels and deg).  xs[i],
 * ys[i] are device pointers for plan i; the pointer arrays are host arrays.
 * Any mix of (ndim, deg) is accepted.  No counterpart in the reference (its
 * batching, Kronmult/kronmult{k}_batch.m, batches the GEMMs of ONE apply). */
typedef struct kmx_batch kmx_batch;
int kmx_batch_create(kmx_plan* const* plans, int count, kmx_batch** batch);
int kmx_batch_apply(kmx_batch* batch, const double* const* xs, double* const* ys, int prec, void* stream);
int kmx_batch_destroy(kmx_batch* batch);

/* Merged dimension sweeps for a batch.  on = 1: for prec (KMX_PREC_FP64 /
 * KMX_PREC_FP32; KMX_PREC_BF16 names the FP32 merged sweep, whose tables BF16
 * applies use when every plan of the batch is a separable member of it), every
 * plan of the batch that has a dimension-sweep path
 * (kmx_plan_get_sweep; plans created with KMX_PLAN_LEVEL_BLOCKS) runs it, and
 * the sweeps of all such plans share one launch per phase (a 3D plan has five
 * phases: two down, one whole operator, two up); the other plans keep the
 * persistent pair launch.  A plan listed twice runs its sweep once per apply
 * for its first slot only; the others stay on the pair path.  FP64 merges only
 * the plans whose sweep allows FP64 (kmx_plan_sweep_info fp64).  Term scales and
 * matrices are the plans' current ones at every apply (kmx_plan_set_term_scale).
 * on = 0 goes back to the pair path.  Synchronous (allocates); not under graph capture.  KMX_E_UNSUPPORTED
 * when no plan has a sweep path.  get: plans in the merged sweep, whether
 * prec uses it, and its launches per apply.  No counterpart in the reference. */
int kmx_batch_set_sweep(kmx_batch* batch, int prec, int on);
int kmx_batch_get_sweep(const kmx_batch* batch, int prec, int* members, int* active, int* phases);

/* Times the batch's merged sweep for prec (on, kmx_batch_set_sweep) with 4, 8 and 16
 * load batches per wave before a tile is split over more waves (best of three
 * rounds of 10 applies on zero vectors each), keeps the fastest and returns it in
 * *wave_batches (may be null).  Rebuilds the sweep's tables (synchronous; not under
 * graph capture: graphs captured before must be captured again).  No counterpart in
 * the reference. */
int kmx_batch_autotune(kmx_batch* batch, int prec, int* wave_batches, void* stream);

/* nsteps explicit RK3 steps of every plan of the batch, y' = A y (the Kutta RK3
 * of time_advance.m:151-204 -- a21 = 1/2, a31 = -1, a32 = 2, b = 1/6, 2/3, 1/6 --
 * without source or boundary terms; contractions in prec, states and stage
 * derivatives in FP64), with the stage arithmetic fused into the
 * merged separable sweeps: per stage one launch of the dimension ops and one that
 * sums them into the stage derivative and writes the next stage's input (plus one
 * launch per call that reads the states in).  fs[i]: plan i's DEVICE state, N*deg^d
 * FP64, advanced in place; dts[i]: its step (host array).  Needs the batch's merged
 * sweep on for prec (kmx_batch_set_sweep) with every plan a separable member (else
 * KMX_E_UNSUPPORTED).  The first call allocates the stage workspace (3 x the
 * values of the members' boxes in FP64, plus one byte each); later calls allocate
 * nothing.  Asynchronous
 * on `stream`.  No counterpart in the reference (its time steps call apply_A three times). */
int kmx_batch_rk3(kmx_batch* batch, double* const* fs, const double* dts, int nsteps, int prec, void* stream);

/* One Arnoldi step of a restarted GMRES on the device, for the matrix-free
 * backward-Euler solve that replaces time_advance.m:328-349's direct solve
 * ((I - dt A) f1 = f0 + dt (s + bc)).  V: DEVICE, rows 0..k+1 of n FP64 values
 * (row j at V + j*n), rows 0..k orthonormal; u: DEVICE, the caller's product
 * A v_k.  Forms w = s0 v_k + s1 u (s0 = 1, s1 = -dt for I - dt A), orthogonalises
 * it against rows 0..k by classical Gram-Schmidt twice, writes h[0..k] (the
 * summed coefficients) and h[k+1] = |w| (DEVICE, k+2 values), and row k+1 of V =
 * w / |w| (zero when |w| = 0: breakdown).  work: DEVICE scratch of
 * KMX_ARNOLDI_WORK(n) doubles, owned by the caller.  Allocates nothing; three
 * launches of fixed shape for given (n, k), asynchronous on `stream` (graph
 * capturable).  |w| after the second pass is taken as sqrt(|w1|^2 - |h2|^2)
 * (w1: w after the first pass, h2: the second pass's coefficients).
 * 0 <= k <= KMX_ARNOLDI_MAX_K. */
#define KMX_ARNOLDI_MAX_K 127
#define KMX_ARNOLDI_WORK(n) ((n) + 32896)
int kmx_arnoldi_step(int64_t n, int k, double* V, const double* u, double s0, double s1, double* h, double* work,
                     void* stream);

/* Y = kron(A1,...,Ak) X for k = 1..KMX_KRON_MAX_FACTORS (kron_multd.m:30-42,
 * Kronmult/kronmult{k}.m; k > 6 as kron_multd.m:43-86's recursion).
 * A[i]: DEVICE pointer, column-major nrow[i] x ncol[i] (rectangular allowed);
 * X: DEVICE, prod(ncol) x nvec column-major; Y: DEVICE, prod(nrow) x nvec.
 * nrow/ncol/A are host arrays of length k. */
#define KMX_KRON_MAX_FACTORS 16
int kmx_kronmult(int k, const int* nrow, const int* ncol, const double* const* A, const double* X, int nvec,
                 double* Y, void* stream);

/* kmx_kronmult without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): no counterpart in
 * the reference.  Same arguments and result as kmx_kronmult; the chain's
 * intermediates live in a library-owned device array of KMX_KRON_SCRATCH_ELEMS
 * doubles (8 MiB, held per GPU until the process ends), so each intermediate
 * may hold at most KMX_KRON_SCRATCH_ELEMS / 2 values (else KMX_E_SHAPE).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (concurrent
 * streams or threads) race.  Asynchronous on `stream`; no synchronisation. */
#define KMX_KRON_SCRATCH_ELEMS (1 << 20)
int kmx_kronmult_noalloc(int k, const int* nrow, const int* ncol, const double* const* A, const double* X, int nvec,
                         double* Y, void* stream);

/* Variable-size batched GEMM (Kronmult/gemm_vbatched.m:1-38, magmablas-style):
 * C_i = alpha*op(A_i)*op(B_i) + beta*C_i with op from transA/transB ('N'/'T'),
 * m_i x k_i times k_i x n_i.  k is taken from k[] (the reference reads nlist at
 * gemm_vbatched.m:25; identical for its square use).  Size/ld arrays and the
 * pointer arrays are HOST arrays; the pointers address DEVICE memory. */
int kmx_gemm_vbatched(char transA, char transB, const int* m, const int* n, const int* k, double alpha,
                      const double* const* A, const int* lda, const double* const* B, const int* ldb, double beta,
                      double* const* C, const int* ldc, int count, void* stream);

/* One GEMM descriptor of kronmult{k}_batch's per-stage lists
 * (Kronmult/kronmult1_batch.m:45-61, kronmult2_batch.m:71-111, kronmult6_batch.m:86-125).
 * Operands are named by buffer + element offset so that stage s's output
 * aliases stage s+1's input (the reference's Clist holds copies instead). */
#define KMX_BUF_X 0      /* the input X                            */
#define KMX_BUF_Y 1      /* the output Y                           */
#define KMX_BUF_W0 2     /* workspace ping                         */
#define KMX_BUF_W1 3     /* workspace pong                         */
#define KMX_BUF_FACTOR 4 /* + i: factor A_{i+1} (0-based i)       */
typedef struct {
    int stage; /* 1..k: batch_list{stage} */
    char transA, transB;
    int m, n, k;
    int lda, ldb, ldc;
    double alpha, beta;
    int bufA, bufB, bufC;
    int64_t offA, offB, offC;
} kmx_gemm_desc;

/* Emit the descriptor lists kronmult{k}_batch would build for factors of the
 * given shapes and nvec vectors, in the reference's order (stage-major, then
 * emission order).  Pass out == NULL to query *count and *work_elems (elements
 * needed by each of W0 and W1).  Method choice per level follows
 * flops_kron{k} (Kronmult/flops_kron2.m:22-46). */
int kmx_kronmult_batch_plan(int k, const int* nrow, const int* ncol, int nvec, kmx_gemm_desc* out, int capacity,
                            int* count, int64_t* work_elems);

/* ---- wavelet space -> real space (wavelet_to_realspace.m:1-79).
 * f_r = sum over elements e of kron(Meval_1(:, block J_1(e)), ..., Meval_d(:, block J_d(e)))
 *       * f_w(block e), where Meval_d(:, block J) are the deg columns of 1D index J.
 * create: elem_idx1d N x d (as in kmx_plan_desc), meval[d] HOST column-major
 * npts[d] x ncols[d] (ncols = deg * cells); copied to the device.  apply: f_w
 * DEVICE N*deg^d, f_r DEVICE prod(npts), overwritten, Kronecker order (dim 1
 * slowest), asynchronous on `stream`. */
typedef struct kmx_realspace kmx_realspace;
int kmx_realspace_create(int ndim, int deg, int num_elem, const int* elem_idx1d, int index_base, const int* npts,
                         const int* ncols, const double* const* meval, kmx_realspace** out);
int kmx_realspace_apply(kmx_realspace* r, const double* f_w, double* f_r, void* stream);
int kmx_realspace_destroy(kmx_realspace* r);

/* ---- 1D operators on the device (get_coeff_mats.m:29-53 for one term and dimension).
 * The forward multiwavelet transform (OperatorTwoScale_wavelet2.m:1-144) as a dense
 * column-major deg*2^lev square on the HOST, from the two-scale filters H0, G0
 * (deg x deg, row-major: two_scale_rel_<deg>.mat). */
int kmx_fmwt_matrix(int deg, int lev, const double* H0, const double* G0, double* F);

/* The Gauss-Legendre points (lgwt.m, default_quad_number.m: nq = max(deg + 1, 10)
 * per cell) of the 2^lev cells of [xmin, xmax], cell by cell, where the pterm
 * functions are to be sampled: *nq always; x (HOST, 2^lev * nq values) if not NULL. */
int kmx_quad_points(int deg, int lev, double xmin, double xmax, int* nq, double* x);

#define KMX_PTERM_MASS 0
#define KMX_PTERM_DIV 1
#define KMX_PTERM_GRAD 2
/* One pterm (MASS.m / DIV.m / GRAD.m), its functions sampled by the caller.  All
 * arrays are HOST arrays.  LF and the internal boundary types are as the pterm
 * object holds them (GRAD.m already swaps D <-> N and negates LF). */
typedef struct {
    int type;                /* KMX_PTERM_*                                          */
    double LF;               /* flux coefficient (DIV / GRAD)                        */
    char ibcl, ibcr;         /* 'D', 'N' or 'P' (DIV / GRAD)                         */
    const double* g_quad;    /* [2^lev * nq] g(x) * dV(x) at the quadrature points    */
    const double* g_face;    /* [2^lev + 1] g(x) * dV(x) at the cell faces (DIV / GRAD) */
    const double* lhs_quad;  /* [2^lev * nq] LHS_mass_g(x) * moment_dV(x) at the points */
} kmx_pterm_desc;

/* The wavelet-space matrix of one dimension of one term: the product, in chain
 * order, of its pterms' M \ L (sd_term_coeff_matrix.m:1-27, pterm_coeff_matrix.m,
 * coeff_matrix.m:105-323), deg*2^lev square, column-major, into DEVICE memory mat.
 * Built in the cell basis (block-diagonal M, block-banded L, the chain's product)
 * and rotated once by the transform (H0, G0 as for kmx_fmwt_matrix).  Allocates
 * its work space; synchronous on `stream`. */
int kmx_sd_term_matrix(int deg, int lev, double xmin, double xmax, const double* H0, const double* G0, int npterms,
                       const kmx_pterm_desc* pterms, double* mat, void* stream);

const char* kmx_last_error(void);
int kmx_sync(void* stream);
/* Library version string. */
const char* kmx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* KMX_H */
