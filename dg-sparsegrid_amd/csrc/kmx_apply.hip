// kmx_apply.hip -- host-side launch logic of the operator apply: variant
// choice (autotuned per plan and precision), argument packing, batches.
// Device code lives in kmx_kernels.h (instantiated in kmx_apply_f64.hip and
// kmx_apply_f32.hip).
#include "kmx_kernels.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

using namespace kmx;

static size_t lds_limit() {
    static size_t lim = [] {
        const char* e = getenv("KMX_LDS_LIMIT");
        return e ? size_t(atol(e)) : size_t(160 * 1024);
    }();
    return lim;
}

template <typename T>
static void fill_args(KArgs<T>& a, const kmx_plan* p, bool transpose, const double* x, double* y) {
    std::memset(&a, 0, sizeof(a));
    if (std::is_same<T, double>::value)
        a.blk = reinterpret_cast<const T*>(transpose ? p->d_blkT64 : p->d_blk64);
    else
        a.blk = reinterpret_cast<const T*>(transpose ? p->d_blkT32 : p->d_blk32);
    for (int i = 0; i < p->T * p->D; ++i) a.off[i] = p->off[i];
    for (int k = 0; k < p->D; ++k) a.nb[k] = p->nb[k];
    for (int t = 0; t < p->T; ++t) a.scale[t] = T(p->scale[t]);
    a.idx1d = p->d_idx1d;
    a.row_ptr = transpose ? p->d_col_ptr : p->d_row_ptr;
    a.cols = transpose ? p->d_rows : p->d_cols;
    a.row_order = transpose ? p->d_col_order : p->d_row_order;
    a.x = x;
    a.y = y;
    a.N = p->N;
    a.nterms = p->T;
    int se = 0;
    for (int k = 0; k < p->D; ++k) se += p->nb[k] * p->deg * p->deg;
    a.strip_elems = se * p->T;
}

static int prec_slot(int prec) { return prec == KMX_PREC_FP64 ? 0 : 1; }

// LDS bytes of a row's coefficient strips: every term and dim (mode 1) or
// every dim but the first (mode 2, whose dim-1 blocks come from L2).
static size_t strip_bytes(const kmx_plan* p, size_t elem, int mode = 1) {
    size_t se = 0;
    for (int k = (mode == 2 ? 1 : 0); k < p->D; ++k) se += size_t(p->nb[k]) * p->deg * p->deg;
    return se * p->T * elem;
}

static bool strip_fits(const kmx_plan* p, size_t elem, int mode = 1) {
    return strip_bytes(p, elem, mode) <= lds_limit();
}

// One launch of the pair-by-pair kernel family with an explicit variant.
static int launch_variant(kmx_plan* p, bool transpose, const double* x, double* y, int prec, kmx_variant v,
                          hipStream_t s) {
    const int D = p->D, deg = p->deg, B = p->B;
    hipError_t err = hipSuccess;
    if (prec == KMX_PREC_FP64 || prec == KMX_PREC_FP32) {
        const bool f64 = prec == KMX_PREC_FP64;
        const size_t elem = f64 ? sizeof(double) : sizeof(float);
        if (!f64 && !(transpose ? p->d_blkT32 : p->d_blk32)) {
            set_error("kmx_apply: plan has no FP32 blocks (create it with KMX_PLAN_FP32)");
            return KMX_E_ARG;
        }
        // LDS mode: 0 none, 1 every strip, 2 every strip but dim 1's (-1: 1 when it fits)
        int mode = v.lds < 0 ? (strip_fits(p, elem) ? 1 : 0) : v.lds;
        if (mode && !strip_fits(p, elem, mode)) mode = 0;
        if (mode == 2 && (v.kind == 1 || D < 2)) mode = 1;
        if (mode == 1 && !strip_fits(p, elem, 1)) mode = 0;
        const bool lds = mode != 0;
        const int nt = v.nt > 0 ? v.nt : 256;
        LaunchFn f = nullptr;
        if (v.kind == 1) f = f64 ? pick_split_f64(D, deg, nt, lds) : pick_split_f32(D, deg, nt, lds);
        if (!f) f = f64 ? pick_f64(D, deg, nt, mode) : pick_f32(D, deg, nt, mode);
        if (!f) f = f64 ? pick_split_f64(D, deg, nt == 128 ? 256 : nt, lds) : pick_split_f32(D, deg, nt == 128 ? 256 : nt, lds);
        if (f) {
            const size_t strip = strip_bytes(p, elem, mode);
            const size_t smem = std::max(lds ? strip : size_t(0), size_t(nt / 64) * B * elem);
            if (f64) {
                KArgs<double> a;
                fill_args(a, p, transpose, x, y);
                err = f(&a, p->N, smem, s);
            } else {
                KArgs<float> a;
                fill_args(a, p, transpose, x, y);
                err = f(&a, p->N, smem, s);
            }
        } else {
            const size_t smem = size_t(3 * B) * elem;
            if (smem > 160 * 1024) {
                set_error("kmx_apply: block too large for the generic kernel");
                return KMX_E_UNSUPPORTED;
            }
            if (f64) {
                KArgs<double> a;
                fill_args(a, p, transpose, x, y);
                err = launch_generic_f64(a, D, deg, p->N, smem, s);
            } else {
                KArgs<float> a;
                fill_args(a, p, transpose, x, y);
                err = launch_generic_f32(a, D, deg, p->N, smem, s);
            }
        }
    } else {
        set_error("kmx_apply: unknown precision");
        return KMX_E_ARG;
    }
    if (err != hipSuccess) {
        set_error(std::string("apply launch: ") + hipGetErrorString(err));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

int kmx_launch_apply(kmx_plan* p, bool transpose, const double* x, double* y, int prec, hipStream_t s) {
    if (p->N == 0) return KMX_OK;
    if (prec == KMX_PREC_BF16) {  // the separable sweep with BF16 operands, whatever kmx_plan_set_sweep says
        if (!kmx_sweep_bf16_ok(p) || (transpose && !kmx_sweep_has_transpose(p))) {
            set_error("kmx_apply: BF16 runs on the separable dimension sweep only (a KMX_PLAN_LEVEL_BLOCKS | "
                      "KMX_PLAN_FP32 plan whose terms each act in one dimension)");
            return KMX_E_UNSUPPORTED;
        }
        return transpose ? kmx_sweep_apply_transpose(p, 2, x, y, s) : kmx_sweep_apply(p, 2, x, y, s);
    }
    if (prec != KMX_PREC_FP64 && prec != KMX_PREC_FP32) {
        set_error("kmx_apply: unknown precision");
        return KMX_E_ARG;
    }
    if (p->sw && p->var[prec_slot(prec)].sweep) {
        if (!transpose) return kmx_sweep_apply(p, prec_slot(prec), x, y, s);
        if (kmx_sweep_has_transpose(p)) return kmx_sweep_apply_transpose(p, prec_slot(prec), x, y, s);
    }
    kmx_variant v = p->var[prec_slot(prec)];
    return launch_variant(p, transpose, x, y, prec, v, s);
}

// ---------------------------------------------------------------- autotuner
// Per (d, deg, terms, precision, pair-density bucket): the workgroup size and
// LDS staging that ran fastest on a real plan of that shape.  Plans created
// later with the same key start from the cached choice.
namespace {
struct TuneKey {
    int D, deg, T, prec, dens;
    bool operator<(const TuneKey& o) const {
        return std::tie(D, deg, T, prec, dens) < std::tie(o.D, o.deg, o.T, o.prec, o.dens);
    }
};
std::map<TuneKey, kmx_variant>& tune_cache() {
    static std::map<TuneKey, kmx_variant> c;
    return c;
}
std::mutex& tune_mutex() {
    static std::mutex m;
    return m;
}
TuneKey key_of(const kmx_plan* p, int prec) {
    const double per_row = p->N ? double(p->P) / p->N : 0.0;
    int dens = 0;
    while ((1 << dens) < per_row && dens < 30) ++dens;
    return TuneKey{p->D, p->deg, p->T, prec, dens};
}
}  // namespace

void kmx_plan_apply_cached_variants(kmx_plan* p) {
    std::lock_guard<std::mutex> g(tune_mutex());
    for (int prec : {KMX_PREC_FP64, KMX_PREC_FP32}) {
        auto it = tune_cache().find(key_of(p, prec));
        if (it != tune_cache().end()) {
            kmx_variant v = it->second;
            v.sweep = v.sweep && p->sw && (prec == KMX_PREC_FP32 || kmx_sweep_fp64_ok(p));
            p->var[prec_slot(prec)] = v;
        }
    }
}

namespace {
// Average time of one call of fn on stream s: best of three rounds, each
// round at least ~1 ms of calls (at least 8) so that short kernels are timed
// over many launches.
template <typename F>
int time_calls(F fn, hipStream_t s, hipEvent_t e0, hipEvent_t e1, float& ms) {
    int rc = KMX_OK;
    for (int w = 0; w < 2 && rc == KMX_OK; ++w) rc = fn();
    ms = 1e30f;
    int reps = 8;
    for (int round = 0; round < 3 && rc == KMX_OK; ++round) {
        hipEventRecord(e0, s);
        for (int r = 0; r < reps && rc == KMX_OK; ++r) rc = fn();
        hipEventRecord(e1, s);
        hipEventSynchronize(e1);
        float t = 0.f;
        hipEventElapsedTime(&t, e0, e1);
        ms = std::min(ms, t / reps);
        if (round == 0 && t < 1.0f) reps = std::min(200, std::max(8, int(reps * 1.0f / std::max(t, 1e-3f)) + 1));
    }
    return rc;
}
}  // namespace

extern "C" int kmx_plan_autotune(kmx_plan* p, void* stream) {
    if (!p) {
        set_error("kmx_plan_autotune: null plan");
        return KMX_E_ARG;
    }
    if (p->N == 0) return KMX_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t n = size_t(p->N) * p->B;
    double *x = nullptr, *y = nullptr;
    if (hipMalloc(&x, n * sizeof(double)) != hipSuccess || hipMalloc(&y, n * sizeof(double)) != hipSuccess) {
        if (x) hipFree(x);
        set_error("kmx_plan_autotune: allocation failed");
        return KMX_E_NOMEM;
    }
    hipMemsetAsync(x, 0, n * sizeof(double), s);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    int rc = KMX_OK;
    for (int prec : {KMX_PREC_FP64, KMX_PREC_FP32}) {
        if (prec == KMX_PREC_FP32 && !p->d_blk32) continue;
        const size_t elem = prec == KMX_PREC_FP64 ? sizeof(double) : sizeof(float);
        const bool f64 = prec == KMX_PREC_FP64;
        const bool has_reg = f64 ? pick_f64(p->D, p->deg, 256, 0) : pick_f32(p->D, p->deg, 256, 0);
        const bool has_split = f64 ? pick_split_f64(p->D, p->deg, 256, false) : pick_split_f32(p->D, p->deg, 256, false);
        kmx_variant bestv = p->var[prec_slot(prec)];
        bestv.sweep = 0;
        float best = 1e30f;
        // 1. pair-by-pair kernel variant, on the whole pair list
        if (has_reg || has_split) {
            for (int kind : {0, 1})
                for (int nt : {128, 256, 512})
                    for (int lds : {1, 2, 0}) {
                        if (lds && !strip_fits(p, elem, lds)) continue;
                        if (lds == 2 && (kind == 1 || p->D < 2)) continue;
                        if ((kind == 0 && !has_reg) || (kind == 1 && (!has_split || nt == 128))) continue;
                        const kmx_variant v{nt, lds, kind, 0};
                        float ms;
                        rc = time_calls([&] { return launch_variant(p, false, x, y, prec, v, s); }, s, e0, e1, ms);
                        if (rc != KMX_OK) break;
                        if (ms < best) {
                            best = ms;
                            bestv = v;
                        }
                    }
        } else {
            rc = time_calls([&] { return launch_variant(p, false, x, y, prec, bestv, s); }, s, e0, e1, best);
        }
        if (rc != KMX_OK) break;
        p->var[prec_slot(prec)] = bestv;
        // 2. the dimension-sweep path against it
        if (p->sw && (prec == KMX_PREC_FP32 || kmx_sweep_fp64_ok(p))) {
            p->var[prec_slot(prec)].sweep = 1;
            float ms;
            rc = time_calls([&] { return kmx_launch_apply(p, false, x, y, prec, s); }, s, e0, e1, ms);
            if (rc != KMX_OK) break;
            bestv.sweep = ms < best ? 1 : 0;
            p->var[prec_slot(prec)] = bestv;
        }
        std::lock_guard<std::mutex> g(tune_mutex());
        tune_cache()[key_of(p, prec)] = bestv;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(x);
    hipFree(y);
    return rc;
}

extern "C" int kmx_plan_set_sweep(kmx_plan* p, int prec, int on) {
    if (p && prec == KMX_PREC_BF16) {  // BF16 has no other path: on is all it takes
        if (on && kmx_sweep_bf16_ok(p)) return KMX_OK;
        set_error(on ? "kmx_plan_set_sweep: this plan has no separable sweep with BF16 values"
                     : "kmx_plan_set_sweep: BF16 applies run on the separable sweep only; it cannot be switched off");
        return KMX_E_UNSUPPORTED;
    }
    if (!p || (prec != KMX_PREC_FP64 && prec != KMX_PREC_FP32)) {
        set_error("kmx_plan_set_sweep: bad arguments");
        return KMX_E_ARG;
    }
    if (on && !p->sw) {
        set_error("kmx_plan_set_sweep: this plan has no dimension-sweep path");
        return KMX_E_UNSUPPORTED;
    }
    if (on && prec == KMX_PREC_FP64 && !kmx_sweep_fp64_ok(p)) {
        set_error("kmx_plan_set_sweep: the separable sweep takes near-identity factors as the identity, which "
                  "FP64 applies would see (declare them in kmx_plan_desc.identity to allow it)");
        return KMX_E_UNSUPPORTED;
    }
    p->var[prec_slot(prec)].sweep = on ? 1 : 0;
    return KMX_OK;
}

extern "C" int kmx_plan_sweep_info(const kmx_plan* p, kmx_sweep_info* info) {
    if (!p || !info) {
        set_error("kmx_plan_sweep_info: bad arguments");
        return KMX_E_ARG;
    }
    kmx_sweep_get_info(p, info);
    return KMX_OK;
}

extern "C" int kmx_plan_get_sweep(const kmx_plan* p, int prec, int* available, int* active, double* flops) {
    if (!p || (prec != KMX_PREC_FP64 && prec != KMX_PREC_FP32 && prec != KMX_PREC_BF16)) {
        set_error("kmx_plan_get_sweep: bad arguments");
        return KMX_E_ARG;
    }
    if (prec == KMX_PREC_BF16) {
        const int ok = kmx_sweep_bf16_ok(p) ? 1 : 0;
        if (available) *available = ok;
        if (active) *active = ok;
        if (flops) *flops = kmx_sweep_flops(p);
        return KMX_OK;
    }
    if (available) *available = (p->sw && (prec == KMX_PREC_FP32 || kmx_sweep_fp64_ok(p))) ? 1 : 0;
    if (active) *active = (p->sw && p->var[prec_slot(prec)].sweep) ? 1 : 0;
    if (flops) *flops = kmx_sweep_flops(p);
    return KMX_OK;
}

extern "C" int kmx_plan_get_variant(const kmx_plan* p, int prec, int* nt, int* lds) {
    // (the lane-split kind is reported through kmx_plan_get_variant's nt sign: negative = split)
    if (!p || (prec != KMX_PREC_FP64 && prec != KMX_PREC_FP32)) {
        set_error("kmx_plan_get_variant: bad arguments");
        return KMX_E_ARG;
    }
    const kmx_variant v = p->var[prec_slot(prec)];
    const size_t elem = prec == KMX_PREC_FP64 ? sizeof(double) : sizeof(float);
    int mode = v.lds < 0 ? (strip_fits(p, elem) ? 1 : 0) : v.lds;
    if (mode && !strip_fits(p, elem, mode)) mode = 0;
    if (mode == 2 && (v.kind == 1 || p->D < 2)) mode = strip_fits(p, elem, 1) ? 1 : 0;
    if (nt) *nt = (v.nt > 0 ? v.nt : 256) * (v.kind == 1 ? -1 : 1);
    if (lds) *lds = mode;
    return KMX_OK;
}

// ------------------------------------------------------------------ batches
// Per precision: the plans whose (d, deg) the persistent kernel's shape switch
// holds in registers go into one work list (instance, row) sorted by cost;
// the others are launched on their own after it.
struct BatchSet {
    std::vector<int> in_kernel;  // plan indices served by the persistent kernel (slot order)
    std::vector<int> solo;       // plans launched on their own
    void* d_args = nullptr;      // KArgs<T>[slots]
    int* d_shape = nullptr;
    int2* d_items = nullptr;
    int nitems = 0;
    int maxB = 1;
    int grid = 0;
};

struct kmx_batch {
    std::vector<kmx_plan*> plans;
    BatchSet set[2];  // [0] FP64, [1] FP32
    bool has32 = true;
    // merged dimension sweeps per precision (kmx_batch_set_sweep): members[s][q] =
    // plan index of io slot q
    kmx_sweep_batch* sweep[2] = {nullptr, nullptr};
    std::vector<int> members[2];
    bool sweep_on[2] = {false, false};
    // plan versions the tables above were built from (kmx_plan::version)
    std::vector<unsigned long long> version;
    // load batches per wave of the merged sweeps (0: the defaults; kmx_batch_autotune)
    int wave_batches[2] = {0, 0};
};

// The switch's register budget is that of its largest case: blocks of up to 27
// values (one lane per pair) share the persistent launch, and in FP64 so do 3D
// degree-4 blocks (four lanes per pair, a row body that needs fewer registers
// than the 27-value one); other shapes run as their own autotuned launches.
static bool batch_shape_ok(const kmx_plan* p, bool fp64) {
    if (p->D > 3 || p->deg > 4) return false;
    return p->B <= 27 || (fp64 && p->D == 3 && p->deg == 4);
}

template <typename T>
static hipError_t build_set(kmx_batch* b, BatchSet& bs, bool fp64) {
    std::vector<KArgs<T>> args;
    std::vector<int> shape;
    struct Item {
        double cost;
        int slot, row;
    };
    std::vector<Item> items;
    std::vector<long long> rp;
    std::vector<char> skip(b->plans.size(), 0);  // plans served by the merged sweeps
    if (b->sweep_on[fp64 ? 0 : 1])
        for (int i : b->members[fp64 ? 0 : 1]) skip[size_t(i)] = 1;
    for (int i = 0; i < int(b->plans.size()); ++i) {
        kmx_plan* p = b->plans[i];
        if (!fp64 && !p->d_blk32) continue;
        if (skip[size_t(i)]) continue;
        if (!batch_shape_ok(p, fp64) || int(bs.in_kernel.size()) >= KMX_BATCH_MAX) {
            bs.solo.push_back(i);
            continue;
        }
        const int slot = int(bs.in_kernel.size());
        bs.in_kernel.push_back(i);
        KArgs<T> a;
        fill_args(a, p, false, nullptr, nullptr);
        args.push_back(a);
        shape.push_back(p->D * 16 + p->deg);
        bs.maxB = std::max(bs.maxB, p->B);
        rp.assign(size_t(p->N) + 1, 0);
        if (p->all_pairs) {
            for (int r = 0; r <= p->N; ++r) rp[r] = (long long)r * p->N;
        } else if (p->N > 0) {
            hipError_t e = hipMemcpy(rp.data(), p->d_row_ptr, sizeof(long long) * (p->N + 1), hipMemcpyDeviceToHost);
            if (e != hipSuccess) return e;
        }
        const double unit = double(p->T) * p->B * p->D * p->deg;
        for (int r = 0; r < p->N; ++r) items.push_back({double(rp[r + 1] - rp[r]) * unit, slot, r});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.cost > y.cost; });
    std::vector<int2> it(items.size());
    for (size_t q = 0; q < items.size(); ++q) it[q] = make_int2(items[q].slot, items[q].row);
    bs.nitems = int(it.size());
    auto up = [](const void* h, size_t bytes, void** d) {
        if (bytes == 0) return hipSuccess;
        hipError_t e = hipMalloc(d, bytes);
        if (e != hipSuccess) return e;
        return hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice);
    };
    hipError_t e = hipSuccess;
    if (!args.empty()) {
        e = up(args.data(), args.size() * sizeof(KArgs<T>), &bs.d_args);
        if (e == hipSuccess) e = up(shape.data(), shape.size() * sizeof(int), reinterpret_cast<void**>(&bs.d_shape));
        if (e == hipSuccess) e = up(it.data(), it.size() * sizeof(int2), reinterpret_cast<void**>(&bs.d_items));
    }
    if (e != hipSuccess) return e;
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const size_t smem = size_t(4 * bs.maxB) * sizeof(T);
    const int occ = fp64 ? batch_occupancy_f64(smem) : batch_occupancy_f32(smem);
    bs.grid = std::max(1, std::min(bs.nitems, cus * std::max(1, occ)));
    return hipSuccess;
}

static void free_set(BatchSet& bs) {
    if (bs.d_args) hipFree(bs.d_args);
    if (bs.d_shape) hipFree(bs.d_shape);
    if (bs.d_items) hipFree(bs.d_items);
    bs = BatchSet();
}

extern "C" int kmx_batch_create(kmx_plan* const* plans, int count, kmx_batch** out) {
    if (!plans || count < 0 || !out) {
        set_error("kmx_batch_create: bad arguments");
        return KMX_E_ARG;
    }
    *out = nullptr;
    kmx_batch* b = new kmx_batch();
    for (int i = 0; i < count; ++i) {
        if (!plans[i]) {
            delete b;
            set_error("kmx_batch_create: null plan");
            return KMX_E_ARG;
        }
        b->plans.push_back(plans[i]);
        if (!plans[i]->d_blk32) b->has32 = false;
    }
    hipError_t e = build_set<double>(b, b->set[0], true);
    if (e == hipSuccess) e = build_set<float>(b, b->set[1], false);
    if (e != hipSuccess) {
        set_error(std::string("kmx_batch_create: ") + hipGetErrorString(e));
        kmx_batch_destroy(b);
        return KMX_E_HIP;
    }
    for (kmx_plan* p : b->plans) b->version.push_back(p->version);
    *out = b;
    return KMX_OK;
}

// The batch's merged sweep for one precision: every plan whose sweep serves that
// precision, each plan once (one arena per plan), at most KMX_BATCH_MAX.
static int build_batch_sweep(kmx_batch* b, int slot) {
    kmx_sweep_batch_free(b->sweep[slot]);
    b->sweep[slot] = nullptr;
    b->members[slot].clear();
    std::vector<kmx_plan*> mp;
    std::vector<int> mi;
    for (int i = 0; i < int(b->plans.size()) && int(mi.size()) < KMX_BATCH_MAX; ++i) {
        kmx_plan* p = b->plans[i];
        if (!p->sw || (slot == 0 && !kmx_sweep_fp64_ok(p)) || (slot == 1 && !p->d_blk32)) continue;
        if (std::find(mp.begin(), mp.end(), p) != mp.end()) continue;
        mp.push_back(p);
        mi.push_back(i);
    }
    if (mp.empty()) {
        set_error("kmx_batch_set_sweep: no plan of the batch has a dimension-sweep path for this precision");
        return KMX_E_UNSUPPORTED;
    }
    try {
        b->sweep[slot] = kmx_sweep_batch_build(mp, b->wave_batches[slot]);
    } catch (const std::exception& ex) {
        set_error(std::string("kmx_batch_set_sweep: ") + ex.what());
        return KMX_E_HIP;
    }
    b->members[slot] = mi;
    return KMX_OK;
}

// Term matrices or scales of a plan changed since the batch's tables were built:
// rebuild them from the plans' current state (synchronous; not under graph capture).
static int refresh_batch(kmx_batch* b) {
    bool changed = false;
    for (size_t i = 0; i < b->plans.size(); ++i) changed = changed || b->plans[i]->version != b->version[i];
    if (!changed) return KMX_OK;
    for (int slot = 0; slot < 2; ++slot) {
        if (b->sweep_on[slot]) {
            const int rc = build_batch_sweep(b, slot);
            if (rc != KMX_OK) return rc;
        }
        free_set(b->set[slot]);
        const hipError_t e = slot == 0 ? build_set<double>(b, b->set[0], true) : build_set<float>(b, b->set[1], false);
        if (e != hipSuccess) {
            set_error(std::string("kmx_batch_apply: refreshing the batch: ") + hipGetErrorString(e));
            return KMX_E_HIP;
        }
    }
    for (size_t i = 0; i < b->plans.size(); ++i) b->version[i] = b->plans[i]->version;
    return KMX_OK;
}

extern "C" int kmx_batch_apply(kmx_batch* b, const double* const* xs, double* const* ys, int prec, void* stream) {
    if (!b || (!b->plans.empty() && (!xs || !ys))) {
        set_error("kmx_batch_apply: bad arguments");
        return KMX_E_ARG;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if ((prec == KMX_PREC_FP32 || prec == KMX_PREC_BF16) && !b->has32) {
        set_error("kmx_batch_apply: a plan has no FP32 blocks");
        return KMX_E_ARG;
    }
    if (prec != KMX_PREC_FP32 && prec != KMX_PREC_FP64 && prec != KMX_PREC_BF16) {
        set_error("kmx_batch_apply: unknown precision");
        return KMX_E_ARG;
    }
    for (size_t i = 0; i < b->plans.size(); ++i)
        if (b->plans[i]->N > 0 && (!xs[i] || !ys[i])) {
            set_error("kmx_batch_apply: null vector");
            return KMX_E_ARG;
        }
    const int slot = prec == KMX_PREC_FP64 ? 0 : 1;
    {
        const int rc = refresh_batch(b);
        if (rc != KMX_OK) return rc;
    }
    if (prec == KMX_PREC_BF16) {  // the FP32 merged sweep with BF16 values; every plan must be in it
        if (!b->sweep_on[1] || b->members[1].size() != b->plans.size()) {
            set_error("kmx_batch_apply: BF16 runs on the FP32 merged sweep (kmx_batch_set_sweep) with every plan "
                      "a separable member");
            return KMX_E_UNSUPPORTED;
        }
        BatchIO io;
        std::memset(&io, 0, sizeof(io));
        const std::vector<int>& mem = b->members[1];
        for (size_t q = 0; q < mem.size(); ++q) {
            io.x[q] = xs[mem[q]];
            io.y[q] = ys[mem[q]];
        }
        return kmx_sweep_batch_apply(b->sweep[1], 2, io, s);
    }
    if (b->sweep_on[slot]) {
        BatchIO io;
        std::memset(&io, 0, sizeof(io));
        const std::vector<int>& mem = b->members[slot];
        for (size_t q = 0; q < mem.size(); ++q) {
            io.x[q] = xs[mem[q]];
            io.y[q] = ys[mem[q]];
        }
        const int rc = kmx_sweep_batch_apply(b->sweep[slot], slot, io, s);
        if (rc != KMX_OK) return rc;
    }
    BatchSet& bs = b->set[slot];
    if (bs.nitems > 0) {
        BatchIO io;
        std::memset(&io, 0, sizeof(io));
        for (size_t q = 0; q < bs.in_kernel.size(); ++q) {
            io.x[q] = xs[bs.in_kernel[q]];
            io.y[q] = ys[bs.in_kernel[q]];
        }
        hipError_t e = (prec == KMX_PREC_FP64)
                           ? launch_batch_f64(reinterpret_cast<const KArgs<double>*>(bs.d_args), bs.d_items, bs.nitems,
                                              bs.d_shape, io, bs.grid, size_t(4 * bs.maxB) * sizeof(double), s)
                           : launch_batch_f32(reinterpret_cast<const KArgs<float>*>(bs.d_args), bs.d_items, bs.nitems,
                                              bs.d_shape, io, bs.grid, size_t(4 * bs.maxB) * sizeof(float), s);
        if (e != hipSuccess) {
            set_error(std::string("kmx_batch_apply: ") + hipGetErrorString(e));
            return KMX_E_HIP;
        }
    }
    for (int i : bs.solo) {
        int rc = kmx_launch_apply(b->plans[i], false, xs[i], ys[i], prec, s);
        if (rc != KMX_OK) return rc;
    }
    return KMX_OK;
}

extern "C" int kmx_batch_destroy(kmx_batch* b) {
    if (!b) return KMX_OK;
    free_set(b->set[0]);
    free_set(b->set[1]);
    kmx_sweep_batch_free(b->sweep[0]);
    kmx_sweep_batch_free(b->sweep[1]);
    delete b;
    return KMX_OK;
}

extern "C" int kmx_batch_set_sweep(kmx_batch* b, int prec, int on) {
    if (!b || (prec != KMX_PREC_FP64 && prec != KMX_PREC_FP32 && prec != KMX_PREC_BF16)) {
        set_error("kmx_batch_set_sweep: bad arguments");
        return KMX_E_ARG;
    }
    const int slot = prec_slot(prec);  // BF16: the FP32 merged sweep, whose tables it shares
    if (bool(on) == b->sweep_on[slot]) return KMX_OK;
    if (on) {
        const int rc = build_batch_sweep(b, slot);
        if (rc != KMX_OK) return rc;
    } else {
        kmx_sweep_batch_free(b->sweep[slot]);
        b->sweep[slot] = nullptr;
        b->members[slot].clear();
    }
    b->sweep_on[slot] = on != 0;
    free_set(b->set[slot]);
    const hipError_t e = slot == 0 ? build_set<double>(b, b->set[0], true) : build_set<float>(b, b->set[1], false);
    if (e != hipSuccess) {
        set_error(std::string("kmx_batch_set_sweep: ") + hipGetErrorString(e));
        return KMX_E_HIP;
    }
    return KMX_OK;
}

extern "C" int kmx_batch_autotune(kmx_batch* b, int prec, int* wave_batches, void* stream) {
    if (!b || (prec != KMX_PREC_FP64 && prec != KMX_PREC_FP32 && prec != KMX_PREC_BF16)) {
        set_error("kmx_batch_autotune: bad arguments");
        return KMX_E_ARG;
    }
    const int slot = prec_slot(prec);
    if (!b->sweep_on[slot]) {
        set_error("kmx_batch_autotune: the merged sweep is off for this precision (kmx_batch_set_sweep)");
        return KMX_E_UNSUPPORTED;
    }
    {
        const int rc = refresh_batch(b);
        if (rc != KMX_OK) return rc;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // scratch x / y of every plan, x all ones (zero operands time up to 8 % faster on
    // the matrix cores than data, and can rank the candidates differently)
    std::vector<double*> xs(b->plans.size(), nullptr), ys(b->plans.size(), nullptr);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = KMX_OK;
    auto cleanup = [&]() {
        for (double* p : xs) if (p) (void)hipFree(p);
        for (double* p : ys) if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    size_t nmax = 1;
    for (const kmx_plan* p : b->plans) nmax = std::max(nmax, size_t(p->N) * p->B);
    const std::vector<double> ones(nmax, 1.0);
    for (size_t i = 0; i < b->plans.size(); ++i) {
        const size_t n = std::max<size_t>(1, size_t(b->plans[i]->N) * b->plans[i]->B);
        if (hipMalloc(&xs[i], n * sizeof(double)) != hipSuccess || hipMalloc(&ys[i], n * sizeof(double)) != hipSuccess ||
            hipMemcpy(xs[i], ones.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            cleanup();
            set_error("kmx_batch_autotune: out of device memory");
            return KMX_E_HIP;
        }
    }
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        cleanup();
        set_error("kmx_batch_autotune: event set-up failed");
        return KMX_E_HIP;
    }
    // load batches per wave before a tile is split: 4, 8 (the separable default), 16
    const int cand[3] = {4, 8, 16};
    int best = b->wave_batches[slot];
    float best_ms = 0.0f;
    for (int c : cand) {
        b->wave_batches[slot] = c;
        rc = build_batch_sweep(b, slot);
        if (rc != KMX_OK) break;
        float ms = 0.0f;
        for (int round = 0; round < 3 && rc == KMX_OK; ++round) {  // best of three rounds of 10 applies
            for (int w = 0; w < (round ? 0 : 3) && rc == KMX_OK; ++w) rc = kmx_batch_apply(b, xs.data(), ys.data(), prec, s);
            if (rc != KMX_OK) break;
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < 10 && rc == KMX_OK; ++r) rc = kmx_batch_apply(b, xs.data(), ys.data(), prec, s);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = KMX_E_HIP;
                set_error("kmx_batch_autotune: timing failed");
                break;
            }
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, e0, e1);
            ms = round ? std::min(ms, t) : t;
        }
        if (rc != KMX_OK) break;
        if (best_ms == 0.0f || ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    if (rc == KMX_OK) {
        b->wave_batches[slot] = best;
        rc = build_batch_sweep(b, slot);
    }
    if (rc == KMX_OK) {
        free_set(b->set[slot]);
        const hipError_t e = slot == 0 ? build_set<double>(b, b->set[0], true) : build_set<float>(b, b->set[1], false);
        if (e != hipSuccess) {
            set_error(std::string("kmx_batch_autotune: ") + hipGetErrorString(e));
            rc = KMX_E_HIP;
        }
    }
    (void)hipDeviceSynchronize();
    cleanup();
    if (rc == KMX_OK && wave_batches) *wave_batches = best;
    return rc;
}

extern "C" int kmx_batch_rk3(kmx_batch* b, double* const* fs, const double* dts, int nsteps, int prec, void* stream) {
    if (!b || !fs || !dts || nsteps < 0 ||
        (prec != KMX_PREC_FP64 && prec != KMX_PREC_FP32 && prec != KMX_PREC_BF16)) {
        set_error("kmx_batch_rk3: bad arguments");
        return KMX_E_ARG;
    }
    if (nsteps == 0 || b->plans.empty()) return KMX_OK;
    const int slot = prec_slot(prec);  // BF16: the FP32 merged sweep's tables
    {
        const int rc = refresh_batch(b);
        if (rc != KMX_OK) return rc;
    }
    const std::vector<int>& mem = b->members[slot];
    if (!b->sweep_on[slot] || mem.size() != b->plans.size()) {
        set_error("kmx_batch_rk3: every plan of the batch must be a member of its merged sweep for this precision "
                  "(kmx_batch_set_sweep; distinct plans)");
        return KMX_E_UNSUPPORTED;
    }
    std::vector<double*> f(mem.size());
    std::vector<double> dt(mem.size());
    for (size_t q = 0; q < mem.size(); ++q) {
        f[q] = fs[mem[q]];
        dt[q] = dts[mem[q]];
        if (!f[q] && b->plans[size_t(mem[q])]->N > 0) {
            set_error("kmx_batch_rk3: null state");
            return KMX_E_ARG;
        }
    }
    return kmx_sweep_batch_rk3(b->sweep[slot], prec == KMX_PREC_BF16 ? 2 : slot, f.data(), dt.data(), nsteps,
                               reinterpret_cast<hipStream_t>(stream));
}

extern "C" int kmx_batch_get_sweep(const kmx_batch* b, int prec, int* members, int* active, int* phases) {
    if (!b || (prec != KMX_PREC_FP64 && prec != KMX_PREC_FP32 && prec != KMX_PREC_BF16)) {
        set_error("kmx_batch_get_sweep: bad arguments");
        return KMX_E_ARG;
    }
    const int slot = prec_slot(prec);
    const bool act = b->sweep_on[slot] &&
                     (prec != KMX_PREC_BF16 ||
                      (b->members[1].size() == b->plans.size() && kmx_sweep_batch_has16(b->sweep[1])));
    if (members) *members = act ? int(b->members[slot].size()) : 0;
    if (active) *active = act ? 1 : 0;
    if (phases) *phases = act ? kmx_sweep_batch_phases(b->sweep[slot]) : 0;
    return KMX_OK;
}
