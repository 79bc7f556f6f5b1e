// Batched form: thousands of independent tensor-product integrals of mixed
// rule, family, dimension and size in one launch per LDS class.
//
// The host plan splits every integral into tiles of ~tile_points points
// (one 64-lane group x a range of loop items, as grid_tile walks them) and
// orders the tile list by decreasing integral size (longest first), the
// tiles of an integral contiguous.  Each wave takes one tile (wave index =
// tile index, so the dispatcher hands out the large tiles first), stages that
// integral's axis maps and loop-axis table in its own LDS slice, runs the
// tile and writes one FP64 partial.  A second kernel reduces each integral's
// tile partials in tile order, so results are deterministic for a given
// tiling.  No memset is issued and no counter is kept, so the launches can be
// captured in a graph and replayed.
#include <algorithm>
#include <mutex>

#include "tq_grid_tile.hpp"

namespace tq {

template <typename T>
struct BatchDev {
    int rule, fam, d, n, m, L, q, nl, lane_count, lane_waves;
    long long items;
    long long gl_offset;
    double scale;
    Axis ax[TQ_MAX_DIM];
    Params<T> p;
};

template <typename T>
struct BatchTab {
    static constexpr int N = (sizeof(T) == 4) ? 1024 : 512;  // general kernel: per-wave loop table entries
};

// Per-wave LDS region of the row-table kernels, in bytes: class 0 (small
// tables, high occupancy) and class 1 (loop tables up to 8 KB).
constexpr int kRegionSmall = 4096;
constexpr int kRegionBig = 8192;
constexpr int kBatchClasses = 4;  // 0: one-axis kernel; 1, 2: row-table kernels; 3: general kernel

// LDS bytes of an integral's tables in the row-table kernels:
// [row weights L*n FP64][loop table S*nl T][row terms slots*L*n T] (S = tab_stride, slots = state_slots)
inline long long rowtab_bytes(int fam, int n, int nl, int L, int tsize) {
    const int S = tab_stride(fam);
    if (L < 0) return 0;
    return (long long)L * n * 8 + (long long)S * nl * tsize + (long long)state_slots(fam) * L * n * tsize;
}


// General kernel: loop-axis table when it fits, row axes recomputed per row.
template <typename T, int FAM, int PATHS>
__device__ __forceinline__ double batch_tile_general(const BatchDev<T>& D, const Axis* ax, T* tab, bool table,
                                            const double* glx, const double* glw, const float* gp32, int lw0,
                                            int lw1, long long t0, long long t1) {
    if (table) {
        const int lane = threadIdx.x & (kWave - 1);
        if (D.q == 2 && D.n <= 32)
            pair_table_wave<T, FAM>(tab, ax, D.L, D.n, D.rule, glx, glw, D.p);
        else
            for (int j = lane; j < D.nl; j += kWave)
                loop_table_entry<T, FAM>(tab, j, ax, D.L, D.q, D.n, D.rule, glx, glw, D.p);
    }
    // make this wave's LDS writes visible to all of its lanes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    double acc = 0.0;
    for (int lw = lw0; lw < lw1; ++lw)
        acc += table ? grid_tile<T, FAM, true, false, true, PATHS>(D.d, D.n, D.nl, D.m, D.L, D.rule, D.lane_count,
                                                                  D.p, ax, tab, glx, glw, lw, t0, t1, nullptr,
                                                                  nullptr, gp32)
                     : grid_tile<T, FAM, false, false, true, PATHS>(D.d, D.n, D.nl, D.m, D.L, D.rule, D.lane_count,
                                                                   D.p, ax, tab, glx, glw, lw, t0, t1, nullptr,
                                                                   nullptr, gp32);
    return acc;
}

// Row-table kernels: the planner guarantees that the loop and row tables of
// every integral they see fit the region (or that it has no loop axis), and
// sends them multi-axis integrals only.
template <typename T, int FAM>
__device__ __forceinline__ double batch_tile_rowtab(const BatchDev<T>& D, const Axis* ax, unsigned char* region,
                                           const double* glx, const double* glw, const float* gp32, int lw0,
                                           int lw1, long long t0, long long t1) {
    const int n = D.n, L = D.L;
    double* rw = reinterpret_cast<double*>(region);
    T* tab = reinterpret_cast<T*>(region + (L > 0 ? (size_t)L * n * sizeof(double) : 0));
    T* rt = tab + (size_t)TabStride<FAM>::v * D.nl;
    if (D.m > 0 && L >= 0) {
        const int lane = threadIdx.x & (kWave - 1);
        // two loop axes (n <= 32 in a batch): nodes and weights formed once
        // per axis, not once per entry
        if (D.q == 2 && n <= 32)
            pair_table_wave<T, FAM>(tab, ax, L, n, D.rule, glx, glw, D.p);
        else
            for (int j = lane; j < D.nl; j += kWave)
                loop_table_entry<T, FAM>(tab, j, ax, L, D.q, n, D.rule, glx, glw, D.p);
        for (int e = lane; e < L * n; e += kWave) {
            const int k = e / n;
            row_table_entry<T, FAM>(rt, rw, k, e - k * n, ax[k], axis_c(D.p, k), n, D.rule, glx, glw, D.p);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    double acc = 0.0;
    for (int lw = lw0; lw < lw1; ++lw)
        acc += grid_tile<T, FAM, true, true, true, kPathsMultiAxis>(D.d, n, D.nl, D.m, L, D.rule, D.lane_count, D.p,
                                                                     ax, tab, glx, glw, lw, t0, t1, rt, rw, gp32);
    return acc;
}

template <typename T, int REGION, int PATHS>
__device__ __forceinline__ double batch_tile(const BatchDev<T>& D, const Axis* ax, unsigned char* region, bool table,
                                    const double* glx, const double* glw, const float* gp32, int lw0, int lw1,
                                    long long t0, long long t1) {
    switch (D.fam) {
#define TQ_BT(F)                                                                                                    \
    case F:                                                                                                         \
        return REGION > 0 ? batch_tile_rowtab<T, F>(D, ax, region, glx, glw, gp32, lw0, lw1, t0, t1)                \
                          : batch_tile_general<T, F, PATHS>(D, ax, reinterpret_cast<T*>(region), table, glx, glw,  \
                                                            gp32, lw0, lw1, t0, t1);
        TQ_BT(TQ_FAM_GAUSSIAN_PEAK)
        TQ_BT(TQ_FAM_OSCILLATORY)
        TQ_BT(TQ_FAM_PRODUCT_PEAK)
        TQ_BT(TQ_FAM_EXP_LINEAR)
        TQ_BT(TQ_FAM_SIN_SUM)
        TQ_BT(TQ_FAM_DOUBLE_GAUSSIAN)
        TQ_BT(TQ_FAM_CONSTANT)
#undef TQ_BT
        default: return 0.0;
    }
}

// One wave per tile of one class, tiles [0, n_tiles) in the planner's order
// (longest integrals first), so the hardware dispatcher hands them out
// largest first as waves finish.  REGION > 0: row-table kernel with a
// per-wave LDS region of REGION bytes; REGION == 0: general kernel.  PATHS
// (grid_tile): the one-axis class runs a kernel with only the one-axis code,
// the others one with only the multi-axis code (83 VGPRs for the 4 KB
// row-table kernel instead of 112 with both).  A
// persistent form (each wave claiming tiles from 16 atomic queues) held
// its claim loop's state and the hoisted per-tile invariants for the whole
// kernel: 164 VGPRs against 112, and the 8192-integral Genz batch took
// 0.441 ms against 0.344.
template <typename T, int REGION, int PATHS>
__global__ __launch_bounds__(kBlock) void k_batch(const BatchDev<T>* descs, const tq_batch_tile* tiles,
                                                  long long n_tiles, const double* glx, const double* glw,
                                                  const float* glp32, long long gl_len, double* tile_part) {
    constexpr int NT = BatchTab<T>::N;
    // one-axis integrals use no LDS table
    constexpr int RB = REGION > 0 ? REGION : (PATHS == kPathsOneAxis ? 8 : (int)(2 * NT * sizeof(T)));
    __shared__ Axis s_ax[kWavesPerBlock][TQ_MAX_DIM];
    __shared__ double s_region[kWavesPerBlock][RB / 8];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    unsigned char* region = reinterpret_cast<unsigned char*>(s_region[wave]);
    const long long t = (long long)blockIdx.x * kWavesPerBlock + wave;
    if (t >= n_tiles) return;
    const tq_batch_tile tile = tiles[t];
    const BatchDev<T>& D = descs[tile.item];
    // A Gauss-Legendre integral whose n nodes at gl_offset are not all
    // inside the gl_len-entry tables reads nothing and yields NaN
    // (tq_batch_plan checks items without the tables; this is where their
    // length is known).
    const bool gl_ok = D.rule != TQ_RULE_GAUSS_LEGENDRE ||
                       (glx && glw && D.gl_offset >= 0 && D.gl_offset + D.n <= gl_len);
    double acc = __longlong_as_double(0x7ff8000000000000ll);
    if (gl_ok) {
        const double* gx = glx ? glx + D.gl_offset : nullptr;
        const double* gw = glw ? glw + D.gl_offset : nullptr;
        // FP32 (node, weight) pairs for one-axis Gauss-Legendre rows of FP32 batches
        const float* gp = (glp32 && sizeof(T) == 4 && D.m == 0) ? glp32 + 2 * D.gl_offset : nullptr;
        for (int k = lane; k < D.d; k += kWave) s_ax[wave][k] = D.ax[k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const bool table = REGION == 0 && PATHS != kPathsOneAxis && D.m > 0 && D.L >= 0 &&
                           D.nl * tab_stride(D.fam) <= 2 * NT;
        acc = batch_tile<T, REGION, PATHS>(D, s_ax[wave], region, table, gx, gw, gp, tile.lane_wave,
                                           tile.lane_wave_end, tile.t0, tile.t1);
        acc = wave_sum(acc);
    }
    if (lane == 0) tile_part[t] = acc;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_batch_finalize(const BatchDev<T>* descs, long long count,
                                                           const int64_t* tile_begin, const int64_t* tile_count,
                                                           const double* tile_part, T* out, double* out64) {
    const long long i = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (i >= count) return;
    const long long b = tile_begin[i], c = tile_count[i];
    double v = 0.0;
    for (long long j = lane; j < c; j += kWave) v += tile_part[b + j];
    v = wave_sum(v);
    if (lane == 0) {
        const double r = v * descs[i].scale;
        if (out) out[i] = (T)r;
        if (out64) out64[i] = r;
    }
}

static int check_item(const tq_batch_item& it) {
    if (it.family < 0 || it.family >= TQ_NUM_FAMILIES) return fail(TQ_E_ARG, "batch item: unknown family");
    if (it.dim < 1 || it.dim > TQ_MAX_DIM) return fail(TQ_E_ARG, "batch item: dim out of range (1..16)");
    switch (it.rule) {
        case TQ_RULE_TRAPEZOID:
            if (it.n < 2) return fail(TQ_E_ARG, "batch item: Trapezoid needs n >= 2");
            break;
        case TQ_RULE_SIMPSON:
            if (it.n < 3 || it.n % 2 != 1) return fail(TQ_E_ARG, "batch item: Simpson needs odd n >= 3");
            break;
        case TQ_RULE_BOOLE:
            if (it.n < 5 || (it.n - 1) % 4 != 0) return fail(TQ_E_ARG, "batch item: Boole needs n = 1 + 4k");
            break;
        case TQ_RULE_GAUSS_LEGENDRE:
            if (it.n < 1 || it.gl_offset < 0) return fail(TQ_E_ARG, "batch item: bad Gauss-Legendre n/offset");
            break;
        default: return fail(TQ_E_ARG, "batch item: unknown rule");
    }
    long long P;
    if (ipow_checked(it.n, it.dim, 1LL << 62, &P) != 0) return fail(TQ_E_ARG, "batch item: n^dim too large");
    return TQ_OK;
}

// Split-mode integrals (one axis, m == 0) walk one 64-node row per item
// (Gauss-Legendre: 64 nodes and their 64 mirrors).  Gauss-Legendre rows
// wait on global loads of the node tables, so shorter tiles spread them
// over more waves; Newton-Cotes rows compute their nodes and take longer
// tiles, which spread the per-tile cost over more points.
constexpr long long kSplitMaxItemsGL = 128;
constexpr long long kSplitMaxItemsNC = 512;

static long long tile_ipw(const GridGeom& G, long long tile_points) {
    long long ipw = tile_points / kWave;
    const long long cap = G.rule == TQ_RULE_GAUSS_LEGENDRE ? kSplitMaxItemsGL : kSplitMaxItemsNC;
    if (G.m == 0 && ipw > cap) ipw = cap;
    if (ipw < 1) ipw = 1;
    if (ipw > G.items) ipw = G.items;
    return ipw;
}

template <typename T>
static int plan_batch(const tq_batch_item* items, long long count, long long tile_points, BatchDev<T>* desc,
                      tq_batch_tile* tiles, int64_t* tile_begin, int64_t* tile_count, long long* n_tiles_out,
                      int64_t* class_offsets = nullptr) {
    struct Entry {
        int cls;
        long long cost;
        long long idx;
    };
    Entry* order = new Entry[count > 0 ? count : 1];
    GridGeom* geo = new GridGeom[count > 0 ? count : 1];
    long long* group = new long long[count > 0 ? count : 1];  // 64-lane groups per tile
    long long total = 0;
    for (long long i = 0; i < count; ++i) {
        const tq_batch_item& it = items[i];
        int rc = check_item(it);
        if (!rc) rc = plan_geom(it.dim, it.n, it.rule, 0, 1, 1, geo[i]);
        if (rc) {
            delete[] order;
            delete[] geo;
            delete[] group;
            return rc;
        }
        // two loop axes when their table keeps the integral in the same
        // row-table class (so the kernel's occupancy does not change)
        if (geo[i].m > 0 && geo[i].L >= 0 && it.dim >= 3 && it.n <= kMergeMaxN) {
            const long long b1 = rowtab_bytes(it.family, it.n, geo[i].nl, geo[i].L, (int)sizeof(T));
            const long long region = b1 <= kRegionSmall ? kRegionSmall : (b1 <= kRegionBig ? kRegionBig : 0);
            GridGeom g2;
            if (region > 0 && plan_geom(it.dim, it.n, it.rule, 0, 1, 1, g2, kLoopTable) == TQ_OK && g2.q == 2 &&
                rowtab_bytes(it.family, it.n, g2.nl, g2.L, (int)sizeof(T)) <= region)
                geo[i] = g2;
        }
        const long long ipw = tile_ipw(geo[i], tile_points);
        const long long chunks = (geo[i].items + ipw - 1) / ipw;
        geo[i].ipw = ipw;
        geo[i].chunks = chunks;
        // a lane group's whole loop is shorter than a tile: several groups per tile
        long long g = 1;
        if (chunks == 1) {
            g = tile_points / ((long long)kWave * geo[i].items);
            if (g < 1) g = 1;
            if (g > geo[i].lane_waves) g = geo[i].lane_waves;
        }
        group[i] = g;
        long long P;
        ipow_checked(it.n, it.dim, 1LL << 62, &P);
        // class 0: one-axis integrals; 1, 2: multi-axis ones whose tables fit
        // 4 KB / 8 KB (or that have no loop axis); 3: the rest
        int cls = geo[i].m == 0 ? 0 : 1;
        if (geo[i].m > 0 && geo[i].L >= 0) {
            const long long b = rowtab_bytes(it.family, it.n, geo[i].nl, geo[i].L, (int)sizeof(T));
            cls = b <= kRegionSmall ? 1 : (b <= kRegionBig ? 2 : 3);
        }
        order[i] = {cls, P, i};
        total += chunks * ((geo[i].lane_waves + g - 1) / g);
        if (desc) {
            BatchDev<T>& D = desc[i];
            D.rule = it.rule;
            D.fam = it.family;
            D.d = it.dim;
            D.n = it.n;
            D.m = geo[i].m;
            D.L = geo[i].L;
            D.q = geo[i].q;
            D.nl = geo[i].nl;
            D.lane_count = geo[i].lane_count;
            D.lane_waves = geo[i].lane_waves;
            D.items = geo[i].items;
            D.gl_offset = it.rule == TQ_RULE_GAUSS_LEGENDRE ? it.gl_offset : 0;
            double scale = it.c0;
            for (int k = 0; k < TQ_MAX_DIM; ++k) {
                if (k < it.dim) {
                    D.ax[k] = make_axis_host(it.domain[2 * k], it.domain[2 * k + 1], it.n, it.rule);
                    scale *= axis_scale(it.domain[2 * k], it.domain[2 * k + 1], it.n, it.rule);
                } else {
                    D.ax[k] = Axis{0.0, 0.0, 0.0};
                }
            }
            D.scale = scale;
            tq_integrand f{};
            f.family = it.family;
            f.dim = it.dim;
            f.c0 = it.c0;
            f.c1 = it.c1;
            f.c2 = it.c2;
            f.c3 = it.c3;
            for (int k = 0; k < TQ_MAX_DIM; ++k) {
                f.a[k] = it.a[k];
                f.u[k] = it.u[k];
            }
            D.p = make_params<T>(f);
        }
    }
    if (tiles) {
        // by kernel class, then longest integrals first; stable on the index
        std::sort(order, order + count, [](const Entry& x, const Entry& y) {
            if (x.cls != y.cls) return x.cls < y.cls;
            return x.cost != y.cost ? x.cost > y.cost : x.idx < y.idx;
        });
        long long t = 0;
        int next_cls = 0;
        for (long long o = 0; o < count; ++o) {
            const long long i = order[o].idx;
            while (class_offsets && next_cls <= order[o].cls) class_offsets[next_cls++] = t;
            const GridGeom& G = geo[i];
            tile_begin[i] = t;
            const int g = (int)group[i];
            for (int lw = 0; lw < G.lane_waves; lw += g) {
                for (long long c = 0; c < G.chunks; ++c) {
                    tq_batch_tile& tl = tiles[t++];
                    tl.item = (int32_t)i;
                    tl.lane_wave = lw;
                    tl.lane_wave_end = (lw + g < G.lane_waves) ? lw + g : G.lane_waves;
                    tl.reserved = 0;
                    tl.t0 = c * G.ipw;
                    tl.t1 = (c + 1) * G.ipw < G.items ? (c + 1) * G.ipw : G.items;
                }
            }
            tile_count[i] = t - tile_begin[i];
        }
        while (class_offsets && next_cls <= kBatchClasses) class_offsets[next_cls++] = t;
    }
    *n_tiles_out = total;
    delete[] order;
    delete[] geo;
    delete[] group;
    return TQ_OK;
}

template <typename T, int REGION, int PATHS>
static int launch_class(const void* desc, const tq_batch_tile* tiles, long long n, const double* glx,
                        const double* glw, const float* glp32, long long gl_len, double* tile_part, hipStream_t s) {
    if (n <= 0) return TQ_OK;
    const long long blocks = (n + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL((k_batch<T, REGION, PATHS>), dim3((unsigned)blocks), dim3(kBlock), 0, s,
                       (const BatchDev<T>*)desc, tiles, n, glx, glw, glp32, gl_len, tile_part);
    return check_launch("k_batch");
}

// Library-owned streams on which classes of a batch run beside the caller's
// stream: the one-axis class (Gauss-Legendre table reads from HBM) and the
// small classes overlap the large multi-axis class (VALU-bound) instead of
// following it.  Created once per device (default priority) and kept; fork
// and join go through events, so a captured run is one forked subgraph.
// Runs on one device enqueue their fork, launches and join under that
// device's lock, since the events are shared.
constexpr int kMaxBatchDevices = 64;
struct BatchAux {
    hipStream_t st[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr;
    hipEvent_t join[2] = {nullptr, nullptr};
    std::mutex mu;
    bool ready = false;
};
static BatchAux g_aux[kMaxBatchDevices];
static std::mutex g_aux_mu;

static int hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return TQ_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return (int)e;
}

static int batch_aux(BatchAux** out) {
    int dev = -1;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxBatchDevices) return fail(TQ_E_UNSUPPORTED, "device index beyond the batch stream table");
    std::lock_guard<std::mutex> lock(g_aux_mu);
    BatchAux& a = g_aux[dev];
    if (!a.ready) {
        for (int i = 0; i < 2 && !rc; ++i)
            rc = hip_check(hipStreamCreateWithFlags(&a.st[i], hipStreamNonBlocking), "hipStreamCreateWithFlags");
        if (!rc) rc = hip_check(hipEventCreateWithFlags(&a.fork, hipEventDisableTiming), "hipEventCreateWithFlags");
        for (int i = 0; i < 2 && !rc; ++i)
            rc = hip_check(hipEventCreateWithFlags(&a.join[i], hipEventDisableTiming), "hipEventCreateWithFlags");
        if (rc) return rc;
        a.ready = true;
    }
    *out = &a;
    return TQ_OK;
}

template <typename T>
static int run_classes(const void* desc, const tq_batch_tile* tiles, const int64_t* off, const double* glx,
                       const double* glw, const float* glp32, long long gl_len, double* tile_part, hipStream_t s0,
                       hipStream_t s1, hipStream_t s2) {
    // s0: the 4 KB multi-axis class; s1: one-axis integrals; s2: the 8 KB and general classes
    int rc = launch_class<T, kRegionSmall, kPathsMultiAxis>(desc, tiles + off[1], off[2] - off[1], glx, glw, glp32,
                                                            gl_len, tile_part + off[1], s0);
    if (!rc)
        rc = launch_class<T, 0, kPathsOneAxis>(desc, tiles + off[0], off[1] - off[0], glx, glw, glp32, gl_len,
                                               tile_part + off[0], s1);
    if (!rc)
        rc = launch_class<T, kRegionBig, kPathsMultiAxis>(desc, tiles + off[2], off[3] - off[2], glx, glw, glp32,
                                                          gl_len, tile_part + off[2], s2);
    if (!rc)
        rc = launch_class<T, 0, kPathsMultiAxis>(desc, tiles + off[3], off[4] - off[3], glx, glw, glp32, gl_len,
                                                 tile_part + off[3], s2);
    return rc;
}

template <typename T>
static int run_batch(const void* desc, long long count, const tq_batch_tile* tiles, long long n_tiles,
                     const int64_t* tile_begin, const int64_t* tile_count, const double* glx, const double* glw,
                     const float* glp32, long long gl_len, double* tile_part, void* out, double* out64,
                     const int64_t* class_offsets, hipStream_t s) {
    int rc = TQ_OK;
    if (!class_offsets) {
        // no classes: every tile on the general kernel with both paths
        rc = launch_class<T, 0, kPathsAll>(desc, tiles, n_tiles, glx, glw, glp32, gl_len, tile_part, s);
    } else {
        const int64_t* off = class_offsets;
        for (int c = 0; c < kBatchClasses; ++c)
            if (off[0] != 0 || off[c] < 0 || off[c] > off[c + 1] || off[kBatchClasses] != n_tiles)
                return fail(TQ_E_ARG, "class_offsets do not partition the tiles");
        int busy = 0;
        for (int c = 0; c < kBatchClasses; ++c) busy += off[c + 1] > off[c];
        if (busy <= 1) {
            rc = run_classes<T>(desc, tiles, off, glx, glw, glp32, gl_len, tile_part, s, s, s);
        } else {
            BatchAux* aux = nullptr;
            rc = batch_aux(&aux);
            if (rc) return rc;
            std::lock_guard<std::mutex> lock(aux->mu);
            rc = hip_check(hipEventRecord(aux->fork, s), "hipEventRecord");
            for (int i = 0; i < 2 && !rc; ++i) rc = hip_check(hipStreamWaitEvent(aux->st[i], aux->fork, 0), "hipStreamWaitEvent");
            if (!rc) rc = run_classes<T>(desc, tiles, off, glx, glw, glp32, gl_len, tile_part, s, aux->st[0], aux->st[1]);
            for (int i = 0; i < 2 && !rc; ++i) {
                rc = hip_check(hipEventRecord(aux->join[i], aux->st[i]), "hipEventRecord");
                if (!rc) rc = hip_check(hipStreamWaitEvent(s, aux->join[i], 0), "hipStreamWaitEvent");
            }
        }
    }
    if (rc) return rc;
    const long long fb = (count + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL((k_batch_finalize<T>), dim3((unsigned)(fb > 0 ? fb : 1)), dim3(kBlock), 0, s,
                       (const BatchDev<T>*)desc, count, tile_begin, tile_count, tile_part, (T*)out, out64);
    return check_launch("k_batch_finalize");
}

}  // namespace tq

using namespace tq;

extern "C" {

int tq_batch_plan_sizes(const tq_batch_item* items, int64_t count, int32_t dtype, int64_t tile_points,
                        int64_t* desc_bytes, int64_t* n_tiles) {
    if (!items || !desc_bytes || !n_tiles) return fail(TQ_E_NULL, "NULL argument");
    if (count < 0 || tile_points < 1) return fail(TQ_E_ARG, "count/tile_points out of range");
    if (dtype != TQ_F32 && dtype != TQ_F64) return fail(TQ_E_ARG, "dtype must be TQ_F32 or TQ_F64");
    long long nt = 0;
    int rc = dtype == TQ_F32 ? plan_batch<float>(items, count, tile_points, nullptr, nullptr, nullptr, nullptr, &nt)
                             : plan_batch<double>(items, count, tile_points, nullptr, nullptr, nullptr, nullptr, &nt);
    if (rc) return rc;
    *n_tiles = nt;
    *desc_bytes = (int64_t)count * (dtype == TQ_F32 ? (int64_t)sizeof(BatchDev<float>) : (int64_t)sizeof(BatchDev<double>));
    return TQ_OK;
}

int tq_batch_plan(const tq_batch_item* items, int64_t count, int32_t dtype, int64_t tile_points, void* desc_out,
                  tq_batch_tile* tiles_out, int64_t* tile_begin, int64_t* tile_count, int64_t* class_offsets) {
    if (!items || !desc_out || !tiles_out || !tile_begin || !tile_count) return fail(TQ_E_NULL, "NULL argument");
    if (count < 0 || tile_points < 1) return fail(TQ_E_ARG, "count/tile_points out of range");
    long long nt = 0;
    if (dtype == TQ_F32)
        return plan_batch<float>(items, count, tile_points, (BatchDev<float>*)desc_out, tiles_out, tile_begin,
                                 tile_count, &nt, class_offsets);
    if (dtype == TQ_F64)
        return plan_batch<double>(items, count, tile_points, (BatchDev<double>*)desc_out, tiles_out, tile_begin,
                                  tile_count, &nt, class_offsets);
    return fail(TQ_E_ARG, "dtype must be TQ_F32 or TQ_F64");
}

int tq_batch_run(int32_t dtype, const void* desc, int64_t count, const tq_batch_tile* tiles, int64_t n_tiles,
                 const int64_t* tile_begin, const int64_t* tile_count, const double* gl_nodes,
                 const double* gl_weights, const float* gl_pairs32, int64_t gl_len, double* tile_partials,
                 void* out, double* out64, const int64_t* class_offsets, void* stream) {
    if (!desc || !tiles || !tile_begin || !tile_count || !tile_partials) return fail(TQ_E_NULL, "NULL argument");
    if (gl_len < 0) return fail(TQ_E_ARG, "gl_len must be >= 0");
    if (gl_len > 0 && (!gl_nodes || !gl_weights)) return fail(TQ_E_NULL, "gl_len > 0 with NULL node/weight tables");
    if (count < 0 || n_tiles < 0 || n_tiles > 0xffffffffLL) return fail(TQ_E_ARG, "count/n_tiles out of range");
    if (!out && !out64) return fail(TQ_E_NULL, "no output pointer");
    if (dtype == TQ_F32)
        return run_batch<float>(desc, count, tiles, n_tiles, tile_begin, tile_count, gl_nodes, gl_weights,
                                gl_pairs32, gl_len, tile_partials, out, out64, class_offsets, as_stream(stream));
    if (dtype == TQ_F64)
        return run_batch<double>(desc, count, tiles, n_tiles, tile_begin, tile_count, gl_nodes, gl_weights,
                                 nullptr, gl_len, tile_partials, out, out64, class_offsets, as_stream(stream));
    return fail(TQ_E_ARG, "dtype must be TQ_F32 or TQ_F64");
}

}  // extern "C"
