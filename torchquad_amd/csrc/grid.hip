// Tensor-product grid rules (Trapezoid, Simpson, Boole, Gauss-Legendre).
//
// Index space.  The n^d grid points (ij order, last axis fastest, as
// IntegrationGrid.points in integration_grid.py:98-99) are split into
//   * lane axes  [d-m, d): one combination per lane, fixed for the lane's
//     lifetime (its partial fold and weight are computed once);
//   * loop axis  L = d-m-1: wave-uniform, iterated in the hot loop with its
//     nodes and weights read from an LDS table (broadcast reads);
//   * row axes   [0, L): wave-uniform, advanced by an odometer once per row.
// A wave owns one group of 64 lane combinations and a contiguous range of
// (row, loop-index) items.  Per point the kernel does: fold of the loop axis
// into the shared row+lane partial, the family's finish (transcendental),
// and one FMA into an FP32 partial that is flushed to FP64 at least every 128 points.
// Block partials are FP64 and the final sum runs in a fixed order, so the
// result is deterministic.
#include <mutex>

#include "tq_host.hpp"
#include "tq_grid_tile.hpp"

namespace tq {

constexpr int kMaxDimPoints = 64;  // point writer / weighted reduce

template <typename T>
struct GridArgs {
    GridGeom g;
    const T* domain;
    const double* glx;
    const double* glw;
    double* partials;
    unsigned int* done;  // blocks finished; zero between launches
    T* out;
    double* out64;
    double c0;
    Params<T> p;
};

// Workspace of the fused kernel: completion counters in the first
// kWsHeader bytes (zero before the first launch, and every launch leaves
// them zero), then one FP64 partial per workgroup.  Workgroup b counts into
// group counter b % kDoneGroups; the last workgroup of each group counts
// into the top counter, whose last arrival sums the partials.  One counter
// for all workgroups serialised their atomics at about 15 ns each (10112
// workgroups: 148 us for 3.9e7 evaluations).
constexpr int kDoneGroups = 16;
constexpr int kDoneStride = 64;  // words between counters (256 bytes; 4 KB measured the same)
constexpr int kDoneTop = kDoneGroups * kDoneStride;
constexpr size_t kWsHeader = ((size_t)(kDoneTop + 1) * sizeof(unsigned int) + 255) & ~(size_t)255;

static int check_grid_desc(const tq_grid_desc* g, int max_dim) {
    if (!g) return fail(TQ_E_NULL, "grid descriptor is NULL");
    if (g->dtype != TQ_F32 && g->dtype != TQ_F64) return fail(TQ_E_ARG, "dtype must be TQ_F32 or TQ_F64");
    if (g->dim < 1 || g->dim > max_dim) return fail(TQ_E_ARG, "grid dim out of range");
    switch (g->rule) {
        case TQ_RULE_TRAPEZOID:
            if (g->n < 2) return fail(TQ_E_ARG, "Trapezoid needs n >= 2 points per axis");
            break;
        case TQ_RULE_SIMPSON:
            if (g->n < 3 || (g->n % 2) != 1) return fail(TQ_E_ARG, "Simpson needs odd n >= 3");
            break;
        case TQ_RULE_BOOLE:
            if (g->n < 5 || ((g->n - 1) % 4) != 0) return fail(TQ_E_ARG, "Boole needs n = 1 + 4k >= 5");
            break;
        case TQ_RULE_GAUSS_LEGENDRE:
            if (g->n < 1) return fail(TQ_E_ARG, "Gauss-Legendre needs n >= 1");
            break;
        default: return fail(TQ_E_ARG, "unknown rule");
    }
    long long P;
    if (ipow_checked(g->n, g->dim, (1LL << 62), &P) != 0) return fail(TQ_E_ARG, "n^dim exceeds 2^62 points");
    return TQ_OK;
}

// Gauss-Legendre node/weight tables are needed wherever nodes or weights
// are formed (everything except unit-weight reductions).
static int check_gl_tables(const tq_grid_desc* g) {
    if (g->rule == TQ_RULE_GAUSS_LEGENDRE && (!g->gl_nodes || !g->gl_weights))
        return fail(TQ_E_NULL, "Gauss-Legendre needs node/weight tables");
    return TQ_OK;
}

// LDS for a merged (two-axis) loop table in the fused kernel; any cap of at
// least 1024 entries gives the same launch size (see kMergeMaxN).
constexpr size_t kFusedMergeLds = 40 * 1024;

inline long long fused_merge_cap(int fam, int tsize) {
    const long long cap = (long long)(kFusedMergeLds / ((size_t)tab_stride(fam) * tsize));
    return cap > 1024 ? cap : 1024;
}

static int make_geom(const tq_grid_desc* gd, GridGeom& G, long long merge_cap = 1024) {
    return plan_geom(gd->dim, gd->n, gd->rule, gd->lane_axes, gd->items_per_wave, 16384, G, merge_cap);
}

// ---------------------------------------------------------------------------
// Fused kernel
// ---------------------------------------------------------------------------
// Dynamic LDS of the fused kernel: [row weights: L*n FP64][loop table:
// S*nl T][row contributions: slots*L*n T] (the row parts only when ROWTAB),
// then for a two-axis loop, 8-byte aligned, the loop axes' table weights (2n FP64: weights or their keys)
// and nodes (2n T) from which the table is built.
template <typename T, int FAM>
__host__ __device__ inline size_t fused_pair_offset(const GridGeom& G, bool rowtab) {
    constexpr int S = TabStride<FAM>::v;
    size_t b = (size_t)S * G.nl * sizeof(T);
    if (rowtab) b += (size_t)G.L * G.n * (sizeof(double) + StateSlots<FAM>::v * sizeof(T));
    return (b + 7) & ~(size_t)7;
}
template <typename T, int FAM>
inline size_t fused_lds_bytes(const GridGeom& G, bool rowtab) {
    return fused_pair_offset<T, FAM>(G, rowtab) + (G.q == 2 ? (size_t)2 * G.n * (sizeof(double) + sizeof(T)) : 0);
}
constexpr size_t kFusedLdsMax = 48 * 1024;

// BW waves per block: 4, or 8 when a large loop table is shared by more
// waves (the LDS per block, not the registers, then bounds occupancy).
template <typename T, int FAM, bool TABLE, bool ROWTAB, int BW>
__global__ __launch_bounds__(BW * kWave) void k_grid_fused(GridArgs<T> A) {
    constexpr int NT = BW * kWave;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ Axis s_axis[TQ_MAX_DIM];
    __shared__ double s_red[BW];
    const GridGeom& g = A.g;
    const int tid = threadIdx.x;
    const int n = g.n, d = g.d, L = g.L, rule = g.rule;
    if (tid < d) s_axis[tid] = make_axis(A.domain, tid, n, rule);
    __syncthreads();
    double* s_rw = reinterpret_cast<double*>(smem_raw);
    T* s_tab = reinterpret_cast<T*>(smem_raw + (ROWTAB ? (size_t)L * n * sizeof(double) : 0));
    T* s_rt = s_tab + (size_t)TabStride<FAM>::v * g.nl;
    if (TABLE) {
        // Two loop axes: their 2n nodes and weights first, so that each of
        // the n^2 entries costs two folds and a product instead of a
        // division for its indices and two node and weight evaluations.
        const bool pair = g.q == 2;
        double* s_pw = reinterpret_cast<double*>(smem_raw + fused_pair_offset<T, FAM>(g, ROWTAB));
        T* s_px = reinterpret_cast<T*>(s_pw + 2 * n);
        if (pair) {
            for (int e = tid; e < 2 * n; e += NT) {
                const int j = e < n ? e : e - n;
                s_px[e] = (T)node_of(s_axis[e < n ? L : L + 1], j, n, rule, A.glx);
                s_pw[e] = table_weight<T, FAM>(weight_of(rule, j, n, A.glw));
            }
        } else {
            for (int j = tid; j < g.nl; j += NT)
                loop_table_entry<T, FAM>(s_tab, j, s_axis, L, 1, n, rule, A.glx, A.glw, A.p);
        }
        if (ROWTAB)
            for (int e = tid; e < L * n; e += NT) {
                const int k = e / n;
                row_table_entry<T, FAM>(s_rt, s_rw, k, e - k * n, s_axis[k], axis_c(A.p, k), n, rule, A.glx, A.glw,
                                        A.p);
            }
        __syncthreads();
        if (pair) {
            const AxisC<T> c1 = axis_c(A.p, L), c2 = axis_c(A.p, L + 1);
            // entry j = j1 n + j2, stepped by NT without a division per entry
            int j1 = tid / n, j2 = tid - j1 * n;
            const int s1 = NT / n, s2 = NT - s1 * n;
            for (int j = tid; j < g.nl; j += NT) {
                pair_table_entry<T, FAM>(s_tab, j, s_px[j1], s_px[n + j2], s_pw[j1], s_pw[n + j2], c1, c2, A.p);
                j1 += s1;
                j2 += s2;
                if (j2 >= n) {
                    j2 -= n;
                    ++j1;
                }
            }
            __syncthreads();
        }
    }
    // readfirstlane: the wave index is uniform, so the tile loop runs on SGPRs
    const long long wave = (long long)blockIdx.x * BW + __builtin_amdgcn_readfirstlane(tid >> 6);
    double acc = 0.0;
    if (wave < g.num_waves) {
        // wave < 2^31 (grid size limit): 32-bit division
        const unsigned uw = (unsigned)wave, ulw = (unsigned)g.lane_waves;
        const unsigned chunk = uw / ulw;
        const int lw = (int)(uw - chunk * ulw);
        const long long t0 = (long long)chunk * g.ipw;
        const long long t1 = (t0 + g.ipw < g.items) ? (t0 + g.ipw) : g.items;
        acc = grid_tile<T, FAM, TABLE, ROWTAB>(d, n, g.nl, g.m, L, rule, g.lane_count, A.p, s_axis, s_tab, A.glx,
                                               A.glw, lw, t0, t1, s_rt, s_rw);
    }
    const double bs = block_sum(acc, s_red);
    // The last block to finish sums every block's partial in a fixed order
    // (independent of which block is last) and applies the per-axis scale:
    // partial store, device-scope release, counter; the last block acquires
    // before reading the partials and resets the counter for the next launch.
    // Release writes this workgroup's partial back to device scope (no L2
    // invalidate, which a full fence in every workgroup would add); only
    // the last workgroup acquires.
    // Counting goes through kDoneGroups group counters: the last arrival of
    // a group acquires its group's partials and releases them on to the top
    // counter, so the final block sees every partial.
    __shared__ bool s_last;
    if (tid == 0) {
        A.partials[blockIdx.x] = bs;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned nb = gridDim.x, grp = blockIdx.x % kDoneGroups;
        const unsigned members = (nb - 1 - grp) / kDoneGroups + 1;  // blocks b < nb with b % kDoneGroups == grp
        unsigned int* gc = A.done + grp * kDoneStride;
        bool last = false;
        if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const unsigned groups = nb < (unsigned)kDoneGroups ? nb : (unsigned)kDoneGroups;
            last = __hip_atomic_fetch_add(A.done + kDoneTop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                   groups - 1;
        }
        s_last = last;
    }
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    double v = 0.0;
    for (unsigned i = tid; i < gridDim.x; i += NT) v += A.partials[i];
    v = block_sum(v, s_red);
    if (tid == 0) {
        double scale = A.c0;
        for (int k = 0; k < d; ++k)
            scale *= axis_scale((double)A.domain[2 * k], (double)A.domain[2 * k + 1], n, rule);
        const double r = v * scale;
        if (A.out) *A.out = (T)r;
        if (A.out64) *A.out64 = r;
        __hip_atomic_store(A.done + kDoneTop, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Blocks of `kernel` (`threads` per block, `shmem` bytes of dynamic LDS)
// that stay resident on the current device at once, and the device's CU
// count; 0 if unknown.  Answers are kept per thread for the last few
// (device, kernel, shmem) keys, since the runtime query costs more than a
// small launch.
static long long resident_blocks(const void* kernel, int threads, size_t shmem, int* cus_out) {
    struct Entry {
        int dev;
        const void* kernel;
        size_t shmem;
        long long blocks;
        int cus;
    };
    constexpr int kEntries = 16;
    thread_local Entry cache[kEntries] = {};
    thread_local int next = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    for (const Entry& e : cache)
        if (e.kernel == kernel && e.dev == dev && e.shmem == shmem) {
            *cus_out = e.cus;
            return e.blocks;
        }
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, shmem) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    const long long r = (long long)cus * per_cu;
    cache[next] = Entry{dev, kernel, shmem, r, cus};
    next = (next + 1) % kEntries;
    *cus_out = cus;
    return r;
}

// Workspace of the fused kernel, in block partials: at least this many, so
// that the automatic items per wave below can fill whole rounds.
constexpr long long kFusedMinBlocks = 8192;

// Launch-time model behind the automatic items per wave, fitted to the FP32
// Boole Gaussian-peak kernel on MI355X (d = 3..6, n = 17..129, items per
// wave 16..4096): a launch takes about
//   blocks x kBlockNs  (each workgroup's completion fence and counter, which
//                       the device serialises: 10112 workgroups of 16-item
//                       waves took 112 us for 3.9e7 evaluations)
//   + sum over rounds of ipw x kItemNs x (w + kItemLatency)
// where a round is one fill of the W resident waves and w its waves per
// SIMD: a wave's item (64 points) takes kItemNs (w + kItemLatency) when w
// waves share the SIMD -- latency-bound alone, issue-bound with many.
constexpr double kBlockNs = 10.0;
constexpr double kItemNs = 5.6;
constexpr double kItemLatency = 1.28;

// Relative time per item of a family and dtype (FP64 transcendentals are
// software sequences; the oscillatory cosine is a polynomial).
inline double item_scale(int fam, int tsize) {
    double s = fam == TQ_FAM_OSCILLATORY ? 1.5 : (fam == TQ_FAM_DOUBLE_GAUSSIAN ? 2.0 : 1.0);
    return tsize == 8 ? 6.0 * s : s;
}

// Automatic items per wave for a launch of `kernel`: the candidate with the
// least modelled time -- the counts that fill r = 1..16 resident rounds (each
// lane wave's items cut into floor(r W / lane_waves) chunks) and the powers
// of two up to the item count.  Keeps G as it is when the occupancy is unknown.
static void align_ipw(GridGeom& G, const void* kernel, int bw, size_t shmem, long long max_blocks, double scale) {
    int cus = 0;
    const long long W = resident_blocks(kernel, bw * kWave, shmem, &cus) * bw;
    if (W <= 0 || cus <= 0 || G.items <= 0) return;
    const double simds = 4.0 * cus;
    const double item_ns = kItemNs * scale;
    double best_t = -1.0;
    long long best_ipw = G.ipw;
    auto consider = [&](long long ipw) {
        if (ipw < 1) ipw = 1;
        if (ipw > G.items) ipw = G.items;
        const long long waves = ((G.items + ipw - 1) / ipw) * (long long)G.lane_waves;
        const long long blocks = (waves + bw - 1) / bw;
        if (blocks > max_blocks) return;
        const long long full = waves / W, rest = waves % W;
        double t = kBlockNs * (double)blocks + (double)full * (double)ipw * item_ns * ((double)W / simds + kItemLatency);
        if (rest > 0) t += (double)ipw * item_ns * ((double)rest / simds + kItemLatency);
        if (best_t < 0.0 || t < best_t) {
            best_t = t;
            best_ipw = ipw;
        }
    };
    for (long long r = 1; r <= 16; ++r) {
        long long per_lw = (r * W) / G.lane_waves;
        if (per_lw < 1) per_lw = 1;
        consider((G.items + per_lw - 1) / per_lw);
    }
    for (long long p = 16; p < G.items; p *= 2) consider(p);
    consider(G.items);
    G.ipw = best_ipw;
    G.chunks = (G.items + best_ipw - 1) / best_ipw;
    G.num_waves = G.chunks * (long long)G.lane_waves;
}

// Dynamic LDS above which 8-wave blocks share the tables.
constexpr size_t kWideBlockLds = 16 * 1024;

template <typename T, int FAM, bool TABLE, bool ROWTAB, int BW>
static void launch_fused(GridGeom& G, GridArgs<T>& A, size_t shmem, bool auto_ipw, long long max_blocks,
                         hipStream_t st) {
    const void* kern = (const void*)(k_grid_fused<T, FAM, TABLE, ROWTAB, BW>);
    if (auto_ipw) align_ipw(G, kern, BW, shmem, max_blocks, item_scale(FAM, (int)sizeof(T)));
    G.blocks = (G.num_waves + BW - 1) / BW;
    A.g = G;
    hipLaunchKernelGGL((k_grid_fused<T, FAM, TABLE, ROWTAB, BW>), dim3((unsigned)G.blocks), dim3(BW * kWave), shmem,
                       st, A);
}

template <typename T>
static int launch_grid(const tq_grid_desc* gd, const tq_integrand* f, const void* domain,
                       void* ws, size_t ws_bytes, void* out, double* out64, hipStream_t st) {
    GridGeom G;
    int rc = make_geom(gd, G, fused_merge_cap(f->family, (int)sizeof(T)));
    if (rc) return rc;
    const size_t need = kWsHeader + (size_t)G.blocks * sizeof(double);
    if (!ws || ws_bytes < need) return fail(TQ_E_WORKSPACE, "grid workspace too small");
    GridArgs<T> A;
    A.g = G;
    A.domain = (const T*)domain;
    A.glx = gd->gl_nodes;
    A.glw = gd->gl_weights;
    A.done = (unsigned int*)ws;
    A.partials = (double*)((unsigned char*)ws + kWsHeader);
    A.out = (T*)out;
    A.out64 = out64;
    A.c0 = f->c0;
    A.p = make_params<T>(*f);
    const bool auto_ipw = gd->items_per_wave <= 0;
    const long long max_blocks = (long long)((ws_bytes - kWsHeader) / sizeof(double));
    TQ_DISPATCH_FAMILY(f->family, {
        const bool rowtab = G.use_table && G.L >= 1 && (fused_lds_bytes<T, FAM>(G, true)) <= kFusedLdsMax;
        const size_t shmem = (fused_lds_bytes<T, FAM>(G, rowtab));
        const bool wide = shmem > kWideBlockLds;
        if (rowtab && wide)
            (launch_fused<T, FAM, true, true, 8>)(G, A, shmem, auto_ipw, max_blocks, st);
        else if (rowtab)
            (launch_fused<T, FAM, true, true, kWavesPerBlock>)(G, A, shmem, auto_ipw, max_blocks, st);
        else if (G.use_table && wide)
            (launch_fused<T, FAM, true, false, 8>)(G, A, shmem, auto_ipw, max_blocks, st);
        else if (G.use_table)
            (launch_fused<T, FAM, true, false, kWavesPerBlock>)(G, A, shmem, auto_ipw, max_blocks, st);
        else
            (launch_fused<T, FAM, false, false, kWavesPerBlock>)(G, A, 0, auto_ipw, max_blocks, st);
    });
    return check_launch("k_grid_fused");
}

// ---------------------------------------------------------------------------
// Point writer (user integrands)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_grid_points(const T* domain, int d, int n, int rule,
                                                        const double* glx, long long P, T* pts) {
    __shared__ Axis s_axis[kMaxDimPoints];
    for (int k = threadIdx.x; k < d; k += blockDim.x) s_axis[k] = make_axis(domain, k, n, rule);
    __syncthreads();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += stride) {
        unsigned long long rem = (unsigned long long)p;
        T* row = pts + p * d;
        for (int k = d - 1; k >= 0; --k) {
            const int j = (int)(rem % (unsigned long long)n);
            rem /= (unsigned long long)n;
            row[k] = (T)node_of(s_axis[k], j, n, rule, glx);
        }
    }
}

template <typename T>
__global__ void k_grid_h(const T* domain, int d, int n, int rule, const double* glx, T* h) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d) return;
    const Axis ax = make_axis(domain, k, n, rule);
    const T x0 = (T)node_of(ax, 0, n, rule, glx);
    const T x1 = (n > 1) ? (T)node_of(ax, 1, n, rule, glx) : x0;
    h[k] = x1 - x0;
}

// ---------------------------------------------------------------------------
// Weighted reduction of user values f[P, S]
// Lanes enumerate (inner point over the last m axes, component s); the wave
// walks the outer index with an odometer; reads are coalesced.
// ---------------------------------------------------------------------------
struct ReduceGeom {
    int d, n, m, rule;
    long long S;
    long long inner;       // n^m
    long long lane_elems;  // inner*S
    long long lane_waves;
    long long outer;       // n^(d-m)
    long long ipw;
    long long chunks;
    long long num_waves;
    long long blocks;
};

static int make_reduce_geom(const tq_grid_desc* gd, long long S, ReduceGeom& R) {
    const int d = gd->dim, n = gd->n;
    R.d = d;
    R.n = n;
    R.rule = gd->rule;
    R.S = S;
    int m = d;
    long long c = 1;
    for (int k = 1; k <= d; ++k) {
        c *= n;
        if (c * S >= 64) {
            m = k;
            break;
        }
    }
    if (m > d) m = d;
    long long inner;
    ipow_checked(n, m, 1LL << 62, &inner);
    R.m = m;
    R.inner = inner;
    R.lane_elems = inner * S;
    R.lane_waves = (R.lane_elems + kWave - 1) / kWave;
    long long outer;
    ipow_checked(n, d - m, 1LL << 62, &outer);
    R.outer = outer;
    const long long target_waves = 16384;
    long long ipw = (outer * R.lane_waves + target_waves - 1) / target_waves;
    if (ipw < 16) ipw = outer < 16 ? outer : 16;
    if (ipw < 1) ipw = 1;
    while (((outer + ipw - 1) / ipw) * R.lane_waves > (1LL << 30)) ipw *= 2;
    R.ipw = ipw;
    R.chunks = (outer + ipw - 1) / ipw;
    R.num_waves = R.chunks * R.lane_waves;
    R.blocks = (R.num_waves + kWavesPerBlock - 1) / kWavesPerBlock;
    return TQ_OK;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_grid_reduce(ReduceGeom R, const T* fv, const double* glw,
                                                        double* partials /* [blocks][S] */) {
    __shared__ double s_acc[kBlock];
    __shared__ long long s_comp[kBlock];
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    const long long wave = (long long)blockIdx.x * kWavesPerBlock + (tid >> 6);
    const int n = R.n, d = R.d, m = R.m, rule = R.rule;
    double acc = 0.0;
    long long comp = -1;
    if (wave < R.num_waves) {
        const long long lw = wave % R.lane_waves;
        const long long chunk = wave / R.lane_waves;
        const long long e = lw * kWave + lane;
        if (e < R.lane_elems) {
            comp = e % R.S;
            long long rem = e / R.S;
            double w_in = 1.0;
            for (int k = d - 1; k >= d - m; --k) {
                const int j = (int)(rem % n);
                rem /= n;
                w_in *= weight_of(rule, j, n, glw);
            }
            const long long o0 = chunk * R.ipw;
            const long long o1 = (o0 + R.ipw < R.outer) ? (o0 + R.ipw) : R.outer;
            int odig[kMaxDimPoints];
            const int no = d - m;
            {
                long long r = o0;
                for (int k = no - 1; k >= 0; --k) {
                    odig[k] = (int)(r % n);
                    r /= n;
                }
            }
            const long long stride = R.lane_elems;
            for (long long o = o0; o < o1; ++o) {
                double w = w_in;
                for (int k = 0; k < no; ++k) w *= weight_of(rule, odig[k], n, glw);
                acc = fma(w, (double)fv[o * stride + e], acc);
                for (int k = no - 1; k >= 0; --k) {
                    if (++odig[k] < n) break;
                    odig[k] = 0;
                }
            }
        }
    }
    s_acc[tid] = acc;
    s_comp[tid] = comp;
    __syncthreads();
    for (long long s = tid; s < R.S; s += blockDim.x) {
        double v = 0.0;
        for (int i = 0; i < (int)blockDim.x; ++i)
            if (s_comp[i] == s) v += s_acc[i];
        partials[(long long)blockIdx.x * R.S + s] = v;
    }
}

// Scale of the weighted reduction: Newton-Cotes uses the caller's mesh
// widths when given (hs[k]/c as _apply_composite_rule does), otherwise the
// domain's; Gauss-Legendre uses (b-a)/2 per axis.
template <typename T>
__device__ inline double reduce_scale(const T* domain, const T* hs, int d, int n, int rule, int unit) {
    double scale = 1.0;
    for (int k = 0; k < d; ++k) {
        const double a = (double)domain[2 * k], b = (double)domain[2 * k + 1];
        if (unit || rule == TQ_RULE_GAUSS_LEGENDRE) {
            scale *= 0.5 * (b - a);
        } else if (hs) {
            const double h = (double)hs[k];
            scale *= (rule == TQ_RULE_TRAPEZOID) ? h / 2.0 : (rule == TQ_RULE_SIMPSON) ? h / 3.0 : h / 22.5;
        } else {
            scale *= axis_scale(a, b, n, rule);
        }
    }
    return scale;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_grid_reduce_finalize(const double* partials, long long blocks,
                                                                 long long S, const T* domain, const T* hs,
                                                                 int unit, int d, int n, int rule, T* out) {
    __shared__ double s_red[kWavesPerBlock];
    const long long s = blockIdx.x;
    double v = 0.0;
    for (long long b = threadIdx.x; b < blocks; b += blockDim.x) v += partials[b * S + s];
    v = block_sum(v, s_red);
    if (threadIdx.x == 0) out[s] = (T)(v * reduce_scale(domain, hs, d, n, rule, unit));
}

// grad_f[p, s] = grad_out[s] * scale * W(p)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_grid_reduce_backward(const T* domain, const T* hs, int unit, int d,
                                                                 int n, int rule, const double* glw,
                                                                 const T* grad_out, long long S, long long P,
                                                                 T* grad_f) {
    const double scale = reduce_scale(domain, hs, d, n, rule, unit);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += stride) {
        double w = scale;
        if (!unit) {
            unsigned long long rem = (unsigned long long)p;
            for (int k = d - 1; k >= 0; --k) {
                const int j = (int)(rem % (unsigned long long)n);
                rem /= (unsigned long long)n;
                w *= weight_of(rule, j, n, glw);
            }
        }
        for (long long s = 0; s < S; ++s) grad_f[p * S + s] = (T)(w * (double)grad_out[s]);
    }
}

// d x_k / d a_k and d x_k / d b_k of node j on axis k.
__device__ inline void node_grad(int j, int n, int rule, const double* glx, double* da, double* db) {
    if (rule == TQ_RULE_GAUSS_LEGENDRE) {
        const double r = glx[j];
        *da = 0.5 * (1.0 - r);
        *db = 0.5 * (1.0 + r);
    } else {
        const double t = (n > 1) ? (double)j / (double)(n - 1) : 0.0;
        *da = 1.0 - t;
        *db = t;
    }
}

// Column-fixed threads over grad_points[P, d]: block size is a multiple of d.
struct Strides {
    long long s[kMaxDimPoints];
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_grid_points_backward(int d, int n, int rule, const double* glx,
                                                                 Strides strides, long long P,
                                                                 const T* gp, double* partials /* [blocks][d][2] */) {
    __shared__ double s_a[kBlock], s_b[kBlock];
    const long long bd = blockDim.x;
    const long long total = P * d;
    const long long stride = (long long)gridDim.x * bd;
    double acc_a = 0.0, acc_b = 0.0;
    const long long e0 = (long long)blockIdx.x * bd + threadIdx.x;
    const int k = (int)(e0 % d);
    const long long st = strides.s[k];
    for (long long e = e0; e < total; e += stride) {
        const long long p = e / d;
        const int j = (int)((p / st) % n);
        double da, db;
        node_grad(j, n, rule, glx, &da, &db);
        const double g = (double)gp[e];
        acc_a = fma(g, da, acc_a);
        acc_b = fma(g, db, acc_b);
    }
    s_a[threadIdx.x] = acc_a;
    s_b[threadIdx.x] = acc_b;
    __syncthreads();
    if (threadIdx.x < d) {
        double va = 0.0, vb = 0.0;
        for (long long i = threadIdx.x; i < bd; i += d) {
            va += s_a[i];
            vb += s_b[i];
        }
        partials[((long long)blockIdx.x * d + threadIdx.x) * 2] = va;
        partials[((long long)blockIdx.x * d + threadIdx.x) * 2 + 1] = vb;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_grid_points_backward_finalize(const double* partials, long long nb,
                                                                          int d, T* grad_domain) {
    __shared__ double s_red[kWavesPerBlock];
    const int e = blockIdx.x;  // (k, side)
    double v = 0.0;
    for (long long b = threadIdx.x; b < nb; b += blockDim.x) v += partials[b * d * 2 + e];
    v = block_sum(v, s_red);
    if (threadIdx.x == 0) grad_domain[e] = (T)v;
}

static void points_backward_shape(int d, long long P, int* bd, long long* nb) {
    *bd = kBlock - (kBlock % d);
    long long b = (P * d + (long long)(*bd) * 32 - 1) / ((long long)(*bd) * 32);
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    *nb = b;
}

// ---------------------------------------------------------------------------
// Allocation-free entry point: the block partials live in a file-scope
// device array instead of a caller workspace.  Each device holds its own copy
// of the array; host code reaches it only through the address that
// hipGetSymbolAddress returns with that device current (the array's name is
// a host-side stand-in, not a device address).  Every block writes its
// partial before the last block reads it; the completion counter in the
// first bytes starts at zero (static storage) and every launch leaves it zero.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)1 << 20;  // the counters and (1 MiB - kWsHeader) / 8 block partials
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device (with that device current) and kept.
static int current_scratch(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("hipGetDevice: %s", hipGetErrorString(e));
        return (int)e;
    }
    if (dev < 0 || dev >= kMaxDevices) return fail(TQ_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return e != hipSuccess ? (int)e : TQ_E_UNSUPPORTED;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TQ_OK;
}

}  // namespace tq

using namespace tq;

extern "C" {

size_t tq_grid_workspace_size(const tq_grid_desc* g) {
    if (check_grid_desc(g, TQ_MAX_DIM)) return 0;
    GridGeom G;
    if (make_geom(g, G)) return 0;
    const long long blocks = G.blocks > kFusedMinBlocks ? G.blocks : kFusedMinBlocks;
    return kWsHeader + (size_t)blocks * sizeof(double);
}

int tq_grid_integrate(const tq_grid_desc* g, const tq_integrand* f, const void* domain, void* workspace,
                      size_t workspace_bytes, void* out, double* out_f64, void* stream) {
    int rc = check_grid_desc(g, TQ_MAX_DIM);
    if (rc) return rc;
    rc = check_gl_tables(g);
    if (rc) return rc;
    rc = check_integrand(f, g->dim);
    if (rc) return rc;
    if (!domain) return fail(TQ_E_NULL, "domain is NULL");
    if (!out && !out_f64) return fail(TQ_E_NULL, "no output pointer");
    if (g->dtype == TQ_F32)
        return launch_grid<float>(g, f, domain, workspace, workspace_bytes, out, out_f64, as_stream(stream));
    return launch_grid<double>(g, f, domain, workspace, workspace_bytes, out, out_f64, as_stream(stream));
}

int tq_grid_integrate_noalloc(const tq_grid_desc* g, const tq_integrand* f, const void* domain, void* out,
                              double* out_f64, void* stream) {
    int rc = check_grid_desc(g, TQ_MAX_DIM);
    if (rc) return rc;
    rc = check_gl_tables(g);
    if (rc) return rc;
    rc = check_integrand(f, g->dim);
    if (rc) return rc;
    if (!domain) return fail(TQ_E_NULL, "domain is NULL");
    if (!out && !out_f64) return fail(TQ_E_NULL, "no output pointer");
    // the same partial count as a workspace of tq_grid_workspace_size bytes,
    // so both entry points launch the same geometry
    const size_t need = tq_grid_workspace_size(g);
    if (need == 0) return fail(TQ_E_ARG, "invalid grid geometry");
    if (need > kNoAllocScratchBytes) {
        set_error("tq_grid_integrate_noalloc: %zu bytes of block partials exceed its %zu-byte scratch; use "
                  "tq_grid_integrate with a workspace",
                  need, kNoAllocScratchBytes);
        return TQ_E_WORKSPACE;
    }
    void* ws = nullptr;
    rc = current_scratch(&ws);
    if (rc) return rc;
    if (g->dtype == TQ_F32)
        return launch_grid<float>(g, f, domain, ws, need, out, out_f64, as_stream(stream));
    return launch_grid<double>(g, f, domain, ws, need, out, out_f64, as_stream(stream));
}

int tq_grid_points(const tq_grid_desc* g, const void* domain, void* points, void* stream) {
    int rc = check_grid_desc(g, kMaxDimPoints);
    if (rc) return rc;
    rc = check_gl_tables(g);
    if (rc) return rc;
    if (!domain || !points) return fail(TQ_E_NULL, "domain/points is NULL");
    long long P;
    ipow_checked(g->n, g->dim, 1LL << 62, &P);
    long long blocks = (P + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    if (g->dtype == TQ_F32)
        hipLaunchKernelGGL((k_grid_points<float>), dim3((unsigned)blocks), dim3(kBlock), 0, as_stream(stream),
                           (const float*)domain, g->dim, g->n, g->rule, g->gl_nodes, P, (float*)points);
    else
        hipLaunchKernelGGL((k_grid_points<double>), dim3((unsigned)blocks), dim3(kBlock), 0, as_stream(stream),
                           (const double*)domain, g->dim, g->n, g->rule, g->gl_nodes, P, (double*)points);
    return check_launch("k_grid_points");
}

int tq_grid_h(const tq_grid_desc* g, const void* domain, void* h_out, void* stream) {
    int rc = check_grid_desc(g, kMaxDimPoints);
    if (rc) return rc;
    rc = check_gl_tables(g);
    if (rc) return rc;
    if (!domain || !h_out) return fail(TQ_E_NULL, "domain/h is NULL");
    const int blocks = (g->dim + 63) / 64;
    if (g->dtype == TQ_F32)
        hipLaunchKernelGGL((k_grid_h<float>), dim3(blocks), dim3(64), 0, as_stream(stream), (const float*)domain,
                           g->dim, g->n, g->rule, g->gl_nodes, (float*)h_out);
    else
        hipLaunchKernelGGL((k_grid_h<double>), dim3(blocks), dim3(64), 0, as_stream(stream), (const double*)domain,
                           g->dim, g->n, g->rule, g->gl_nodes, (double*)h_out);
    return check_launch("k_grid_h");
}

size_t tq_grid_reduce_workspace_size(const tq_grid_desc* g, int64_t S) {
    if (check_grid_desc(g, kMaxDimPoints) || S < 1) return 0;
    ReduceGeom R;
    make_reduce_geom(g, S, R);
    return (size_t)(R.blocks * S) * sizeof(double);
}

int tq_grid_reduce(const tq_grid_desc* g, const void* domain, const void* hs, int32_t unit_weights,
                   const void* fvals, int64_t S, void* workspace, size_t workspace_bytes, void* out,
                   void* stream) {
    int rc = check_grid_desc(g, kMaxDimPoints);
    if (rc) return rc;
    if (!unit_weights) {
        rc = check_gl_tables(g);
        if (rc) return rc;
    }
    if (S < 1) return fail(TQ_E_ARG, "S must be >= 1");
    if (!domain || !fvals || !out) return fail(TQ_E_NULL, "domain/fvals/out is NULL");
    ReduceGeom R;
    make_reduce_geom(g, S, R);
    if (R.S > 65535) return fail(TQ_E_UNSUPPORTED, "more than 65535 integrand components");
    const size_t need = (size_t)(R.blocks * S) * sizeof(double);
    if (!workspace || workspace_bytes < need) return fail(TQ_E_WORKSPACE, "reduce workspace too small");
    // unit weights: every point has weight 1 (the reduce kernel still runs the
    // rule's pattern, so route through the Gauss-Legendre branch with ones).
    if (unit_weights) R.rule = -1;
    hipStream_t st = as_stream(stream);
    const int unit = unit_weights ? 1 : 0;
    if (g->dtype == TQ_F32) {
        hipLaunchKernelGGL((k_grid_reduce<float>), dim3((unsigned)R.blocks), dim3(kBlock), 0, st, R,
                           (const float*)fvals, g->gl_weights, (double*)workspace);
        rc = check_launch("k_grid_reduce");
        if (rc) return rc;
        hipLaunchKernelGGL((k_grid_reduce_finalize<float>), dim3((unsigned)S), dim3(kBlock), 0, st,
                           (const double*)workspace, R.blocks, (long long)S, (const float*)domain,
                           (const float*)hs, unit, g->dim, g->n, g->rule, (float*)out);
    } else {
        hipLaunchKernelGGL((k_grid_reduce<double>), dim3((unsigned)R.blocks), dim3(kBlock), 0, st, R,
                           (const double*)fvals, g->gl_weights, (double*)workspace);
        rc = check_launch("k_grid_reduce");
        if (rc) return rc;
        hipLaunchKernelGGL((k_grid_reduce_finalize<double>), dim3((unsigned)S), dim3(kBlock), 0, st,
                           (const double*)workspace, R.blocks, (long long)S, (const double*)domain,
                           (const double*)hs, unit, g->dim, g->n, g->rule, (double*)out);
    }
    return check_launch("k_grid_reduce_finalize");
}

int tq_grid_reduce_backward(const tq_grid_desc* g, const void* domain, const void* hs, int32_t unit_weights,
                            const void* grad_out, int64_t S, void* grad_f, void* stream) {
    int rc = check_grid_desc(g, kMaxDimPoints);
    if (rc) return rc;
    if (!unit_weights) {
        rc = check_gl_tables(g);
        if (rc) return rc;
    }
    if (S < 1) return fail(TQ_E_ARG, "S must be >= 1");
    if (!domain || !grad_out || !grad_f) return fail(TQ_E_NULL, "domain/grad_out/grad_f is NULL");
    long long P;
    ipow_checked(g->n, g->dim, 1LL << 62, &P);
    long long blocks = (P + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    const int unit = unit_weights ? 1 : 0;
    if (g->dtype == TQ_F32)
        hipLaunchKernelGGL((k_grid_reduce_backward<float>), dim3((unsigned)blocks), dim3(kBlock), 0,
                           as_stream(stream), (const float*)domain, (const float*)hs, unit, g->dim, g->n, g->rule,
                           g->gl_weights, (const float*)grad_out, (long long)S, P, (float*)grad_f);
    else
        hipLaunchKernelGGL((k_grid_reduce_backward<double>), dim3((unsigned)blocks), dim3(kBlock), 0,
                           as_stream(stream), (const double*)domain, (const double*)hs, unit, g->dim, g->n,
                           g->rule, g->gl_weights, (const double*)grad_out, (long long)S, P, (double*)grad_f);
    return check_launch("k_grid_reduce_backward");
}

size_t tq_grid_points_backward_workspace_size(const tq_grid_desc* g) {
    if (check_grid_desc(g, kMaxDimPoints)) return 0;
    long long P;
    ipow_checked(g->n, g->dim, 1LL << 62, &P);
    int bd;
    long long nb;
    points_backward_shape(g->dim, P, &bd, &nb);
    return (size_t)(nb * g->dim * 2) * sizeof(double);
}

int tq_grid_points_backward(const tq_grid_desc* g, const void* domain, const void* grad_points, void* workspace,
                            size_t workspace_bytes, void* grad_domain, void* stream) {
    int rc = check_grid_desc(g, kMaxDimPoints);
    if (rc) return rc;
    rc = check_gl_tables(g);
    if (rc) return rc;
    if (!domain || !grad_points || !grad_domain) return fail(TQ_E_NULL, "domain/grad_points/grad_domain is NULL");
    long long P;
    ipow_checked(g->n, g->dim, 1LL << 62, &P);
    int bd;
    long long nb;
    points_backward_shape(g->dim, P, &bd, &nb);
    const size_t need = (size_t)(nb * g->dim * 2) * sizeof(double);
    if (!workspace || workspace_bytes < need) return fail(TQ_E_WORKSPACE, "points backward workspace too small");
    Strides strides;
    long long st = 1;
    for (int k = g->dim - 1; k >= 0; --k) {
        strides.s[k] = st;
        st *= g->n;
    }
    hipStream_t s = as_stream(stream);
    if (g->dtype == TQ_F32) {
        hipLaunchKernelGGL((k_grid_points_backward<float>), dim3((unsigned)nb), dim3(bd), 0, s, g->dim, g->n,
                           g->rule, g->gl_nodes, strides, P, (const float*)grad_points,
                           (double*)workspace);
        rc = check_launch("k_grid_points_backward");
        if (rc) return rc;
        hipLaunchKernelGGL((k_grid_points_backward_finalize<float>), dim3(g->dim * 2), dim3(kBlock), 0, s,
                           (const double*)workspace, nb, g->dim, (float*)grad_domain);
    } else {
        hipLaunchKernelGGL((k_grid_points_backward<double>), dim3((unsigned)nb), dim3(bd), 0, s, g->dim, g->n,
                           g->rule, g->gl_nodes, strides, P, (const double*)grad_points,
                           (double*)workspace);
        rc = check_launch("k_grid_points_backward");
        if (rc) return rc;
        hipLaunchKernelGGL((k_grid_points_backward_finalize<double>), dim3(g->dim * 2), dim3(kBlock), 0, s,
                           (const double*)workspace, nb, g->dim, (double*)grad_domain);
    }
    return check_launch("k_grid_points_backward_finalize");
}

}  // extern "C"
