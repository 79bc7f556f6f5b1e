// VEGAS+ on the device: adaptive map, stratification and the fused
// iteration body, with all state resident in caller-owned buffers.
//
// Iteration pipeline (no host synchronisation inside):
//   k_vegas_nh      nh = max(2, floor(dh * n_exp)) per cube + block scan
//   k_vegas_scan    scan of the block totals
//   k_vegas_offsets cube offsets, total, cube of the first sample of every
//                   1024-sample block (so blocks find their cubes in LDS)
//   k_vegas_sample  y, map, Jacobian, integrand, jf, map histograms
//   k_vegas_strat   per-cube JF/JF2 (segmented, fixed order), ih, sigma^2, dh
//   k_vegas_strat_finalize   iteration result and dh normaliser
//   k_vegas_map_reduce + k_vegas_map_smooth + k_vegas_map_edges   update_map
// Map histograms that fit in LDS are exact per block (integer atomics on a
// per-bin fixed-point sum relative to the bin's largest exponent, see
// HistSlots) and reduced over blocks in block order: the whole iteration is
// deterministic.  Larger maps aggregate each wave's equal bins (sorted keys,
// fixed-order scan) and add them with FP64 atomics.  Integer-valued and rounding-sensitive steps
// (bin ids, y, nh) are written operation by operation as the reference does
// them, with floating-point contraction disabled in this file.
#pragma clang fp contract(off)
#include "tq_host.hpp"

namespace tq {

constexpr int kSB = TQ_VEGAS_BLOCK_SAMPLES;
constexpr int kSPT = kSB / kBlock;
constexpr int kScanItems = 8;
constexpr int kScanBlockCubes = kBlock * kScanItems;
constexpr int kHistLds = 768;         // dim*n_intervals handled with LDS histograms
constexpr int kMapLds = 1024;         // doubles of map tables staged in LDS
constexpr int kOffLds = kSB / 2 + 8;  // cube offsets of one sample block
constexpr size_t kMapUpdateLds = 48 * 1024;  // largest update_map scratch kept in LDS
// Histogram keys are (bin << 6 | lane) in 32 bits: bins up to 2^26 - 2, the
// largest value marks "no bin".
constexpr unsigned kNoBin = (1u << 26) - 1;

static int check_state(const tq_vegas_state* st) {
    if (!st) return fail(TQ_E_NULL, "vegas state is NULL");
    if (st->dim < 1 || st->dim > TQ_MAX_DIM) return fail(TQ_E_ARG, "vegas dim out of range (1..16)");
    if (st->n_intervals < 2) return fail(TQ_E_ARG, "n_intervals must be >= 2");
    if (st->n_intervals >= (int)kNoBin) return fail(TQ_E_ARG, "n_intervals must be < 2^26 - 1");
    if (st->n_strat < 1 || st->n_cubes < 1) return fail(TQ_E_ARG, "n_strat/n_cubes must be >= 1");
    if (st->sample_dtype != TQ_F32 && st->sample_dtype != TQ_F64) return fail(TQ_E_ARG, "bad sample dtype");
    if (st->max_blocks < 4) return fail(TQ_E_ARG, "max_blocks must be >= 4");
    if (!st->x_edges || !st->dx_edges || !st->weights || !st->counts || !st->dh || !st->jf_cube ||
        !st->jf2_cube || !st->nh || !st->offsets || !st->chunk_cube || !st->jf || !st->hist_w || !st->hist_c ||
        !st->block_sums || !st->scan_sums || !st->iter_stats || !st->flags)
        return fail(TQ_E_NULL, "vegas state has a NULL buffer");
    return TQ_OK;
}

// ---------------------------------------------------------------------------
// init
// ---------------------------------------------------------------------------
__global__ void k_vegas_init(tq_vegas_state st) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ni = st.n_intervals;
    const double step = 1.0 / (double)ni;
    for (long long i = i0; i < (long long)st.dim * (ni + 1); i += stride) {
        const int j = (int)(i % (ni + 1));
        st.x_edges[i] = (j == ni) ? 1.0 : (double)j * step;
    }
    for (long long i = i0; i < (long long)st.dim * ni; i += stride) {
        st.dx_edges[i] = 1.0 / (double)ni;
        st.weights[i] = 0.0;
        st.counts[i] = 0.0;
    }
    const double dh0 = 1.0 / (double)st.n_cubes;
    for (long long h = i0; h < st.n_cubes; h += stride) st.dh[h] = dh0;
    if (i0 < 4) st.flags[i0] = 0;
}

// ---------------------------------------------------------------------------
// get_NH + scan
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_vegas_nh(tq_vegas_state st, double n_exp, int normalize_it,
                                                     const int64_t* nh_in) {
    __shared__ long long s_scan[kBlock];
    const int tid = threadIdx.x;
    double dsum = 0.0;
    bool norm = false;
    if (normalize_it >= 0) {
        dsum = st.iter_stats[4 * normalize_it + 3];
        norm = (dsum != 0.0);
    }
    const long long base = (long long)blockIdx.x * kScanBlockCubes + (long long)tid * kScanItems;
    long long local[kScanItems];
    long long tsum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const long long h = base + i;
        long long v = 0;
        if (h < st.n_cubes) {
            if (nh_in) {
                v = nh_in[h];
                if (v < 0) v = 0;
            } else {
                double dh = st.dh[h];
                if (norm) {
                    dh = dh / dsum;
                    st.dh[h] = dh;
                }
                double x = floor(dh * n_exp);
                if (!(x >= 2.0)) x = 2.0;  // clip(., 2) ; NaN -> 2 keeps indices in range
                if (x > 4.0e18) x = 4.0e18;
                v = (long long)x;
            }
            st.nh[h] = v;
        }
        local[i] = tsum;
        tsum += v;
    }
    s_scan[tid] = tsum;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const long long t = (tid >= off) ? s_scan[tid - off] : 0;
        __syncthreads();
        s_scan[tid] += t;
        __syncthreads();
    }
    const long long excl = s_scan[tid] - tsum;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const long long h = base + i;
        if (h < st.n_cubes) st.offsets[h] = excl + local[i];
    }
    if (tid == kBlock - 1) st.scan_sums[blockIdx.x] = s_scan[tid];
}

__global__ __launch_bounds__(1024) void k_vegas_scan(int64_t* sums, int nb) {
    __shared__ long long s[1024];
    const int tid = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = tid * per;
    long long loc = 0;
    for (int i = 0; i < per; ++i)
        if (b0 + i < nb) loc += sums[b0 + i];
    s[tid] = loc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const long long t = (tid >= off) ? s[tid - off] : 0;
        __syncthreads();
        s[tid] += t;
        __syncthreads();
    }
    long long run = s[tid] - loc;
    for (int i = 0; i < per; ++i) {
        if (b0 + i < nb) {
            const long long v = sums[b0 + i];
            sums[b0 + i] = run;
            run += v;
        }
    }
    if (tid == 1023) sums[nb] = s[1023];
}

// raw != 0: scan_sums holds k_vegas_nh's block sums unscanned (at most kBlock
// of them) and each block adds the ones before its cubes itself (integer
// sums: any order gives the same value), which saves the k_vegas_scan launch.
__global__ __launch_bounds__(kBlock) void k_vegas_offsets(tq_vegas_state st, int it, int nb_scan, int raw) {
    __shared__ long long s_part[kWavesPerBlock];
    const long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long prefix = 0, total = 0;
    if (raw) {
        // this block's cubes all lie in nh block j0 (kBlock divides kScanBlockCubes)
        const long long j0 = ((long long)blockIdx.x * kBlock) / kScanBlockCubes;
        const long long lim = blockIdx.x == 0 ? nb_scan : j0;
        long long v = (long long)threadIdx.x < lim ? st.scan_sums[threadIdx.x] : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
        if ((threadIdx.x & (kWave - 1)) == 0) s_part[threadIdx.x >> 6] = v;
        __syncthreads();
        long long r = 0;
        for (int w = 0; w < kWavesPerBlock; ++w) r += s_part[w];
        if (blockIdx.x == 0) total = r;
        else prefix = r;
    } else if (h < st.n_cubes) {
        prefix = st.scan_sums[h / kScanBlockCubes];
    }
    if (h < st.n_cubes) {
        const long long off = st.offsets[h] + prefix;
        st.offsets[h] = off;
        const long long n = st.nh[h];
        if (n > 0) {
            const long long c0 = (off + kSB - 1) / kSB;
            const long long c1 = (off + n - 1) / kSB;
            for (long long c = c0; c <= c1 && c <= st.max_blocks; ++c) st.chunk_cube[c] = h;
        }
    }
    if (h == 0) {
        if (!raw) total = st.scan_sums[nb_scan];
        st.offsets[st.n_cubes] = total;
        st.iter_stats[4 * it + 2] = (double)total;
        if (total > st.max_samples || (total + kSB - 1) / kSB > st.max_blocks) st.flags[3] = 1;
    }
}

// ---------------------------------------------------------------------------
// histograms
// ---------------------------------------------------------------------------
// Per-bin sums of the wave's (bin, value) pairs.  The 32-bit keys
// (bin << 6 | lane) are sorted with a bitonic network over the 64 lanes,
// each lane then fetches the value of the lane its sorted key came from,
// and a segmented inclusive scan (DPP, no LDS) sums each run of equal bins;
// the last lane of a run gets head = true with the run's bin, value sum and
// lane count.
// Lanes with id < 0 take no part.  The summation order depends only on the
// lanes' bins, so the sums are reproducible (no floating-point atomics).
struct BinSum {
    long long bin;
    double sum;
    int count;
    bool head;
};

// x of lane (lane ^ S).  Strides 1, 2, 4 and 8 stay inside 16-lane rows and
// use DPP moves (VALU, no LDS round trip): quad_perm [1,0,3,2] and
// [2,3,0,1]; row_half_mirror (lane ^ 7) then quad_perm [3,2,1,0] (^ 3) for
// 4; row_ror:8 for 8.  16 and 32 cross rows and go through ds_bpermute.
template <int S>
__device__ inline unsigned wave_xor(unsigned x) {
    if constexpr (S == 1) {
        return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
    } else if constexpr (S == 2) {
        return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
    } else if constexpr (S == 4) {
        const unsigned y = __builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false);
        return __builtin_amdgcn_mov_dpp(y, 0x1B, 0xF, 0xF, false);
    } else if constexpr (S == 8) {
        return __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);
    } else {
        // lane ^ S stays inside the wave: ds_bpermute with the byte address
        // of the partner lane (no width clamp needed)
        const int lane = threadIdx.x & (kWave - 1);
        return (unsigned)__builtin_amdgcn_ds_bpermute((lane ^ S) << 2, (int)x);
    }
}

// One compare-exchange stage of the bitonic network (merge size SIZE,
// partner distance STRIDE): ascending overall.
template <int SIZE, int STRIDE>
__device__ inline unsigned bitonic_stage(unsigned key, int lane) {
    const unsigned other = wave_xor<STRIDE>(key);
    const bool up = (lane & SIZE) == 0;
    const bool lower = (lane & STRIDE) == 0;
    // branch-free: min and max of the pair, then one select on the lane's
    // (compile-time per stage) role -- the DPP move, v_min, v_max, v_cndmask
    const unsigned mn = __builtin_elementwise_min(key, other);
    const unsigned mx = __builtin_elementwise_max(key, other);
    return (lower == up) ? mn : mx;
}

template <int SIZE, int STRIDE>
__device__ inline unsigned bitonic_merge(unsigned key, int lane) {
    key = bitonic_stage<SIZE, STRIDE>(key, lane);
    if constexpr (STRIDE > 1) key = bitonic_merge<SIZE, STRIDE / 2>(key, lane);
    return key;
}

template <int SIZE>
__device__ inline unsigned bitonic_sort(unsigned key, int lane) {
    if constexpr (SIZE > 2) key = bitonic_sort<SIZE / 2>(key, lane);
    return bitonic_merge<SIZE, SIZE / 2>(key, lane);
}

// One step of the run scan: take (bin, sum, count) from the lane that the
// DPP control CTRL selects (rows outside ROW_MASK and lanes whose source
// leaves the row read nothing) and add it when the bins are equal.
template <int CTRL, int ROW_MASK>
__device__ inline void run_scan_step(unsigned bin, double& sum, int& cnt) {
    const unsigned pb = (unsigned)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)bin, CTRL, ROW_MASK, 0xF, false);
    const int pc = __builtin_amdgcn_update_dpp(0, cnt, CTRL, ROW_MASK, 0xF, false);
    union {
        double d;
        int i[2];
    } mine, other;
    mine.d = sum;
    other.i[0] = __builtin_amdgcn_update_dpp(0, mine.i[0], CTRL, ROW_MASK, 0xF, false);
    other.i[1] = __builtin_amdgcn_update_dpp(0, mine.i[1], CTRL, ROW_MASK, 0xF, false);
    if (pb == bin) {
        sum += other.d;
        cnt += pc;
    }
}

__device__ inline BinSum wave_bin_sums(int id, double v) {
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned key = bitonic_sort<kWave>(((id >= 0 ? (unsigned)id : kNoBin) << 6) | (unsigned)lane, lane);
    const unsigned bin = key >> 6;
    const double mine = id >= 0 ? v : 0.0;
    double sum = __shfl(mine, (int)(key & (kWave - 1)), kWave);
    int cnt = 1;
    // Inclusive scan over runs of equal bins, which are contiguous because
    // the keys are sorted: Hillis-Steele inside each 16-lane row (row_shr 1,
    // 2, 4, 8), then row r+1 takes lane 16r+15 of rows 0 and 2 (row_bcast15)
    // and rows 2 and 3 take lane 31 (row_bcast31), each added only when the
    // bin continues.
    run_scan_step<0x111, 0xF>(bin, sum, cnt);
    run_scan_step<0x112, 0xF>(bin, sum, cnt);
    run_scan_step<0x114, 0xF>(bin, sum, cnt);
    run_scan_step<0x118, 0xF>(bin, sum, cnt);
    run_scan_step<0x142, 0xA>(bin, sum, cnt);
    run_scan_step<0x143, 0xC>(bin, sum, cnt);
    const unsigned nb = (unsigned)__shfl_down((int)bin, 1, kWave);
    BinSum r;
    r.bin = (long long)bin;
    r.sum = sum;
    r.count = cnt;
    r.head = (lane == kWave - 1 || nb != bin) && bin != kNoBin;
    return r;
}

// Exact block histograms (maps that fit in LDS).  Each bin keeps a count, the
// largest binary exponent among its values and a 64-bit fixed-point sum of
// the values relative to that exponent.  All three are updated with integer
// LDS atomics, which commute, so the block's sums do not depend on the order
// in which lanes and waves arrive: results are reproducible run to run.
//
// A value v >= 0 (a square) is m 2^(e - 1075) with m < 2^53 and e its biased
// exponent field (e = 1 with no hidden bit for subnormals).  Pass 1 takes
// the largest e per bin (E); pass 2 adds floor(m 2^(e - E)) in units of
// 2^(E - 1075).  A bin receives at most kSB values per block (one per sample
// and axis), so the sum stays below 2^63; each value loses under one unit,
// i.e. under 2^-52 of the bin's largest value.  Non-finite values set
// E = 0x7ff and mark the sum (1: +inf seen, 2: NaN seen) instead.
constexpr uint32_t kExpNonFinite = 0x7ffu;

struct HistSlots {
    unsigned long long* sum;  // [W]
    uint32_t* exp;            // [W]
    uint32_t* cnt;            // [W]
};

__device__ inline size_t hist_slot_bytes(int W) { return (size_t)W * (8 + 4 + 4); }

__device__ inline HistSlots hist_slots(unsigned char* base, int W) {
    HistSlots h;
    h.sum = reinterpret_cast<unsigned long long*>(base);
    h.exp = reinterpret_cast<uint32_t*>(base + (size_t)W * 8);
    h.cnt = reinterpret_cast<uint32_t*>(base + (size_t)W * 12);
    return h;
}

__device__ inline void hist_zero(const HistSlots& h, int W) {
    for (int i = threadIdx.x; i < W; i += blockDim.x) {
        h.sum[i] = 0ull;
        h.exp[i] = 0u;
        h.cnt[i] = 0u;
    }
}

// biased exponent field and integer significand of v >= 0
__device__ inline uint32_t f64_exp(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const uint32_t e = (uint32_t)(b >> 52) & 0x7ffu;
    return e == 0u ? 1u : e;
}
__device__ inline unsigned long long f64_sig(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned long long f = b & ((1ull << 52) - 1ull);
    return ((b >> 52) & 0x7ffull) ? (f | (1ull << 52)) : f;
}

// pass 1 for one (bin, value): count and largest exponent
__device__ inline void hist_pass1(const HistSlots& h, int bin, double v) {
    atomicAdd(&h.cnt[bin], 1u);
    if (v != 0.0) atomicMax(&h.exp[bin], f64_exp(v));
}

// pass 2 for one (bin, value), after every pass 1 of the block
__device__ inline void hist_pass2(const HistSlots& h, int bin, double v) {
    if (v == 0.0) return;
    const uint32_t E = h.exp[bin];
    const uint32_t e = f64_exp(v);
    if (e == kExpNonFinite) {
        atomicOr(&h.sum[bin], (v != v) ? 2ull : 1ull);
        return;
    }
    if (E == kExpNonFinite) return;
    const uint32_t sh = E - e;
    if (sh < 64u) atomicAdd(&h.sum[bin], f64_sig(v) >> sh);
}

// the bin's FP64 value after pass 2
__device__ inline double hist_value(const HistSlots& h, int bin) {
    const uint32_t E = h.exp[bin];
    const unsigned long long s = h.sum[bin];
    if (E == kExpNonFinite) return (s & 2ull) ? __longlong_as_double(0x7ff8000000000000ll)
                                              : __longlong_as_double(0x7ff0000000000000ll);
    return ldexp((double)s, (int)E - 1075);
}

// Same aggregation, written with FP64 atomics to the global accumulators
// (maps too large for LDS; the summation order across waves is then not fixed).
__device__ inline void wave_hist_add_global(double* w, double* c, int k, int ni, int id, double v) {
    if (!__ballot(id >= 0)) return;
    const BinSum r = wave_bin_sums(id, v);
    if (r.head) {
        atomicAdd(&w[(long long)k * ni + r.bin], r.sum);
        atomicAdd(&c[(long long)k * ni + r.bin], (double)r.count);
    }
}

// ---------------------------------------------------------------------------
// sample kernel
// ---------------------------------------------------------------------------
enum { MODE_FUSED = 0, MODE_POINTS = 1, MODE_Y = 2 };

template <typename T>
struct SampleArgs {
    tq_vegas_state st;
    const T* domain;
    uint64_t seed, offset;
    int warmup;
    long long n_warm;
    int hist;  // accumulate map histograms
    int it;
    T* x_out;
    double* jac_out;
    int32_t* ids_out;
    T* y_out;
    Params<T> p;
    const T* u_in;  // optional caller uniforms [samples][dim] (MODE_POINTS)
};

// Coordinate k of sample s from the Philox stream: FP32 uses word k%4 of
// block k/4, FP64 words 2(k%2), 2(k%2)+1 of block k/2.  `blk` carries the
// block between the coordinates that share it; callers walk k upwards from 0
// with a compile-time k, so the block is generated once per 4 (2) coordinates.
template <typename T>
__device__ inline T uniform_coord(uint64_t seed, uint64_t s, int k, u32x4& blk);
template <>
__device__ inline float uniform_coord<float>(uint64_t seed, uint64_t s, int k, u32x4& blk) {
    if ((k & 3) == 0) blk = philox_block(seed, s, (uint32_t)(k >> 2));
    return u01f(blk.v[k & 3]);
}
template <>
__device__ inline double uniform_coord<double>(uint64_t seed, uint64_t s, int k, u32x4& blk) {
    if ((k & 1) == 0) blk = philox_block(seed, s, (uint32_t)(k >> 1));
    return (k & 1) ? u01d(blk.v[2], blk.v[3]) : u01d(blk.v[0], blk.v[1]);
}

// MAXD: compile-time bound on the dimension; the fused launches use 8 for
// d <= 8 so the unrolled per-axis work is not predicated over 16 axes.
template <typename T, int FAM, int MODE, bool LDS_HIST, int MAXD = TQ_MAX_DIM>
__global__ __launch_bounds__(kBlock) void k_vegas_sample(SampleArgs<T> A) {
    static_assert(MAXD % 2 == 0, "bin ids are packed in pairs");
    using F = Family<FAM, T>;
    __shared__ int s_off[kOffLds];  // cube offsets relative to the block's first sample, clamped
    __shared__ double s_map[kMapLds];
    // exact block histogram in dynamic LDS sized to the map (hist_slot_bytes)
    extern __shared__ __align__(16) unsigned char s_hist_raw[];
    __shared__ long long s_hrange[2];
    const tq_vegas_state& st = A.st;
    const int tid = threadIdx.x;
    const int d = st.dim, ni = st.n_intervals;
    const int W = d * ni;
    const HistSlots hs = hist_slots(s_hist_raw, W);
    // the block's histogram values and packed bin ids, kept for pass 2
    double r_jf2[kSPT];
    uint32_t r_ids[kSPT][(MAXD + 1) / 2];

    long long total;
    if (A.warmup) {
        total = A.n_warm;
    } else {
        total = st.offsets[st.n_cubes];
        if (total > st.max_samples) total = st.max_samples;
    }
    const long long s0 = (long long)blockIdx.x * kSB;
    if (s0 >= total) return;
    const long long nblocks = (total + kSB - 1) / kSB;

    // map tables
    const bool map_lds = (d * (2 * ni + 1) <= kMapLds);
    const double* xe = st.x_edges;
    const double* dxe = st.dx_edges;
    if (MODE != MODE_Y && map_lds) {
        for (int i = tid; i < d * (ni + 1); i += kBlock) s_map[i] = st.x_edges[i];
        for (int i = tid; i < W; i += kBlock) s_map[d * (ni + 1) + i] = st.dx_edges[i];
        xe = s_map;
        dxe = s_map + d * (ni + 1);
    }
    if (MODE == MODE_FUSED && LDS_HIST && A.hist) hist_zero(hs, W);
    // cube range of this block
    long long h_lo = 0, n_off = 0;
    bool off_lds = false;
    if (!A.warmup) {
        if (tid == 0) {
            s_hrange[0] = st.chunk_cube[blockIdx.x];
            s_hrange[1] = (blockIdx.x + 1 < nblocks) ? st.chunk_cube[blockIdx.x + 1] : (st.n_cubes - 1);
        }
        __syncthreads();
        h_lo = s_hrange[0];
        const long long h_hi = s_hrange[1];
        n_off = h_hi - h_lo + 2;  // offsets h_lo .. h_hi+1
        off_lds = (n_off <= kOffLds);
        // relative offsets clamped to [-1, kSB + 1]: the search compares them
        // with s - s0 in [0, kSB), so the clamp changes no comparison
        if (off_lds)
            for (long long i = tid; i < n_off; i += kBlock) {
                long long r = st.offsets[h_lo + i] - s0;
                r = r < -1 ? -1 : (r > kSB + 1 ? kSB + 1 : r);
                s_off[i] = (int)r;
            }
    }
    __syncthreads();

    // domain transform (vegas.py:104-112)
    T start[MAXD], size[MAXD];
    T vol = T(1);
#pragma unroll
    for (int k = 0; k < MAXD; ++k) {
        if (k < d && MODE != MODE_Y) {
            start[k] = A.domain[2 * k];
            size[k] = A.domain[2 * k + 1] - A.domain[2 * k];
            vol = vol * size[k];
        } else {
            start[k] = T(0);
            size[k] = T(0);
        }
    }
    const T tns = (T)st.n_strat;
    const T tni = (T)ni;
    const T ymax = (T)0.999999;
    // Cube-index digits: division by n_strat as a multiply-high when every
    // cube index is below 2^31 (instead of a 64-bit division per axis).
    // With l = ceil(log2 n_strat) and M = ceil(2^(31+l) / n_strat) < 2^32,
    // floor(r M / 2^(31+l)) = floor(r / n_strat) for all r < 2^31: the
    // excess r (M n_strat - 2^(31+l)) / (n_strat 2^(31+l)) is below
    // r / 2^(31+l) < 2^-l <= 1 / n_strat.
    const bool fast_div = st.n_cubes <= 0x7fffffffLL;
    uint32_t div_m = 1u;
    int div_sh = 0;
    if (st.n_strat > 1) {
        const int l = 32 - __builtin_clz((unsigned)(st.n_strat - 1));
        div_sh = l - 1;
        div_m = (uint32_t)(((1ull << (31 + l)) + (unsigned long long)st.n_strat - 1ull) /
                           (unsigned long long)st.n_strat);
    }

#pragma unroll
    for (int j = 0; j < kSPT; ++j) {
        const long long s = s0 + (long long)j * kBlock + tid;
        const bool active = s < total;
        long long h = 0;
        if (active && !A.warmup) {
            if (off_lds) {
                const int sr = (int)(s - s0);
                int lo = 0, hi = (int)n_off - 1;  // find last i with s_off[i] <= s - s0
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= sr) lo = mid;
                    else hi = mid - 1;
                }
                h = h_lo + lo;
            } else {
                long long lo = h_lo, hi = st.n_cubes - 1;
                while (lo < hi) {
                    const long long mid = (lo + hi + 1) >> 1;
                    if (st.offsets[mid] <= s) lo = mid;
                    else hi = mid - 1;
                }
                h = lo;
            }
        }
        T x[MAXD];
        int ids[MAXD];
        double jac = 1.0;
        long long rem = h;
        u32x4 blk;
#pragma unroll
        for (int k = 0; k < MAXD; ++k) {
            ids[k] = -1;
            x[k] = T(0);
            if (k < d) {
                const T u = (MODE != MODE_FUSED && A.u_in)
                                ? (active ? A.u_in[s * d + k] : T(0))
                                : uniform_coord<T>(A.seed, A.offset + (uint64_t)(active ? s : 0), k, blk);
                T y;
                if (A.warmup) {
                    y = u * ymax;
                } else {
                    long long dig;
                    if (fast_div) {
                        const uint32_t r32 = (uint32_t)rem;
                        const uint32_t q = st.n_strat > 1 ? (__umulhi(r32, div_m) >> div_sh) : r32;
                        dig = (long long)(r32 - q * (uint32_t)st.n_strat);
                        rem = (long long)q;
                    } else {
                        dig = rem % st.n_strat;
                        rem /= st.n_strat;
                    }
                    y = ((T)dig + u) / tns;
                    if (y >= T(1)) y = ymax;
                }
                if (MODE == MODE_Y) {
                    if (active) A.y_out[s * d + k] = y;
                    continue;
                }
                if (MODE == MODE_POINTS && A.y_out && active) A.y_out[s * d + k] = y;
                const T yn = y * tni;
                const T fl = floor(yn);
                int id = (int)fl;
                id = id < 0 ? 0 : (id >= ni ? ni - 1 : id);
                const T off = yn - fl;
                const double xm = xe[k * (ni + 1) + id] + dxe[k * ni + id] * (double)off;
                jac *= (double)ni * dxe[k * ni + id];
                x[k] = (T)xm * size[k] + start[k];
                ids[k] = id;
            }
        }
        if (MODE == MODE_Y) continue;
        if (MODE == MODE_POINTS) {
            if (active) {
                for (int k = 0; k < d; ++k) {
                    A.x_out[s * d + k] = x[k];
                    A.ids_out[s * d + k] = ids[k];
                }
                A.jac_out[s] = jac;
            }
            continue;
        }
        // fused: integrand
        T fv = T(0);
        if (active) fv = (T)(A.p.c0 * eval_point<FAM, T, MAXD>(x, d, A.p)) * vol;
        const double jf = (double)fv * jac;
        if (active && !A.warmup) st.jf[s] = jf;
        if (A.hist) {
            const double jf2 = jf * jf;
            // unrolled with a uniform bound check, so ids[] stays in registers
            if (LDS_HIST) {
                r_jf2[j] = active ? jf2 : 0.0;
#pragma unroll
                for (int k = 0; k < MAXD; k += 2) {
                    const uint32_t lo = (k < d) ? (uint32_t)ids[k] : 0u;
                    const uint32_t hi = (k + 1 < d) ? (uint32_t)ids[k + 1] : 0u;
                    r_ids[j][k >> 1] = lo | (hi << 16);
                }
                if (active) {
#pragma unroll
                    for (int k = 0; k < MAXD; ++k)
                        if (k < d) hist_pass1(hs, k * ni + ids[k], jf2);
                }
            } else {
#pragma unroll
                for (int k = 0; k < MAXD; ++k) {
                    if (k < d) {
                        const int id = active ? ids[k] : -1;
                        wave_hist_add_global(st.weights, st.counts, k, ni, id, jf2);
                    }
                }
            }
        }
    }
    if (MODE == MODE_FUSED && LDS_HIST && A.hist) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSPT; ++j) {
            if (r_jf2[j] == 0.0) continue;
#pragma unroll
            for (int k = 0; k < MAXD; ++k)
                if (k < d) hist_pass2(hs, k * ni + (int)((r_ids[j][k >> 1] >> (16 * (k & 1))) & 0xffffu), r_jf2[j]);
        }
        __syncthreads();
        for (int e = tid; e < W; e += kBlock) {
            st.hist_w[(long long)blockIdx.x * W + e] = hist_value(hs, e);
            st.hist_c[(long long)blockIdx.x * W + e] = hs.cnt[e];
        }
    }
}

// user integrand values -> jf and histograms
template <typename T, bool LDS_HIST>
__global__ __launch_bounds__(kBlock) void k_vegas_consume(tq_vegas_state st, const T* fv, const T* domain,
                                                          const double* jacs, const int32_t* idsg, int warmup,
                                                          long long n_warm, int hist) {
    __shared__ unsigned long long s_sum[LDS_HIST ? kHistLds : 1];
    __shared__ uint32_t s_exp[LDS_HIST ? kHistLds : 1];
    __shared__ uint32_t s_cnt[LDS_HIST ? kHistLds : 1];
    const HistSlots hs{s_sum, s_exp, s_cnt};
    const int tid = threadIdx.x;
    const int d = st.dim, ni = st.n_intervals, W = d * ni;
    long long total = warmup ? n_warm : st.offsets[st.n_cubes];
    if (!warmup && total > st.max_samples) total = st.max_samples;
    const long long s0 = (long long)blockIdx.x * kSB;
    if (s0 >= total) return;
    if (LDS_HIST && hist) hist_zero(hs, W);
    T vol = T(1);
    for (int k = 0; k < d; ++k) vol = vol * (domain[2 * k + 1] - domain[2 * k]);
    __syncthreads();
    auto jf_of = [&](long long s) { return (double)(fv[s] * vol) * jacs[s]; };
    auto id_of = [&](long long s, int k) {
        const int id = idsg[s * d + k];
        return id >= ni ? ni - 1 : id;
    };
    for (int j = 0; j < kSPT; ++j) {
        const long long s = s0 + (long long)j * kBlock + tid;
        const bool active = s < total;
        double jf = 0.0;
        if (active) {
            jf = jf_of(s);
            if (!warmup) st.jf[s] = jf;
        }
        if (hist) {
            const double jf2 = jf * jf;
            for (int k = 0; k < d; ++k) {
                if (LDS_HIST) {
                    if (active) hist_pass1(hs, k * ni + id_of(s, k), jf2);
                } else {
                    wave_hist_add_global(st.weights, st.counts, k, ni, active ? id_of(s, k) : -1, jf2);
                }
            }
        }
    }
    if (LDS_HIST && hist) {
        __syncthreads();
        for (int j = 0; j < kSPT; ++j) {
            const long long s = s0 + (long long)j * kBlock + tid;
            if (s >= total) break;
            const double jf = jf_of(s);
            for (int k = 0; k < d; ++k) hist_pass2(hs, k * ni + id_of(s, k), jf * jf);
        }
        __syncthreads();
        for (int e = tid; e < W; e += kBlock) {
            st.hist_w[(long long)blockIdx.x * W + e] = hist_value(hs, e);
            st.hist_c[(long long)blockIdx.x * W + e] = hs.cnt[e];
        }
    }
}

// accumulate_weight on given y / jf2 (step API): block partials
template <typename T>
__global__ __launch_bounds__(kBlock) void k_vegas_accum_given(tq_vegas_state st, const T* y, const T* jf2v,
                                                              long long n) {
    __shared__ unsigned long long s_sum[kHistLds];
    __shared__ uint32_t s_exp[kHistLds];
    __shared__ uint32_t s_cnt[kHistLds];
    const HistSlots hs{s_sum, s_exp, s_cnt};
    const int tid = threadIdx.x;
    const int d = st.dim, ni = st.n_intervals, W = d * ni;
    const long long s0 = (long long)blockIdx.x * kSB;
    hist_zero(hs, W);
    __syncthreads();
    const T tni = (T)ni;
    auto id_of = [&](long long s, int k) {
        int id = (int)floor(y[s * d + k] * tni);
        return id < 0 ? 0 : (id >= ni ? ni - 1 : id);
    };
    for (int pass = 1; pass <= 2; ++pass) {
        for (int j = 0; j < kSPT; ++j) {
            const long long s = s0 + (long long)j * kBlock + tid;
            if (s >= n) break;
            const double v = (double)jf2v[s];
            for (int k = 0; k < d; ++k) {
                if (pass == 1) hist_pass1(hs, k * ni + id_of(s, k), v);
                else hist_pass2(hs, k * ni + id_of(s, k), v);
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < W; e += kBlock) {
        st.hist_w[(long long)blockIdx.x * W + e] = hist_value(hs, e);
        st.hist_c[(long long)blockIdx.x * W + e] = hs.cnt[e];
    }
}

// ---------------------------------------------------------------------------
// stratification: per-cube sums and update_DH
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_vegas_strat(tq_vegas_state st, const double* w, int update_dh) {
    __shared__ double s_red[kWavesPerBlock];
    const long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double ih = 0.0, sc = 0.0, dr = 0.0;
    if (h < st.n_cubes) {
        long long a = st.offsets[h], b = st.offsets[h + 1];
        if (b > st.max_samples) b = st.max_samples;
        if (a > b) a = b;
        double JF = 0.0, JF2 = 0.0;
        // sequential sums (the oracle's order); loads issued 8 at a time so
        // that cubes holding many samples are not latency-bound
        long long s = a;
        for (; s + 8 <= b; s += 8) {
            double v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = w[s + i];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                JF += v[i];
                JF2 += v[i] * v[i];
            }
        }
        for (; s < b; ++s) {
            const double v = w[s];
            JF += v;
            JF2 += v * v;
        }
        st.jf_cube[h] = JF;
        st.jf2_cube[h] = JF2;
        const double cnt = (double)st.nh[h];
        const double V = st.v_cubes;
        const double ninv = 1.0 / cnt;
        ih = JF * (ninv * V);
        sc = fabs(JF2 * ninv * (V * V) - ih * ih) * ninv;
        if (update_dh) {
            const double V2 = V * V;
            const double t = V * JF / cnt;
            double dt = V2 * JF2 / cnt - t * t;
            if (dt < 0.0) dt = 0.0;
            dr = pow(dt, st.beta);
            st.dh[h] = dr;
        }
    }
    const double b1 = block_sum(ih, s_red);
    const double b2 = block_sum(sc, s_red);
    const double b3 = block_sum(dr, s_red);
    if (threadIdx.x == 0) {
        st.block_sums[4 * blockIdx.x] = b1;
        st.block_sums[4 * blockIdx.x + 1] = b2;
        st.block_sums[4 * blockIdx.x + 2] = b3;
    }
}

// update_DH (vegas_stratification.py:72-90) from stored JF/JF2/nh, without
// the normalisation; block sums of the damped counts into block_sums[.][2].
__global__ __launch_bounds__(kBlock) void k_vegas_update_dh(tq_vegas_state st) {
    __shared__ double s_red[kWavesPerBlock];
    const long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double dr = 0.0;
    if (h < st.n_cubes) {
        const double JF = st.jf_cube[h], JF2 = st.jf2_cube[h];
        const double cnt = (double)st.nh[h];
        const double V = st.v_cubes;
        const double V2 = V * V;
        const double t = V * JF / cnt;
        double dt = V2 * JF2 / cnt - t * t;
        if (dt < 0.0) dt = 0.0;
        dr = pow(dt, st.beta);
        st.dh[h] = dr;
    }
    const double b3 = block_sum(dr, s_red);
    if (threadIdx.x == 0) {
        st.block_sums[4 * blockIdx.x] = 0.0;
        st.block_sums[4 * blockIdx.x + 1] = 0.0;
        st.block_sums[4 * blockIdx.x + 2] = b3;
    }
}

// The iteration statistics from k_vegas_strat's block partials, by one
// 1024-thread block (its own kernel, or an extra block of k_vegas_map_rows).
__device__ inline void strat_finalize_block(const tq_vegas_state& st, long long nb, int it, double* s_red) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (long long i = threadIdx.x; i < nb; i += blockDim.x) {
        a += st.block_sums[4 * i];
        b += st.block_sums[4 * i + 1];
        c += st.block_sums[4 * i + 2];
    }
    a = block_sum(a, s_red);
    b = block_sum(b, s_red);
    c = block_sum(c, s_red);
    if (threadIdx.x == 0) {
        st.iter_stats[4 * it] = a;
        st.iter_stats[4 * it + 1] = b;
        st.iter_stats[4 * it + 3] = c;
    }
}

__global__ __launch_bounds__(1024) void k_vegas_strat_finalize(tq_vegas_state st, long long nb, int it) {
    __shared__ double s_red[1024 / kWave];
    strat_finalize_block(st, nb, it, s_red);
}

__global__ __launch_bounds__(kBlock) void k_vegas_normalize(tq_vegas_state st, int it) {
    const double s = st.iter_stats[4 * it + 3];
    if (s == 0.0) return;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < st.n_cubes; h += stride)
        st.dh[h] = st.dh[h] / s;
}

// ---------------------------------------------------------------------------
// map update
// ---------------------------------------------------------------------------
// weights/counts [dim][ni] = sum over sample blocks of the block partials
// (one wave per bin, lanes over blocks, fixed order); add != 0 accumulates.
__global__ __launch_bounds__(kBlock) void k_vegas_map_reduce(tq_vegas_state st, long long nsamp, int it, int add) {
    const int W = st.dim * st.n_intervals;
    const long long pair = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    long long total = nsamp >= 0 ? nsamp : (long long)st.iter_stats[4 * it + 2];
    if (nsamp < 0 && total > st.max_samples) total = st.max_samples;
    long long nb = (total + kSB - 1) / kSB;
    if (nb > st.max_blocks) nb = st.max_blocks;
    if (pair >= W) return;
    double w = 0.0, c = 0.0;
    for (long long b = lane; b < nb; b += kWave) {
        w += st.hist_w[b * W + pair];
        c += (double)st.hist_c[b * W + pair];
    }
    w = wave_sum(w);
    c = wave_sum(c);
    if (lane == 0) {
        if (add) {
            st.weights[pair] += w;
            st.counts[pair] += c;
        } else {
            st.weights[pair] = w;
            st.counts[pair] = c;
        }
    }
}

// _smooth_map for one axis (vegas_map.py:113-172) up to the weight sum:
// zero-count fill and division; returns the row sum.  w, z: scratch rows.
__device__ inline double smooth_fill_axis(const double* weights, const double* counts, int n, double* w,
                                          double* z, bool any_zero) {
    for (int j = 0; j < n; ++j) {
        z[j] = (counts[j] == 0.0) ? 1.0 : 0.0;
        w[j] = (counts[j] != 0.0) ? weights[j] / counts[j] : weights[j];
        if (!any_zero) w[j] = weights[j] / counts[j];
    }
    if (any_zero) {
        for (int pass = 0; pass < 10; ++pass) {
            bool left = false;
            for (int j = 0; j < n - 1; ++j)
                if (z[j] != 0.0) w[j] = w[j + 1];
            for (int j = 0; j < n - 1; ++j) z[j] = (z[j] != 0.0 && z[j + 1] != 0.0) ? 1.0 : 0.0;
            for (int j = n - 1; j >= 1; --j)
                if (z[j] != 0.0) w[j] = w[j - 1];
            for (int j = n - 1; j >= 1; --j) z[j] = (z[j] != 0.0 && z[j - 1] != 0.0) ? 1.0 : 0.0;
            for (int j = 0; j < n; ++j) left = left || (z[j] != 0.0);
            if (!left) break;
        }
    }
    double sum = 0.0;
    for (int j = 0; j < n; ++j) sum += w[j];
    return sum;
}

// [1,6,1]/8 smoothing (edges [7,1]/8) and range compression (EQ 18-19).
__device__ inline void smooth_finish_axis(const double* w, int n, double sum, double alpha, double* out) {
    for (int j = 0; j < n; ++j) {
        double d;
        if (j == 0) d = 7.0 * w[0] + w[1];
        else if (j == n - 1) d = w[n - 2] + 7.0 * w[n - 1];
        else d = w[j - 1] + 6.0 * w[j] + w[j + 1];
        d = d / (8.0 * sum);
        if (d != 0.0) {
            const double r = (d - 1.0) / log(d);
            d = (alpha == 0.5) ? sqrt(r) : pow(r, alpha);
        }
        out[j] = d;
    }
}

// update_map for one axis from smoothed weights sw (vegas_map.py:198-259).
// scratch: num/val [n+1], newx [n+1].  Returns flags bit0 = non-finite edges
// were found, bit1 = some remained after repair.
__device__ inline int update_axis(double* x_edges, double* dx_edges, const double* sw, int n, double* num,
                                  double* val, double* newx) {
    double total = 0.0;
    for (int j = 0; j < n; ++j) total += sw[j];
    const double delta_d = total / (double)n;
    for (int m = 0; m <= n; ++m) {
        num[m] = 0.0;
        val[m] = 0.0;
    }
    double cum = 0.0;
    for (int j = 0; j < n - 1; ++j) {
        cum += sw[j];
        const double q = cum / delta_d;
        long long mlt = (q >= 0.0 && q <= (double)n) ? (long long)q : -1;
        if (mlt < 0) return 3;  // NaN / out of range: edges cannot be formed
        num[mlt] += 1.0;
        val[mlt] += sw[j];
    }
    long long idx = 0;
    double accu = 0.0;
    for (int i = 0; i < n - 1; ++i) {
        idx += (long long)num[i];
        accu += delta_d - val[i];
        long long ix = idx > n - 1 ? n - 1 : idx;
        newx[i + 1] = x_edges[ix] + accu / sw[ix] * dx_edges[ix];
    }
    newx[0] = x_edges[0];
    newx[n] = x_edges[n];
    int flags = 0;
    bool all_finite = true;
    for (int j = 0; j <= n; ++j) all_finite = all_finite && isfinite(newx[j]);
    if (!all_finite) {
        flags |= 1;
        // replace by the mean of the neighbours (vegas_map.py:252-255)
        for (int j = 1; j < n; ++j) val[j] = isfinite(newx[j]) ? 1.0 : 0.0;
        for (int j = 1; j < n; ++j) num[j] = 0.5 * (newx[j - 1] + newx[j + 1]);
        for (int j = 1; j < n; ++j)
            if (val[j] == 0.0) newx[j] = num[j];
        for (int j = 0; j <= n; ++j)
            if (!isfinite(newx[j])) flags |= 2;
        if (flags & 2) return flags;
    }
    for (int j = 0; j <= n; ++j) x_edges[j] = newx[j];
    for (int j = 0; j < n; ++j) dx_edges[j] = x_edges[j + 1] - x_edges[j];
    return flags;
}

// LDS writes of this wave visible to all of its lanes
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---- update_map: one block per axis, two kernels (vegas_map.py:113-261) ----
// k_vegas_map_smooth: weight/count (or the raw weight when a count is zero
// anywhere), the zero-count fill, the row sum and the smoothed weights.
// k_vegas_map_edges: the "cannot update" test over all axes' sums, then the
// new edges from cumulative sums, all as block-wide scans, so the work per
// axis is O(n / 1024) steps rather than a walk of n bins on one lane.
// Scratch per axis (hist_w, 5 rows of n+1 doubles): row 1 slot n holds the
// row sum for the second kernel, row 2 the smoothed weights; rows 0, 3, 4
// are work space (in LDS instead when 5 (n+1) doubles fit).
constexpr int kMapBlock = 1024;
constexpr int kMapWaves = kMapBlock / kWave;

// Sum over the block in a fixed order; every thread gets the result.
__device__ inline double map_block_sum(double v, double* s_red) {
    v = wave_sum(v);
    if ((threadIdx.x & (kWave - 1)) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < kMapWaves; ++w) r += s_red[w];
    __syncthreads();
    return r;
}

// Inclusive scan of one value per thread, plus `carry`; `carry` becomes the
// running total after this chunk.
__device__ inline double map_block_scan(double v, double& carry, double* s_red) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
    for (int off = 1; off < kWave; off <<= 1) {
        const double o = __shfl_up(v, off, kWave);
        if (lane >= off) v += o;
    }
    if (lane == kWave - 1) s_red[w] = v;
    __syncthreads();
    double pre = carry, all = carry;
    for (int i = 0; i < kMapWaves; ++i) {
        if (i < w) pre += s_red[i];
        all += s_red[i];
    }
    __syncthreads();
    carry = all;
    return v + pre;
}

__device__ inline double axis_weight(const tq_vegas_state& st, int k, int j, bool any_zero) {
    const int n = st.n_intervals;
    const double wt = st.weights[k * n + j], ct = st.counts[k * n + j];
    return (!any_zero || ct != 0.0) ? wt / ct : wt;
}

__global__ __launch_bounds__(kMapBlock) void k_vegas_map_smooth(tq_vegas_state st, int lds) {
    extern __shared__ double s_dyn[];
    __shared__ double s_red[kMapWaves];
    const int k = blockIdx.x, tid = threadIdx.x;
    const int d = st.dim, n = st.n_intervals, W = d * n;
    double* g = st.hist_w + (long long)k * 5 * (n + 1);
    double* base = lds ? s_dyn : g;
    double* w = base;
    double* z = base + (n + 1);
    double* w2 = base + 3 * (n + 1);
    double* z2 = base + 4 * (n + 1);
    int zero_here = 0;
    for (int e = tid; e < W; e += kMapBlock) zero_here |= (st.counts[e] == 0.0) ? 1 : 0;
    const bool any_zero = __syncthreads_or(zero_here) != 0;
    for (int j = tid; j < n; j += kMapBlock) {
        z[j] = (st.counts[k * n + j] == 0.0) ? 1.0 : 0.0;
        w[j] = axis_weight(st, k, j, any_zero);
    }
    __syncthreads();
    if (any_zero) {
        // up to 10 passes, each reading only values from before the pass
        for (int pass = 0; pass < 10; ++pass) {
            for (int j = tid; j < n; j += kMapBlock) {
                const bool zj = z[j] != 0.0;
                w2[j] = (j < n - 1 && zj) ? w[j + 1] : w[j];
                z2[j] = (j < n - 1) ? ((zj && z[j + 1] != 0.0) ? 1.0 : 0.0) : z[j];
            }
            __syncthreads();
            int left = 0;
            for (int j = tid; j < n; j += kMapBlock) {
                const bool zj = z2[j] != 0.0;
                w[j] = (j >= 1 && zj) ? w2[j - 1] : w2[j];
                const bool zn = (j >= 1) ? (zj && z2[j - 1] != 0.0) : zj;
                z[j] = zn ? 1.0 : 0.0;
                left |= zn ? 1 : 0;
            }
            if (!__syncthreads_or(left)) break;
        }
    }
    double part = 0.0;
    for (int j = tid; j < n; j += kMapBlock) part += w[j];
    const double sum = map_block_sum(part, s_red);
    // [1,6,1]/8 smoothing (edges [7,1]/8) and range compression (EQ 18-19)
    const double alpha = st.alpha;
    double* sw = g + 2 * (n + 1);
    for (int j = tid; j < n; j += kMapBlock) {
        double dd;
        if (j == 0) dd = 7.0 * w[0] + w[1];
        else if (j == n - 1) dd = w[n - 2] + 7.0 * w[n - 1];
        else dd = w[j - 1] + 6.0 * w[j] + w[j + 1];
        dd = dd / (8.0 * sum);
        if (dd != 0.0) {
            const double r = (dd - 1.0) / log(dd);
            dd = (alpha == 0.5) ? sqrt(r) : pow(r, alpha);
        }
        sw[j] = dd;
    }
    if (tid == 0) g[(n + 1) + n] = sum;
}

__global__ __launch_bounds__(kMapBlock) void k_vegas_map_edges(tq_vegas_state st, int reset, int lds) {
    extern __shared__ double s_dyn[];
    __shared__ double s_red[kMapWaves];
    __shared__ unsigned s_rk[kMapWaves], s_ck[kMapWaves + 1];
    __shared__ double s_rs[kMapWaves], s_cs[kMapWaves + 1];
    __shared__ int s_rc[kMapWaves], s_cc[kMapWaves + 1];
    const int k = blockIdx.x, tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid >> 6;
    const int d = st.dim, n = st.n_intervals;
    double* g = st.hist_w + (long long)k * 5 * (n + 1);
    bool cannot = false;
    for (int a = 0; a < d; ++a) cannot = cannot || (st.hist_w[(long long)a * 5 * (n + 1) + (n + 1) + n] == 0.0);
    if (cannot) {
        if (k == 0 && tid == 0) st.flags[0] += 1;
    } else {
        double* base = lds ? s_dyn : g;
        double* sw = base + 2 * (n + 1);
        if (lds) {
            for (int j = tid; j < n; j += kMapBlock) sw[j] = g[2 * (n + 1) + j];
            __syncthreads();
        }
        int* mlt = reinterpret_cast<int*>(base);  // row 0, later newx
        double* newx = base;
        double* num = base + 3 * (n + 1);
        double* val = base + 4 * (n + 1);
        double* xe = st.x_edges + k * (n + 1);
        double* dxe = st.dx_edges + k * n;
        double part = 0.0;
        for (int j = tid; j < n; j += kMapBlock) part += sw[j];
        const double total = map_block_sum(part, s_red);
        const double delta_d = total / (double)n;
        // bin of each cumulative sum: mlt[j] = floor(cum_j / delta_d), j < n-1
        int bad = 0;
        double carry = 0.0;
        for (int j0 = 0; j0 < n - 1; j0 += kMapBlock) {
            const int j = j0 + tid;
            const double cum = map_block_scan(j < n - 1 ? sw[j] : 0.0, carry, s_red);
            if (j < n - 1) {
                const double q = cum / delta_d;
                if (q >= 0.0 && q <= (double)n) {
                    mlt[j] = (int)q;
                } else {
                    mlt[j] = 0;
                    bad = 1;
                }
            }
        }
        for (int m = tid; m <= n; m += kMapBlock) {
            num[m] = 0.0;
            val[m] = 0.0;
        }
        const int status = __syncthreads_or(bad) ? 3 : 0;
        if (status == 0) {
            // per-bin counts and weight sums of the cumulative bins: runs of
            // equal mlt are contiguous (cum is nondecreasing); DPP run scan in
            // each wave, carried across waves and chunks in wave order
            unsigned ck = 0xFFFFFFFFu;
            double cs = 0.0;
            int cc = 0;
            for (int j0 = 0; j0 < n - 1; j0 += kMapBlock) {
                const int j = j0 + tid;
                const bool valid = j < n - 1;
                const unsigned key = valid ? (unsigned)mlt[j] : 0xFFFFFFFEu;
                double sum = valid ? sw[j] : 0.0;
                int cnt = valid ? 1 : 0;
                run_scan_step<0x111, 0xF>(key, sum, cnt);
                run_scan_step<0x112, 0xF>(key, sum, cnt);
                run_scan_step<0x114, 0xF>(key, sum, cnt);
                run_scan_step<0x118, 0xF>(key, sum, cnt);
                run_scan_step<0x142, 0xA>(key, sum, cnt);
                run_scan_step<0x143, 0xC>(key, sum, cnt);
                if (lane == kWave - 1) {
                    s_rk[wv] = key;
                    s_rs[wv] = sum;
                    s_rc[wv] = cnt;
                }
                __syncthreads();
                if (tid == 0) {
                    for (int i = 0; i < kMapWaves; ++i) {
                        s_ck[i] = ck;
                        s_cs[i] = cs;
                        s_cc[i] = cc;
                        if (s_rk[i] == ck) {
                            cs = s_rs[i] + cs;
                            cc = s_rc[i] + cc;
                        } else {
                            ck = s_rk[i];
                            cs = s_rs[i];
                            cc = s_rc[i];
                        }
                    }
                    s_ck[kMapWaves] = ck;
                    s_cs[kMapWaves] = cs;
                    s_cc[kMapWaves] = cc;
                }
                __syncthreads();
                if (key == s_ck[wv]) {
                    sum += s_cs[wv];
                    cnt += s_cc[wv];
                }
                if (valid && (j == n - 2 || mlt[j + 1] != (int)key)) {
                    num[key] = (double)cnt;
                    val[key] = sum;
                }
                ck = s_ck[kMapWaves];
                cs = s_cs[kMapWaves];
                cc = s_cc[kMapWaves];
                __syncthreads();
            }
            // new edge i+1 from the running count and remainder (vegas_map.py:238-249)
            double cidx = 0.0, cacc = 0.0;
            for (int i0 = 0; i0 < n - 1; i0 += kMapBlock) {
                const int i = i0 + tid;
                const bool valid = i < n - 1;
                const double idx = map_block_scan(valid ? num[i] : 0.0, cidx, s_red);
                const double accu = map_block_scan(valid ? delta_d - val[i] : 0.0, cacc, s_red);
                if (valid) {
                    const long long ix0 = (long long)idx;
                    const int ix = (int)(ix0 > n - 1 ? n - 1 : ix0);
                    val[i] = xe[ix] + accu / sw[ix] * dxe[ix];  // staged: newx shares row 0 with mlt
                }
            }
            __syncthreads();
            for (int i = tid; i <= n; i += kMapBlock)
                newx[i] = (i == 0) ? xe[0] : ((i == n) ? xe[n] : val[i - 1]);
            __syncthreads();
            int nonfinite = 0;
            for (int i = tid; i <= n; i += kMapBlock) nonfinite |= isfinite(newx[i]) ? 0 : 1;
            int fl = 0;
            if (__syncthreads_or(nonfinite)) {
                fl |= 1;
                // replace by the mean of the neighbours (vegas_map.py:252-255)
                for (int i = tid; i <= n; i += kMapBlock)
                    num[i] = (i >= 1 && i < n && !isfinite(newx[i])) ? 0.5 * (newx[i - 1] + newx[i + 1]) : newx[i];
                __syncthreads();
                int still = 0;
                for (int i = tid; i <= n; i += kMapBlock) {
                    newx[i] = num[i];
                    still |= isfinite(num[i]) ? 0 : 1;
                }
                if (__syncthreads_or(still)) fl |= 2;
            }
            if (!(fl & 2)) {
                for (int i = tid; i <= n; i += kMapBlock) xe[i] = newx[i];
                for (int i = tid; i < n; i += kMapBlock) dxe[i] = newx[i + 1] - newx[i];
            }
            if (tid == 0) {
                if (fl & 1) atomicAdd(&st.flags[1], 1);
                if (fl & 2) atomicAdd(&st.flags[2], 1);
            }
        } else if (tid == 0) {
            atomicAdd(&st.flags[1], 1);
            atomicAdd(&st.flags[2], 1);
        }
    }
    if (reset) {
        // the other axes' blocks read only the row sums and their own rows
        for (int j = tid; j < n; j += kMapBlock) {
            st.weights[k * n + j] = 0.0;
            st.counts[k * n + j] = 0.0;
        }
    }
}

// ---- update_map for small maps (vegas_map.py:113-261) ----
// At most 64 bins per axis and 16 axes.  Histogram mode: G blocks sum the
// sample blocks' histogram rows into G partial rows (two levels of about
// sqrt(rows) loads per thread instead of one long walk), then one block sums
// those and runs the update with wave k < dim on axis k, one lane per bin:
// weight/count, zero-count fill, smoothing, and the new edges from wave sums
// and scans.  These are the one-wave case of the block kernels above (a
// block sum over one busy wave adds only zeros to that wave's sum), so from
// the same weights and counts both give the same edges.  The histogram row
// sums group the blocks differently from k_vegas_map_reduce (last-bit
// differences in the weights).
constexpr int kSmallMapBins = kWave;
constexpr int kSmallMapAxes = kMapWaves;
constexpr int kSmallMapEdges = kSmallMapBins + 1;

// Inclusive scan over the wave's lanes (Hillis-Steele, as map_block_scan).
__device__ inline double wave_scan(double v, int lane) {
    for (int off = 1; off < kWave; off <<= 1) {
        const double o = __shfl_up(v, off, kWave);
        if (lane >= off) v += o;
    }
    return v;
}

// Column e of rows r0, r0 + step, ... < end of a [rows][W] histogram, added
// in row order onto (w, c); loads go out kRowBatch at a time, so a thread
// waits about once per batch instead of once per row.
constexpr int kRowBatch = 8;
__device__ inline void column_sum(const double* hw, const uint32_t* hc, long long W, int e, long long r0,
                                  long long step, long long end, double& w, unsigned long long& c) {
    const long long stride = step * W;
    const double* pw = hw + r0 * W + e;
    const uint32_t* pc = hc + r0 * W + e;
    for (long long r = r0; r < end; r += kRowBatch * step, pw += kRowBatch * stride, pc += kRowBatch * stride) {
        double a[kRowBatch];
        uint32_t b[kRowBatch];
#pragma unroll
        for (int i = 0; i < kRowBatch; ++i) {
            const bool in = r + i * step < end;
            a[i] = in ? pw[i * stride] : 0.0;
            b[i] = in ? pc[i * stride] : 0u;
        }
#pragma unroll
        for (int i = 0; i < kRowBatch; ++i) {
            w += a[i];
            c += b[i];
        }
    }
}

// Histogram mode, first kernel: block g of G sums rows g, g+G, g+2G, ... of
// the sample blocks' hist_w/hist_c (S = 1024 / W threads per column, each
// taking every S-th of those rows; the S sums added in order) and stores the
// result over its own row g, which no other block reads.
// Block `rows` (present when nb_strat > 0) computes the iteration statistics
// from nb_strat k_vegas_strat partials instead, as k_vegas_strat_finalize.
__global__ __launch_bounds__(kMapBlock) void k_vegas_map_rows(tq_vegas_state st, long long nsamp, int it, int rows,
                                                              long long nb_strat) {
    __shared__ double s_pw[kMapBlock], s_pc[kMapBlock];
    __shared__ double s_red[kMapWaves];
    if ((int)blockIdx.x == rows) {
        strat_finalize_block(st, nb_strat, it, s_red);
        return;
    }
    const int tid = threadIdx.x;
    const int W = st.dim * st.n_intervals, S = kMapBlock / W, e = tid % W, sub = tid / W;
    long long total = nsamp >= 0 ? nsamp : (long long)st.iter_stats[4 * it + 2];
    if (nsamp < 0 && total > st.max_samples) total = st.max_samples;
    long long nb = (total + kSB - 1) / kSB;
    if (nb > st.max_blocks) nb = st.max_blocks;
    const long long G = rows, g = blockIdx.x;
    double w = 0.0;
    unsigned long long c = 0;
    if (sub < S) column_sum(st.hist_w, st.hist_c, W, e, g + sub * G, S * G, nb, w, c);
    s_pw[tid] = w;
    s_pc[tid] = (double)c;
    __syncthreads();
    if (tid < W) {
        double wr = 0.0;
        unsigned long long cr = 0;
        for (int q = 0; q < S; ++q) {
            wr += s_pw[q * W + tid];
            cr += (unsigned long long)s_pc[q * W + tid];
        }
        st.hist_w[g * W + tid] = wr;
        st.hist_c[g * W + tid] = (uint32_t)cr;  // < 2^32: the host requires max_samples < 2^32
    }
}

// Second kernel, one block: rows > 0 sums the first `rows` rows of
// hist_w/hist_c into weights/counts; rows == 0 takes the weights and counts
// as they stand and zeroes them (_reset_weight).  Then the map update.
__global__ __launch_bounds__(kMapBlock) void k_vegas_map_small(tq_vegas_state st, int rows) {
    __shared__ double s_tw[kMapBlock], s_tc[kMapBlock];  // totals per (axis, bin)
    __shared__ double s_rows[kMapWaves][4][kSmallMapEdges];  // per-wave rows; first the sub-sums
    __shared__ double s_sum[kMapWaves];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), k = tid >> 6;
    const int d = st.dim, n = st.n_intervals, W = d * n;
    if (rows > 0) {
        double* s_pw = &s_rows[0][0][0];
        double* s_pc = s_pw + kMapBlock;
        const int S = kMapBlock / W, e = tid % W, sub = tid / W;
        double w = 0.0;
        unsigned long long c = 0;
        if (sub < S) column_sum(st.hist_w, st.hist_c, W, e, sub, S, rows, w, c);
        s_pw[tid] = w;
        s_pc[tid] = (double)c;
        __syncthreads();
        if (tid < W) {
            double wr = 0.0, cr = 0.0;
            for (int q = 0; q < S; ++q) {
                wr += s_pw[q * W + tid];
                cr += s_pc[q * W + tid];
            }
            s_tw[tid] = wr;
            s_tc[tid] = cr;
            st.weights[tid] = wr;
            st.counts[tid] = cr;
        }
    } else {
        for (int e = tid; e < W; e += kMapBlock) {
            s_tw[e] = st.weights[e];
            s_tc[e] = st.counts[e];
            st.weights[e] = 0.0;
            st.counts[e] = 0.0;
        }
    }
    const bool any_zero = __syncthreads_or(tid < W && s_tc[tid] == 0.0) != 0;

    // smoothing (vegas_map.py:113-172), lane j = bin j of axis k
    const bool axis = k < d;
    const int j = lane;
    const bool bin = axis && j < n;
    double w = 0.0;
    bool z = false;
    if (bin) {
        const double wt = s_tw[k * n + j], ct = s_tc[k * n + j];
        z = ct == 0.0;
        w = (!any_zero || ct != 0.0) ? wt / ct : wt;
    }
    if (axis && any_zero) {
        // up to 10 passes, each reading only values from before the pass
        for (int pass = 0; pass < 10; ++pass) {
            const double wn = __shfl_down(w, 1, kWave);
            const int zn = __shfl_down(z ? 1 : 0, 1, kWave);
            const double w2 = (j < n - 1 && z) ? wn : w;
            const bool z2 = (j < n - 1) ? (z && zn != 0) : z;
            const double wp = __shfl_up(w2, 1, kWave);
            const int zp = __shfl_up(z2 ? 1 : 0, 1, kWave);
            w = (j >= 1 && z2) ? wp : w2;
            z = (j >= 1) ? (z2 && zp != 0) : z2;
            if (__ballot(bin && z) == 0) break;
        }
    }
    const double sum = wave_sum(bin ? w : 0.0);
    const double wl = __shfl_up(w, 1, kWave), wr = __shfl_down(w, 1, kWave);
    double sw = 0.0;
    if (bin) {
        // [1,6,1]/8 smoothing (edges [7,1]/8) and range compression (EQ 18-19)
        double dd;
        if (j == 0) dd = 7.0 * w + wr;
        else if (j == n - 1) dd = wl + 7.0 * w;
        else dd = wl + 6.0 * w + wr;
        dd = dd / (8.0 * sum);
        if (dd != 0.0) {
            const double r = (dd - 1.0) / log(dd);
            dd = (st.alpha == 0.5) ? sqrt(r) : pow(r, st.alpha);
        }
        sw = dd;
    }
    if (axis && lane == 0) s_sum[k] = sum;
    __syncthreads();
    bool cannot = false;
    for (int a = 0; a < d; ++a) cannot = cannot || (s_sum[a] == 0.0);
    if (cannot) {
        if (tid == 0) st.flags[0] += 1;
        return;
    }
    if (!axis) return;

    // new edges (vegas_map.py:198-259); from here on each wave works alone
    double* num = s_rows[k][0];
    double* val = s_rows[k][1];
    double* newx = s_rows[k][2];
    double* tmp = s_rows[k][3];
    double* xe = st.x_edges + k * (n + 1);
    double* dxe = st.dx_edges + k * n;
    const double total = wave_sum(sw);
    const double delta_d = total / (double)n;
    const bool first = j < n - 1;  // the n-1 cumulative sums
    const double cum = wave_scan(first ? sw : 0.0, lane);
    int mlt = 0;
    bool bad = false;
    if (first) {
        const double q = cum / delta_d;
        if (q >= 0.0 && q <= (double)n) mlt = (int)q;
        else bad = true;
    }
    for (int m = lane; m <= n; m += kWave) {
        num[m] = 0.0;
        val[m] = 0.0;
    }
    if (__ballot(bad) != 0) {
        if (lane == 0) {
            atomicAdd(&st.flags[1], 1);
            atomicAdd(&st.flags[2], 1);
        }
        return;
    }
    wave_lds_sync();
    // count and weight sum of each cumulative bin: runs of equal mlt are
    // contiguous (cum is nondecreasing); DPP run scan, written at run ends
    unsigned key = first ? (unsigned)mlt : 0xFFFFFFFEu;
    double rs = first ? sw : 0.0;
    int cnt = first ? 1 : 0;
    run_scan_step<0x111, 0xF>(key, rs, cnt);
    run_scan_step<0x112, 0xF>(key, rs, cnt);
    run_scan_step<0x114, 0xF>(key, rs, cnt);
    run_scan_step<0x118, 0xF>(key, rs, cnt);
    run_scan_step<0x142, 0xA>(key, rs, cnt);
    run_scan_step<0x143, 0xC>(key, rs, cnt);
    const int next_mlt = __shfl_down(mlt, 1, kWave);
    if (first && (j == n - 2 || next_mlt != mlt)) {
        num[key] = (double)cnt;
        val[key] = rs;
    }
    wave_lds_sync();
    // new edge j+1 from the running count and remainder (vegas_map.py:238-249)
    const double idx = wave_scan(first ? num[j] : 0.0, lane);
    const double accu = wave_scan(first ? delta_d - val[j] : 0.0, lane);
    const long long ix0 = first ? (long long)idx : 0;
    const int ix = (int)(ix0 > n - 1 ? n - 1 : ix0);
    const double sw_ix = __shfl(sw, ix, kWave);
    if (first) tmp[j] = xe[ix] + accu / sw_ix * dxe[ix];
    wave_lds_sync();
    for (int m = lane; m <= n; m += kWave) newx[m] = (m == 0) ? xe[0] : ((m == n) ? xe[n] : tmp[m - 1]);
    wave_lds_sync();
    bool nonfinite = false;
    for (int m = lane; m <= n; m += kWave) nonfinite = nonfinite || !isfinite(newx[m]);
    int fl = 0;
    if (__ballot(nonfinite) != 0) {
        fl |= 1;
        // replace by the mean of the neighbours (vegas_map.py:252-255)
        for (int m = lane; m <= n; m += kWave)
            tmp[m] = (m >= 1 && m < n && !isfinite(newx[m])) ? 0.5 * (newx[m - 1] + newx[m + 1]) : newx[m];
        wave_lds_sync();
        bool still = false;
        for (int m = lane; m <= n; m += kWave) {
            newx[m] = tmp[m];
            still = still || !isfinite(tmp[m]);
        }
        if (__ballot(still) != 0) fl |= 2;
        wave_lds_sync();
    }
    if (!(fl & 2)) {
        for (int m = lane; m <= n; m += kWave) xe[m] = newx[m];
        for (int m = lane; m < n; m += kWave) dxe[m] = newx[m + 1] - newx[m];
    }
    if (lane == 0) {
        if (fl & 1) atomicAdd(&st.flags[1], 1);
        if (fl & 2) atomicAdd(&st.flags[2], 1);
    }
}

__global__ __launch_bounds__(64) void k_vegas_smooth(int d, int n, const double* weights, const double* counts,
                                                     double alpha, double* out, double* scratch, int32_t* flag) {
    __shared__ double s_sum[TQ_MAX_DIM];
    __shared__ int s_anyzero;
    const int k = threadIdx.x;
    if (k == 0) s_anyzero = 0;
    __syncthreads();
    if (k < d)
        for (int j = 0; j < n; ++j)
            if (counts[k * n + j] == 0.0) atomicOr(&s_anyzero, 1);
    __syncthreads();
    double* w = scratch + (long long)k * 2 * n;
    double* z = w + n;
    if (k < d) s_sum[k] = smooth_fill_axis(weights + k * n, counts + k * n, n, w, z, s_anyzero != 0);
    __syncthreads();
    bool cannot = false;
    for (int i = 0; i < d; ++i) cannot = cannot || (s_sum[i] == 0.0);
    if (k == 0) *flag = cannot ? 1 : 0;
    if (!cannot && k < d) smooth_finish_axis(w, n, s_sum[k], alpha, out + k * n);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_vegas_map_eval(tq_vegas_state st, const T* y, long long n,
                                                           int64_t* ids_out, T* off_out, T* x_out, T* jac_out) {
    const int d = st.dim, ni = st.n_intervals;
    const T tni = (T)ni;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += stride) {
        T jac = T(1);
        for (int k = 0; k < d; ++k) {
            const T yn = y[s * d + k] * tni;
            const T fl = floor(yn);
            int id = (int)fl;
            const int idc = id < 0 ? 0 : (id >= ni ? ni - 1 : id);
            const T off = yn - fl;
            if (ids_out) ids_out[s * d + k] = (int64_t)id;
            if (off_out) off_out[s * d + k] = off;
            if (x_out) x_out[s * d + k] = (T)(st.x_edges[k * (ni + 1) + idc] + st.dx_edges[k * ni + idc] * (double)off);
            jac = jac * (T)((double)ni * st.dx_edges[k * ni + idc]);
        }
        if (jac_out) jac_out[s] = jac;
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static int nh_blocks(const tq_vegas_state* st) { return (int)((st->n_cubes + kScanBlockCubes - 1) / kScanBlockCubes); }

static int launch_nh(const tq_vegas_state* st, int64_t n_exp, int normalize_it, int it, const int64_t* nh_in,
                     hipStream_t s) {
    const int nb = nh_blocks(st);
    hipLaunchKernelGGL(k_vegas_nh, dim3(nb), dim3(kBlock), 0, s, *st, (double)n_exp, normalize_it, nh_in);
    int rc = check_launch("k_vegas_nh");
    if (rc) return rc;
    const int raw = nb <= kBlock ? 1 : 0;
    if (!raw) {
        hipLaunchKernelGGL(k_vegas_scan, dim3(1), dim3(1024), 0, s, st->scan_sums, nb);
        rc = check_launch("k_vegas_scan");
        if (rc) return rc;
    }
    const int ob = (int)((st->n_cubes + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_vegas_offsets, dim3(ob), dim3(kBlock), 0, s, *st, it, nb, raw);
    return check_launch("k_vegas_offsets");
}

static long long sample_blocks(const tq_vegas_state* st, int warmup, long long n_warm) {
    long long b = warmup ? (n_warm + kSB - 1) / kSB : st->max_blocks;
    if (b > st->max_blocks) b = st->max_blocks;
    return b < 1 ? 1 : b;
}

static long long strat_blocks(const tq_vegas_state* st) { return (st->n_cubes + kBlock - 1) / kBlock; }

// finalize == false leaves the iteration statistics to the caller's next
// launch (the extra block of k_vegas_map_rows).
static int launch_strat(const tq_vegas_state* st, const double* w, int it, int update_dh, hipStream_t s,
                        bool finalize = true) {
    const long long nb = strat_blocks(st);
    hipLaunchKernelGGL(k_vegas_strat, dim3((unsigned)nb), dim3(kBlock), 0, s, *st, w, update_dh);
    int rc = check_launch("k_vegas_strat");
    if (rc || !finalize) return rc;
    hipLaunchKernelGGL(k_vegas_strat_finalize, dim3(1), dim3(1024), 0, s, *st, nb, it);
    return check_launch("k_vegas_strat_finalize");
}

// Maps that k_vegas_map_small updates: at most 64 bins and 16 axes, and
// per-bin counts below 2^32 (its partial rows keep them in hist_c).
static bool small_map(const tq_vegas_state* st) {
    return st->n_intervals <= kSmallMapBins && st->dim <= kSmallMapAxes && st->max_samples < (1LL << 32);
}

// Partial rows of the one-launch update: about sqrt(sample blocks), so the
// blocks' row sums and the last block's sum of the rows take similar time.
static int small_map_grid(const tq_vegas_state* st) {
    long long g = (long long)std::ceil(std::sqrt((double)st->max_blocks));
    if (g > 256) g = 256;
    if (g > st->max_blocks) g = st->max_blocks;
    return (int)(g < 1 ? 1 : g);
}

// update_map on the accumulated weights/counts: smoothing, then the edges;
// `reset` zeroes the weights and counts afterwards (_reset_weight).
static int launch_map_kernels(const tq_vegas_state* st, int reset, hipStream_t s) {
    if (reset && small_map(st)) {
        hipLaunchKernelGGL(k_vegas_map_small, dim3(1), dim3(kMapBlock), 0, s, *st, 0);
        return check_launch("k_vegas_map_small");
    }
    const size_t scr = (size_t)5 * (st->n_intervals + 1) * sizeof(double);
    const int lds = scr <= kMapUpdateLds ? 1 : 0;
    hipLaunchKernelGGL(k_vegas_map_smooth, dim3((unsigned)st->dim), dim3(kMapBlock), lds ? scr : 0, s, *st, lds);
    int rc = check_launch("k_vegas_map_smooth");
    if (rc) return rc;
    hipLaunchKernelGGL(k_vegas_map_edges, dim3((unsigned)st->dim), dim3(kMapBlock), lds ? scr : 0, s, *st, reset,
                       lds);
    return check_launch("k_vegas_map_edges");
}

// nb_strat > 0: also finish the iteration statistics from that many
// k_vegas_strat partials (small maps only; see launch_iteration_tail).
static int launch_map_update(const tq_vegas_state* st, long long nsamp, int it, hipStream_t s,
                             long long nb_strat = 0) {
    const int W = st->dim * st->n_intervals;
    if (W <= kHistLds && small_map(st)) {
        const int rows = small_map_grid(st);
        hipLaunchKernelGGL(k_vegas_map_rows, dim3(rows + (nb_strat > 0 ? 1 : 0)), dim3(kMapBlock), 0, s, *st, nsamp,
                           it, rows, nb_strat);
        const int rc = check_launch("k_vegas_map_rows");
        if (rc) return rc;
        hipLaunchKernelGGL(k_vegas_map_small, dim3(1), dim3(kMapBlock), 0, s, *st, rows);
        return check_launch("k_vegas_map_small");
    }
    if (W <= kHistLds) {
        const int rb = (W + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL(k_vegas_map_reduce, dim3(rb), dim3(kBlock), 0, s, *st, nsamp, it, 0);
        int rc = check_launch("k_vegas_map_reduce");
        if (rc) return rc;
    }
    return launch_map_kernels(st, W <= kHistLds ? 0 : 1, s);
}

// The end of an iteration: per-cube sums and the iteration statistics (not
// in warm-up), then the map update.  With the small-map update the
// statistics' one-block reduction runs as an extra block of the row-sum
// kernel: both only need the sample kernel's output.
static int launch_iteration_tail(const tq_vegas_state* st, int it, int hist, int warmup, long long n_warm,
                                 hipStream_t s) {
    const bool fold = hist && !warmup && st->dim * st->n_intervals <= kHistLds && small_map(st);
    if (!warmup) {
        const int rc = launch_strat(st, st->jf, it, 1, s, !fold);
        if (rc) return rc;
    }
    if (hist) return launch_map_update(st, warmup ? n_warm : -1, it, s, fold ? strat_blocks(st) : 0);
    return TQ_OK;
}

__global__ void k_zero(double* p, long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0;
}

template <typename T>
static int vegas_iterate(const tq_vegas_state* st, const tq_integrand* f, const void* domain, uint64_t seed,
                         uint64_t offset, int it, int use_grid_improve, int warmup, long long n_warm,
                         hipStream_t s) {
    const int W = st->dim * st->n_intervals;
    const bool lds_hist = W <= kHistLds;
    const int hist = (use_grid_improve || warmup) ? 1 : 0;
    if (hist && !lds_hist) {
        hipLaunchKernelGGL(k_zero, dim3(64), dim3(kBlock), 0, s, st->weights, (long long)W);
        hipLaunchKernelGGL(k_zero, dim3(64), dim3(kBlock), 0, s, st->counts, (long long)W);
    }
    SampleArgs<T> A;
    A.st = *st;
    A.domain = (const T*)domain;
    A.seed = seed;
    A.offset = offset;
    A.warmup = warmup;
    A.n_warm = n_warm;
    A.hist = hist;
    A.it = it;
    A.x_out = nullptr;
    A.jac_out = nullptr;
    A.ids_out = nullptr;
    A.y_out = nullptr;
    A.p = make_params<T>(*f);
    A.u_in = nullptr;
    const long long nb = sample_blocks(st, warmup, n_warm);
    // exact block histogram: 64-bit sum, exponent and count per bin
    const size_t hist_lds = (size_t)W * (8 + 4 + 4);
    TQ_DISPATCH_FAMILY(f->family, {
        if (lds_hist && st->dim <= 8)
            hipLaunchKernelGGL((k_vegas_sample<T, FAM, MODE_FUSED, true, 8>), dim3((unsigned)nb), dim3(kBlock),
                               hist_lds, s, A);
        else if (lds_hist)
            hipLaunchKernelGGL((k_vegas_sample<T, FAM, MODE_FUSED, true>), dim3((unsigned)nb), dim3(kBlock), hist_lds,
                               s, A);
        else if (st->dim <= 8)
            hipLaunchKernelGGL((k_vegas_sample<T, FAM, MODE_FUSED, false, 8>), dim3((unsigned)nb), dim3(kBlock), 0,
                               s, A);
        else
            hipLaunchKernelGGL((k_vegas_sample<T, FAM, MODE_FUSED, false>), dim3((unsigned)nb), dim3(kBlock), 0, s, A);
    });
    int rc = check_launch("k_vegas_sample");
    if (rc) return rc;
    return launch_iteration_tail(st, it, hist, warmup, n_warm, s);
}

}  // namespace tq

using namespace tq;

extern "C" {

int tq_vegas_init(const tq_vegas_state* st, void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_vegas_init, dim3(256), dim3(kBlock), 0, as_stream(stream), *st);
    return check_launch("k_vegas_init");
}

int tq_vegas_nh(const tq_vegas_state* st, int64_t n_exp, int32_t normalize_it, int32_t it, const int64_t* nh_in,
                void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (it < 0 || it >= TQ_VEGAS_MAX_ITERS || normalize_it >= TQ_VEGAS_MAX_ITERS)
        return fail(TQ_E_ARG, "iteration index out of range");
    if (n_exp < 0) return fail(TQ_E_ARG, "n_exp must be >= 0");
    return launch_nh(st, n_exp, normalize_it, it, nh_in, as_stream(stream));
}

int tq_vegas_iterate(const tq_vegas_state* st, const tq_integrand* f, const void* domain, uint64_t seed,
                     uint64_t offset, int32_t it, int32_t use_grid_improve, int32_t warmup, int64_t n_warm,
                     void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    rc = check_integrand(f, st->dim);
    if (rc) return rc;
    if (!domain) return fail(TQ_E_NULL, "domain is NULL");
    if (it < 0 || it >= TQ_VEGAS_MAX_ITERS) return fail(TQ_E_ARG, "iteration index out of range");
    if (warmup && (n_warm < 0 || (n_warm + kSB - 1) / kSB > st->max_blocks))
        return fail(TQ_E_ARG, "n_warm exceeds the block capacity");
    if (st->sample_dtype == TQ_F32)
        return vegas_iterate<float>(st, f, domain, seed, offset, it, use_grid_improve, warmup, n_warm,
                                    as_stream(stream));
    return vegas_iterate<double>(st, f, domain, seed, offset, it, use_grid_improve, warmup, n_warm,
                                 as_stream(stream));
}

int tq_vegas_points(const tq_vegas_state* st, const void* domain, uint64_t seed, uint64_t offset, int32_t warmup,
                    int64_t n_warm, const void* u_in, void* x_out, double* jac_out, int32_t* ids_out, void* y_out,
                    void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (!domain || !x_out || !jac_out || !ids_out) return fail(TQ_E_NULL, "domain/x/jac/ids is NULL");
    if (warmup && (n_warm < 0 || (n_warm + kSB - 1) / kSB > st->max_blocks))
        return fail(TQ_E_ARG, "n_warm exceeds the block capacity");
    const long long nb = sample_blocks(st, warmup, n_warm);
    hipStream_t s = as_stream(stream);
    tq_integrand dummy{};
    dummy.family = TQ_FAM_CONSTANT;
    dummy.dim = st->dim;
    if (st->sample_dtype == TQ_F32) {
        SampleArgs<float> A{*st, (const float*)domain, seed, offset, warmup, n_warm, 0, 0, (float*)x_out, jac_out,
                            ids_out, (float*)y_out, make_params<float>(dummy), (const float*)u_in};
        hipLaunchKernelGGL((k_vegas_sample<float, TQ_FAM_CONSTANT, MODE_POINTS, false>), dim3((unsigned)nb),
                           dim3(kBlock), 0, s, A);
    } else {
        SampleArgs<double> A{*st, (const double*)domain, seed, offset, warmup, n_warm, 0, 0, (double*)x_out,
                             jac_out, ids_out, (double*)y_out, make_params<double>(dummy), (const double*)u_in};
        hipLaunchKernelGGL((k_vegas_sample<double, TQ_FAM_CONSTANT, MODE_POINTS, false>), dim3((unsigned)nb),
                           dim3(kBlock), 0, s, A);
    }
    return check_launch("k_vegas_sample(points)");
}

int tq_vegas_consume(const tq_vegas_state* st, const void* fvals, const void* domain, const double* jac,
                     const int32_t* ids, int32_t it, int32_t use_grid_improve, int32_t warmup, int64_t n_warm,
                     void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (!fvals || !domain || !jac || !ids) return fail(TQ_E_NULL, "fvals/domain/jac/ids is NULL");
    if (it < 0 || it >= TQ_VEGAS_MAX_ITERS) return fail(TQ_E_ARG, "iteration index out of range");
    hipStream_t s = as_stream(stream);
    const int W = st->dim * st->n_intervals;
    const bool lds_hist = W <= kHistLds;
    const int hist = (use_grid_improve || warmup) ? 1 : 0;
    if (hist && !lds_hist) {
        hipLaunchKernelGGL(k_zero, dim3(64), dim3(kBlock), 0, s, st->weights, (long long)W);
        hipLaunchKernelGGL(k_zero, dim3(64), dim3(kBlock), 0, s, st->counts, (long long)W);
    }
    const long long nb = sample_blocks(st, warmup, n_warm);
    if (st->sample_dtype == TQ_F32) {
        if (lds_hist)
            hipLaunchKernelGGL((k_vegas_consume<float, true>), dim3((unsigned)nb), dim3(kBlock), 0, s, *st,
                               (const float*)fvals, (const float*)domain, jac, ids, warmup, (long long)n_warm, hist);
        else
            hipLaunchKernelGGL((k_vegas_consume<float, false>), dim3((unsigned)nb), dim3(kBlock), 0, s, *st,
                               (const float*)fvals, (const float*)domain, jac, ids, warmup, (long long)n_warm, hist);
    } else {
        if (lds_hist)
            hipLaunchKernelGGL((k_vegas_consume<double, true>), dim3((unsigned)nb), dim3(kBlock), 0, s, *st,
                               (const double*)fvals, (const double*)domain, jac, ids, warmup, (long long)n_warm,
                               hist);
        else
            hipLaunchKernelGGL((k_vegas_consume<double, false>), dim3((unsigned)nb), dim3(kBlock), 0, s, *st,
                               (const double*)fvals, (const double*)domain, jac, ids, warmup, (long long)n_warm,
                               hist);
    }
    rc = check_launch("k_vegas_consume");
    if (rc) return rc;
    return launch_iteration_tail(st, it, hist, warmup, n_warm, s);
}

int tq_vegas_update_map(const tq_vegas_state* st, int64_t n_samples, int32_t it, void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    return launch_map_update(st, n_samples, it, as_stream(stream));
}

int tq_vegas_map_eval(const tq_vegas_state* st, const void* y, int64_t n, int64_t* ids_out, void* off_out,
                      void* x_out, void* jac_out, void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (!y) return fail(TQ_E_NULL, "y is NULL");
    if (n <= 0) return TQ_OK;
    long long blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    if (st->sample_dtype == TQ_F32)
        hipLaunchKernelGGL((k_vegas_map_eval<float>), dim3((unsigned)blocks), dim3(kBlock), 0, as_stream(stream),
                           *st, (const float*)y, (long long)n, ids_out, (float*)off_out, (float*)x_out,
                           (float*)jac_out);
    else
        hipLaunchKernelGGL((k_vegas_map_eval<double>), dim3((unsigned)blocks), dim3(kBlock), 0, as_stream(stream),
                           *st, (const double*)y, (long long)n, ids_out, (double*)off_out, (double*)x_out,
                           (double*)jac_out);
    return check_launch("k_vegas_map_eval");
}

int tq_vegas_map_accumulate(const tq_vegas_state* st, const void* y, const void* jf2, int64_t n, void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (!y || !jf2) return fail(TQ_E_NULL, "y/jf2 is NULL");
    if (st->dim * st->n_intervals > kHistLds) return fail(TQ_E_UNSUPPORTED, "dim*n_intervals > 768");
    if (n <= 0) return TQ_OK;
    hipStream_t s = as_stream(stream);
    long long done = 0;
    while (done < n) {
        long long chunk = n - done;
        const long long cap = st->max_blocks * (long long)kSB;
        if (chunk > cap) chunk = cap;
        const long long nb = (chunk + kSB - 1) / kSB;
        if (st->sample_dtype == TQ_F32)
            hipLaunchKernelGGL((k_vegas_accum_given<float>), dim3((unsigned)nb), dim3(kBlock), 0, s, *st,
                               (const float*)y + done * st->dim, (const float*)jf2 + done, chunk);
        else
            hipLaunchKernelGGL((k_vegas_accum_given<double>), dim3((unsigned)nb), dim3(kBlock), 0, s, *st,
                               (const double*)y + done * st->dim, (const double*)jf2 + done, chunk);
        rc = check_launch("k_vegas_accum_given");
        if (rc) return rc;
        const int W = st->dim * st->n_intervals;
        const int rb = (W + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL(k_vegas_map_reduce, dim3(rb), dim3(kBlock), 0, s, *st, chunk, 0, 1);
        rc = check_launch("k_vegas_map_reduce");
        if (rc) return rc;
        done += chunk;
    }
    return TQ_OK;
}

int tq_vegas_map_update_accumulated(const tq_vegas_state* st, void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    return launch_map_kernels(st, 1, as_stream(stream));
}

int tq_vegas_smooth(int32_t dim, int32_t n, const double* weights, const double* counts, double alpha, double* out,
                    double* scratch, int32_t* not_updatable, void* stream) {
    if (dim < 1 || dim > TQ_MAX_DIM || n < 2) return fail(TQ_E_ARG, "dim/n out of range");
    if (!weights || !counts || !out || !scratch || !not_updatable) return fail(TQ_E_NULL, "NULL argument");
    hipLaunchKernelGGL(k_vegas_smooth, dim3(1), dim3(64), 0, as_stream(stream), (int)dim, (int)n, weights, counts,
                       alpha, out, scratch, not_updatable);
    return check_launch("k_vegas_smooth");
}

int tq_vegas_strat_y(const tq_vegas_state* st, uint64_t seed, uint64_t offset, const void* u_in, void* y_out,
                     void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (!y_out) return fail(TQ_E_NULL, "y is NULL");
    const long long nb = sample_blocks(st, 0, 0);
    hipStream_t s = as_stream(stream);
    tq_integrand dummy{};
    dummy.family = TQ_FAM_CONSTANT;
    dummy.dim = st->dim;
    if (st->sample_dtype == TQ_F32) {
        SampleArgs<float> A{*st, nullptr, seed, offset, 0, 0, 0, 0, nullptr, nullptr, nullptr, (float*)y_out,
                            make_params<float>(dummy), (const float*)u_in};
        hipLaunchKernelGGL((k_vegas_sample<float, TQ_FAM_CONSTANT, MODE_Y, false>), dim3((unsigned)nb),
                           dim3(kBlock), 0, s, A);
    } else {
        SampleArgs<double> A{*st, nullptr, seed, offset, 0, 0, 0, 0, nullptr, nullptr, nullptr, (double*)y_out,
                             make_params<double>(dummy), (const double*)u_in};
        hipLaunchKernelGGL((k_vegas_sample<double, TQ_FAM_CONSTANT, MODE_Y, false>), dim3((unsigned)nb),
                           dim3(kBlock), 0, s, A);
    }
    return check_launch("k_vegas_sample(y)");
}

int tq_vegas_strat_accumulate(const tq_vegas_state* st, const double* w, int32_t it, void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (!w) return fail(TQ_E_NULL, "w is NULL");
    if (it < 0 || it >= TQ_VEGAS_MAX_ITERS) return fail(TQ_E_ARG, "iteration index out of range");
    return launch_strat(st, w, it, 0, as_stream(stream));
}

int tq_vegas_update_dh(const tq_vegas_state* st, int32_t it, void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (it < 0 || it >= TQ_VEGAS_MAX_ITERS) return fail(TQ_E_ARG, "iteration index out of range");
    hipStream_t s = as_stream(stream);
    const long long nb = (st->n_cubes + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_vegas_update_dh, dim3((unsigned)nb), dim3(kBlock), 0, s, *st);
    rc = check_launch("k_vegas_update_dh");
    if (rc) return rc;
    hipLaunchKernelGGL(k_vegas_strat_finalize, dim3(1), dim3(1024), 0, s, *st, nb, it);
    return check_launch("k_vegas_strat_finalize");
}

int tq_vegas_normalize_dh(const tq_vegas_state* st, int32_t it, void* stream) {
    int rc = check_state(st);
    if (rc) return rc;
    if (it < 0 || it >= TQ_VEGAS_MAX_ITERS) return fail(TQ_E_ARG, "iteration index out of range");
    hipLaunchKernelGGL(k_vegas_normalize, dim3(256), dim3(kBlock), 0, as_stream(stream), *st, it);
    return check_launch("k_vegas_normalize");
}

}  // extern "C"
