// Tensor-product grid building blocks shared by the single-integral and the
// batched kernels: per-axis maps, rule weights and the tile loop.
#pragma once
#include "tq_host.hpp"

namespace tq {

constexpr int kFlush = 128;      // FP32 partial length before FP64 flush (at most 128 terms)
constexpr int kLoopTable = 2048;  // longest one-axis loop kept in the LDS loop table

// Per-axis affine map.  Newton-Cotes: x_j = a + j*h (x_{n-1} = b exactly, as
// linspace).  Gauss-Legendre: x_j = alpha*r_j + beta (gaussian.py:159-162).
struct Axis {
    double a, h, b;
};

template <typename T>
__device__ inline Axis make_axis(const T* domain, int k, int n, int rule) {
    const double a = (double)domain[2 * k], b = (double)domain[2 * k + 1];
    Axis ax;
    if (rule == TQ_RULE_GAUSS_LEGENDRE) {
        ax.a = 0.5 * (b - a);
        ax.h = 0.5 * (a + b);
        ax.b = b;
    } else {
        ax.a = a;
        ax.h = (b - a) / (double)(n - 1);
        ax.b = b;
    }
    return ax;
}

inline Axis make_axis_host(double a, double b, int n, int rule) {
    Axis ax;
    if (rule == TQ_RULE_GAUSS_LEGENDRE) {
        ax.a = 0.5 * (b - a);
        ax.h = 0.5 * (a + b);
        ax.b = b;
    } else {
        ax.a = a;
        ax.h = (b - a) / (double)(n - 1);
        ax.b = b;
    }
    return ax;
}

__device__ inline double node_of(const Axis& ax, int j, int n, int rule, const double* glx) {
    if (rule == TQ_RULE_GAUSS_LEGENDRE) return fma(ax.a, glx[j], ax.h);
    return (j == n - 1) ? ax.b : fma((double)j, ax.h, ax.a);
}

// Composite Newton-Cotes weight pattern without the per-axis scale factor
// (trapezoid.py:34-36 -> h/2*[1,2,..,2,1]; simpson.py:35-45 -> h/3*[1,4,2,..,4,1];
// boole.py:35-47 -> h/22.5*[7,32,12,32,14,..,32,7]); Gauss-Legendre weight.
__device__ __host__ inline double weight_of(int rule, int j, int n, const double* glw) {
    const bool end = (j == 0) || (j == n - 1);
    switch (rule) {
        case -1: return 1.0;
        case TQ_RULE_TRAPEZOID: return end ? 1.0 : 2.0;
        case TQ_RULE_SIMPSON: return end ? 1.0 : ((j & 1) ? 4.0 : 2.0);
        case TQ_RULE_BOOLE: return end ? 7.0 : ((j & 1) ? 32.0 : (((j & 3) == 2) ? 12.0 : 14.0));
        default: return glw[j];
    }
}

// Per-axis scale so that integral = prod_k scale_k * sum W f.
__device__ __host__ inline double axis_scale(double a, double b, int n, int rule) {
    switch (rule) {
        case TQ_RULE_TRAPEZOID: return (b - a) / (double)(n - 1) / 2.0;
        case TQ_RULE_SIMPSON: return (b - a) / (double)(n - 1) / 3.0;
        case TQ_RULE_BOOLE: return (b - a) / (double)(n - 1) / 22.5;
        default: return 0.5 * (b - a);
    }
}

// Loop LDS table: entry j holds the loop's contribution to the integrand
// state at its point j (fold_c(zero, c_L, x_j) for one loop axis, both axes
// folded for two), then the point's weight; per point the tile loop only
// merges it into the row/lane state.  Families with a two-component state
// take three slots per entry.
//
// Families whose finish is exp2 of an additive state (Gaussian peak,
// exp-linear) or a reciprocal of a multiplicative one (product peak) carry
// the weight inside the state instead (Family::fold_key with the weight's
// key: + log2 w, or * 1/w), so an entry is the state slot alone and a point
// is merge, finish and add.  The table's LDS reads then move 4 bytes per point instead of 8 (the
// loop of the d=5 n=65 Boole Gaussian peak was bound by them: 142 -> 130 us
// per launch), at the cost of one more rounding of the entry (at most half an
// ulp of the exponent, the size of the rounding of state + entry itself).
template <int FAM>
struct StateSlots {
    static constexpr int v = (FAM == TQ_FAM_DOUBLE_GAUSSIAN) ? 2 : 1;
};
template <int FAM>
struct FoldsWeight {
    static constexpr bool v = FAM == TQ_FAM_GAUSSIAN_PEAK || FAM == TQ_FAM_EXP_LINEAR || FAM == TQ_FAM_PRODUCT_PEAK;
};
template <int FAM>
struct TabStride {
    static constexpr int v = StateSlots<FAM>::v + (FoldsWeight<FAM>::v ? 0 : 1);
};

__host__ __device__ inline int state_slots(int fam) { return fam == TQ_FAM_DOUBLE_GAUSSIAN ? 2 : 1; }
__host__ __device__ inline bool folds_weight(int fam) {
    return fam == TQ_FAM_GAUSSIAN_PEAK || fam == TQ_FAM_EXP_LINEAR || fam == TQ_FAM_PRODUCT_PEAK;
}
__host__ __device__ inline int tab_stride(int fam) { return state_slots(fam) + (folds_weight(fam) ? 0 : 1); }

// A table build's FP64 value of a weight: the weight, or its key when the
// family folds it into the state; the value of a product of weights.
template <typename T, int FAM>
__device__ inline double table_weight(double w) {
    if constexpr (FoldsWeight<FAM>::v) return Family<FAM, T>::weight_key(w);
    else return w;
}
template <typename T, int FAM>
__device__ inline double combine_weights(double a, double b) {
    if constexpr (FoldsWeight<FAM>::v) return Family<FAM, T>::combine_keys(a, b);
    else return a * b;
}

// Stores table entry j from its state and the table_weight value w.
template <typename T, int FAM>
__device__ inline void store_table_entry(T* tab, int j, const State<T>& s, double w) {
    constexpr int S = TabStride<FAM>::v;
    if constexpr (FoldsWeight<FAM>::v) {
        tab[S * j] = Family<FAM, T>::fold_key(s.s0, w);
    } else {
        tab[S * j] = s.s0;
        if (StateSlots<FAM>::v == 2) tab[S * j + 1] = s.s1;
        tab[S * j + S - 1] = (T)w;
    }
}

// Loop-table entry j when the loop runs over q consecutive axes L .. L+q-1
// (q = 1 or 2; for q = 2, j = j1 n + j2 with j1 on axis L): the axes'
// contributions folded together, then the product of their weights.
template <typename T, int FAM>
__device__ inline void loop_table_entry(T* tab, int j, const Axis* ax, int L, int q, int n, int rule,
                                        const double* glx, const double* glw, const Params<T>& p) {
    using F = Family<FAM, T>;
    const int j1 = q == 2 ? j / n : j;
    State<T> s = F::fold_c(F::zero(), axis_c(p, L), (T)node_of(ax[L], j1, n, rule, glx), p);
    double w = weight_of(rule, j1, n, glw);
    if (q == 2) {
        const int j2 = j - j1 * n;
        s = F::fold_c(s, axis_c(p, L + 1), (T)node_of(ax[L + 1], j2, n, rule, glx), p);
        w *= weight_of(rule, j2, n, glw);
    }
    store_table_entry<T, FAM>(tab, j, s, table_weight<T, FAM>(w));
}

// Two-axis loop-table entry j = j1 n + j2 from the nodes of its two axes
// (x1 on axis L, x2 on axis L + 1, already rounded to T as loop_table_entry
// rounds node_of) and their table_weight values: the same folds as
// loop_table_entry with q = 2, while the nodes and weights (or weight keys)
// are formed once per axis instead of once per entry.
template <typename T, int FAM>
__device__ inline void pair_table_entry(T* tab, int j, T x1, T x2, double w1, double w2, const AxisC<T>& c1,
                                        const AxisC<T>& c2, const Params<T>& p) {
    using F = Family<FAM, T>;
    State<T> s = F::fold_c(F::zero(), c1, x1, p);
    s = F::fold_c(s, c2, x2, p);
    store_table_entry<T, FAM>(tab, j, s, combine_weights<T, FAM>(w1, w2));
}

// Two-axis loop table (axes L and L + 1, n <= 32) built by whole waves: the
// calling wave stores entries first + lane, first + stride + lane, ... (one
// wave: first 0, stride 64; the waves of a block: first 64 w, stride 64 BW).
// Lane l < 32 forms node l of axis L and lane 32 + l node l of axis L + 1,
// and each entry takes its two nodes and weights from those lanes
// (ds_bpermute).  Every lane runs every trip, so the lanes read from are
// active; only entries j < n^2 are stored.
template <typename T, int FAM>
__device__ inline void pair_table_wave(T* tab, const Axis* ax, int L, int n, int rule, const double* glx,
                                       const double* glw, const Params<T>& p, int first = 0, int stride = kWave) {
    const int lane = threadIdx.x & (kWave - 1);
    const int jn = lane & 31;
    const bool has = jn < n;
    const int js = has ? jn : 0;  // in-range index for the lanes past n
    const T xn = (T)node_of(ax[lane < 32 ? L : L + 1], js, n, rule, glx);
    const double wn = table_weight<T, FAM>(weight_of(rule, js, n, glw));
    const T xa = has ? xn : T(0);
    const double wa = has ? wn : 0.0;
    const AxisC<T> c1 = axis_c(p, L), c2 = axis_c(p, L + 1);
    const int nl = n * n;
    int j1 = (first + lane) / n, j2 = first + lane - j1 * n;
    const int s1 = stride / n, s2 = stride - s1 * n;
    for (int jb = first; jb < nl; jb += stride) {
        const T x1 = __shfl(xa, j1 & 31), x2 = __shfl(xa, 32 + j2);
        const double w1 = __shfl(wa, j1 & 31), w2 = __shfl(wa, 32 + j2);
        if (jb + lane < nl) pair_table_entry<T, FAM>(tab, jb + lane, x1, x2, w1, w2, c1, c2, p);
        j1 += s1;
        j2 += s2;
        if (j2 >= n) {
            j2 -= n;
            ++j1;
        }
    }
}

// Row-axis tables (ROWTAB): for the row axes k < L, entry k*n + j holds the
// axis's contribution at node j (the state slots of loop_table_entry) in `rt` and
// its weight in FP64 in `rw`, so a row's state and weight are merged from
// the tables instead of being recomputed from the axis maps.
template <typename T, int FAM>
__device__ inline void row_table_entry(T* rt, double* rw, int k, int j, const Axis& ax, const AxisC<T>& c, int n,
                                       int rule, const double* glx, const double* glw, const Params<T>& p) {
    using F = Family<FAM, T>;
    constexpr int S1 = StateSlots<FAM>::v;
    const State<T> s = F::fold_c(F::zero(), c, (T)node_of(ax, j, n, rule, glx), p);
    const int e = k * n + j;
    rt[S1 * e] = s.s0;
    if (S1 == 2) rt[S1 * e + 1] = s.s1;
    rw[e] = weight_of(rule, j, n, glw);
}

// One tile of the tensor-product index space.  m == 0 (one axis): node
// 64 t + lane for loop items t in [t_begin, t_end); with GL_FOLD, a
// Gauss-Legendre axis instead walks the lower half of its symmetric table,
// item t, lane l evaluating node k = 64 t + l < ceil(n/2) and its mirror
// n - 1 - k (x_{n-1-k} = -x_k, w_{n-1-k} = w_k), so each table load serves
// two points.  Otherwise the lanes of the calling wave
// take lane combinations lw*64 .. lw*64+63 of the last m axes, the wave walks
// loop items [t_begin, t_end) (row axes [0, L) x a loop of nl points over
// the axes between the row and lane axes: one axis, nl = n, or two, nl = n^2,
// which needs the table).  Returns the lane's weighted sum.  `ax` holds the
// per-axis maps, `tab` the loop table (loop_table_entry) when TABLE.
// Code paths a grid_tile instantiation carries: one-axis integrals (m == 0),
// multi-axis ones (m > 0), or both.  The batched kernels run the two kinds in
// separate launches, each without the other's code and live registers; an
// integral sent to an instantiation without its path yields NaN.
constexpr int kPathsOneAxis = 1;
constexpr int kPathsMultiAxis = 2;
constexpr int kPathsAll = 3;

template <typename T, int FAM, bool TABLE, bool ROWTAB = false, bool GL_FOLD = true, int PATHS = kPathsAll>
__device__ inline double grid_tile(int d, int n, int nl, int m, int L, int rule, int lane_count, const Params<T>& p,
                                   const Axis* ax, const T* tab, const double* glx, const double* glw, int lw,
                                   long long t_begin, long long t_end, const T* rt = nullptr,
                                   const double* rw = nullptr, const float* gl_pairs32 = nullptr) {
    using F = Family<FAM, T>;
    const int lane = threadIdx.x & (kWave - 1);
    double acc = 0.0;
    if constexpr ((PATHS & kPathsOneAxis) != 0) {
        if (m == 0 && GL_FOLD && rule == TQ_RULE_GAUSS_LEGENDRE) {
            const State<T> s0 = F::start(p);
            const AxisC<T> ac0 = axis_c(p, 0);
            const Axis a0 = ax[0];
            const int h = (n + 1) >> 1;  // nodes 0 .. h-1; for odd n node h-1 is the middle one, counted once
            for (long long tb = t_begin; tb < t_end; tb += kFlush) {
                const long long te = (tb + kFlush < t_end) ? (tb + kFlush) : t_end;
                T fa = T(0);
                if (gl_pairs32) {
                    // FP32 (node, weight) pairs: one 8-byte load per two points
                    // (FP32 results only; the node and weight carry 2^-24
                    // relative rounding, far inside the FP32 tolerance)
                    const float2* pairs = reinterpret_cast<const float2*>(gl_pairs32);
                    const float sa = (float)a0.a, sh = (float)a0.h;
    #pragma unroll 2
                    for (long long t = tb; t < te; ++t) {
                        const int k = (int)(t * kWave) + lane;
                        const float2 e = pairs[k < h ? k : h - 1];
                        const T w1 = (k < h) ? (T)e.y : T(0);
                        const T w2 = (k < h && 2 * k + 1 != n) ? (T)e.y : T(0);
                        const T x1 = (T)fmaf(sa, e.x, sh);
                        const T x2 = (T)fmaf(-sa, e.x, sh);
                        fa = fma(w1, F::finish(F::fold_c(s0, ac0, x1, p), p), fa);
                        fa = fma(w2, F::finish(F::fold_c(s0, ac0, x2, p), p), fa);
                    }
                } else {
    #pragma unroll 2
                    for (long long t = tb; t < te; ++t) {
                        const int k = (int)(t * kWave) + lane;
                        const int kk = k < h ? k : h - 1;
                        const double xn = glx[kk], wk = glw[kk];
                        const T w1 = (k < h) ? (T)wk : T(0);
                        const T w2 = (k < h && 2 * k + 1 != n) ? (T)wk : T(0);
                        const T x1 = (T)fma(a0.a, xn, a0.h);
                        const T x2 = (T)fma(-a0.a, xn, a0.h);
                        fa = fma(w1, F::finish(F::fold_c(s0, ac0, x1, p), p), fa);
                        fa = fma(w2, F::finish(F::fold_c(s0, ac0, x2, p), p), fa);
                    }
                }
                acc += (double)fa;
            }
            return acc;
        }
        if (m == 0) {
            // one axis split into rows of 64 consecutive nodes: node k = 64 t + lane
            const State<T> s0 = F::start(p);
            const AxisC<T> ac0 = axis_c(p, 0);
            const Axis a0 = ax[0];
            // Items holding an end node (t = 0 and the item of node n - 1, which
            // may also run past n) take the general body: clamped index, zero
            // weight past n, end weights.
            const long long t_end_item = (long long)(n - 1) / kWave;
            auto edge_item = [&](long long t) -> double {
                const long long k = t * kWave + lane;
                const int kk = (int)(k < n ? k : n - 1);
                const double w = (k < n) ? weight_of(rule, kk, n, glw) : 0.0;
                const T f = F::finish(F::fold_c(s0, ac0, (T)node_of(a0, kk, n, rule, glx), p), p);
                return w * (double)f;
            };
            long long ib = t_begin, ie = t_end;  // interior items [ib, ie)
            if (ib == 0 && ie > 0) {
                acc += edge_item(0);
                ib = 1;
            }
            if (ie > t_end_item && t_end_item >= ib) {
                acc += edge_item(t_end_item);
                ie = t_end_item;
            }
            if (rule == TQ_RULE_GAUSS_LEGENDRE) {
                for (long long tb = ib; tb < ie; tb += kFlush) {
                    const long long te = (tb + kFlush < ie) ? (tb + kFlush) : ie;
                    T fa = T(0);
    #pragma unroll 4
                    for (long long t = tb; t < te; ++t) {
                        const int k = (int)(t * kWave) + lane;
                        const T x = (T)fma(a0.a, glx[k], a0.h);
                        fa = fma((T)glw[k], F::finish(F::fold_c(s0, ac0, x, p), p), fa);
                    }
                    acc += (double)fa;
                }
            } else {
                // Newton-Cotes interior weights repeat with period 4 in the node
                // index, and 64 t + lane = lane (mod 4): one weight per lane.
                const double w_lane = weight_of(rule, kWave + lane, 0x7fffffff, glw);
                for (long long tb = ib; tb < ie; tb += kFlush) {
                    const long long te = (tb + kFlush < ie) ? (tb + kFlush) : ie;
                    double kd = (double)(tb * kWave + lane);
                    T fa = T(0);
    #pragma unroll 4
                    for (long long t = tb; t < te; ++t) {
                        const T x = (T)fma(kd, a0.h, a0.a);
                        kd += (double)kWave;
                        fa += F::finish(F::fold_c(s0, ac0, x, p), p);
                    }
                    acc += w_lane * (double)fa;
                }
            }
            return acc;
        }
    } else {
        if (m == 0) return __longlong_as_double(0x7ff8000000000000ll);
    }
    if constexpr ((PATHS & kPathsMultiAxis) == 0) return __longlong_as_double(0x7ff8000000000000ll);
    const int li = lw * kWave + lane;
    const bool active = li < lane_count;

    // lane axes: fold and weight once
    State<T> s_lane = F::zero();
    double w_lane = active ? 1.0 : 0.0;
    {
        int rem = active ? li : 0;
        for (int k = d - 1; k >= d - m; --k) {
            const int j = rem % n;
            rem /= n;
            s_lane = F::fold_c(s_lane, axis_c(p, k), (T)node_of(ax[k], j, n, rule, glx), p);
            w_lane *= weight_of(rule, j, n, glw);
        }
    }
    if (L < 0) {
        if (t_begin == 0 && t_end > 0) {
            const T f = F::finish(F::merge(F::start(p), s_lane), p);
            acc = w_lane * (double)f;
        }
        return acc;
    }
    const AxisC<T> acL = axis_c(p, L);
    long long t = t_begin;
    int j;
    // Row odometer: lane k of `dig` holds the digit of row axis k (k < L <
    // 16), read with readlane at a uniform k.  Digits kept as a register array
    // with its loops unrolled over all 16 possible axes made the compiler
    // hoist 32 "k < L" masks out of the row loop, spilling SGPRs into a VGPR
    // and raising the kernel to 67 VGPRs (6 waves per SIMD instead of 8).
    int dig = 0;
    if (t < 0x7fffffffLL) {
        // 32-bit divisions when the item index fits
        const unsigned un = (unsigned)n, unl = (unsigned)nl;
        unsigned r = (unsigned)t / unl;
        j = (int)((unsigned)t - r * unl);
        for (int k = L - 1; k >= 0; --k) {
            const unsigned q = r / un;
            dig = (lane == k) ? (int)(r - q * un) : dig;
            r = q;
        }
    } else {
        long long r = t / nl;
        j = (int)(t - r * nl);
        for (int k = L - 1; k >= 0; --k) {
            const long long q = r / n;
            dig = (lane == k) ? (int)(r - q * n) : dig;
            r = q;
        }
    }
    while (t < t_end) {
        State<T> s_row = F::start(p);
        double w_row = w_lane;
        for (int k = 0; k < L; ++k) {
            const int dk = __builtin_amdgcn_readlane(dig, k);
            if (ROWTAB) {
                constexpr int S1 = StateSlots<FAM>::v;
                const int e = k * n + dk;
                State<T> term;
                term.s0 = rt[S1 * e];
                term.s1 = (S1 == 2) ? rt[S1 * e + 1] : T(0);
                s_row = F::merge(s_row, term);
                w_row *= rw[e];
            } else {
                s_row = F::fold_c(s_row, axis_c(p, k), (T)node_of(ax[k], dk, n, rule, glx), p);
                w_row *= weight_of(rule, dk, n, glw);
            }
        }
        const State<T> s_rl = F::merge(s_row, s_lane);
        const long long left = t_end - t;
        const int jend = (left < (long long)(nl - j)) ? (j + (int)left) : nl;
        double racc = 0.0;
        for (int jb = j; jb < jend; jb += kFlush) {
            const int je = (jb + kFlush < jend) ? (jb + kFlush) : jend;
            T fa = T(0);
            // 16 points per trip: the wave-uniform table address is moved to
            // a VGPR once per 16 points (ds_read2 offsets cover the rest), so
            // a point costs the add, the exp and the FMA plus 1/16 of a move
            // (d=5 n=65: 157 -> 149 us per launch against 4 points per trip)
#pragma unroll 16
            for (int jj = jb; jj < je; ++jj) {
                if (TABLE) {
                    constexpr int S = TabStride<FAM>::v;
                    State<T> term;
                    term.s0 = tab[S * jj];
                    term.s1 = (StateSlots<FAM>::v == 2) ? tab[S * jj + 1] : T(0);
                    const T f = F::finish(F::merge(s_rl, term), p);
                    if constexpr (FoldsWeight<FAM>::v)
                        fa += f;  // the weight is inside f
                    else
                        fa = fma(tab[S * jj + S - 1], f, fa);
                } else {
                    const T xL = (T)node_of(ax[L], jj, n, rule, glx);
                    const T wL = (T)weight_of(rule, jj, n, glw);
                    fa = fma(wL, F::finish(F::fold_c(s_rl, acL, xL, p), p), fa);
                }
            }
            racc += (double)fa;
        }
        acc = fma(w_row, racc, acc);
        t += jend - j;
        j = jend;
        if (j == nl) {
            j = 0;
            for (int k = L - 1; k >= 0; --k) {
                const int dk = __builtin_amdgcn_readlane(dig, k) + 1;
                const bool wrap = dk >= n;
                dig = (lane == k) ? (wrap ? 0 : dk) : dig;
                if (!wrap) break;
            }
        }
    }
    return acc;
}

// ---- host-side planning ---------------------------------------------------
struct GridGeom {
    int d, n, m, L;       // L: row axes [0, L), loop axes [L, L + q)
    int q, nl;            // loop axes (1 or 2) and loop length n^q
    int rule;
    int lane_count;       // n^m
    int lane_waves;       // ceil(lane_count / 64)
    int use_table;        // loop axis nodes/weights in LDS
    int gl_fold;          // m == 0 Gauss-Legendre: items cover nodes [0, ceil(n/2)) and their mirrors
    long long items;      // n^(d-m) (1 when there is no loop axis)
    long long ipw;        // items per wave
    long long chunks;     // ceil(items / ipw)
    long long num_waves;  // lane_waves * chunks
    long long blocks;
};

inline int ipow_checked(long long base, int e, long long limit, long long* out) {
    long long r = 1;
    for (int i = 0; i < e; ++i) {
        if (r > limit / base) return -1;
        r *= base;
    }
    *out = r;
    return 0;
}

// Largest n for which the planner gives up a lane axis to merge two loop
// axes (the choice of lane axes, and so the launch size, then does not
// depend on merge caps of 1024 or more).
constexpr int kMergeMaxN = 32;

// Splits the n^d index space of one integral (see grid_tile).  lane_axes /
// items_per_wave override the automatic choice (0 = automatic).  When
// merge_cap > 0, n^2 <= merge_cap (the caller's loop-table capacity) and
// there are at least two axes between the row axes and the lane axes, the
// loop runs over the last two non-lane axes (n^2 points per row, one table
// entry per point pair), so a short axis does not pay the per-row cost
// every n points.  trade_lane_axis = false keeps the lane axes (and so the
// launch size) of the unmerged plan.  gl_fold: a one-axis (m == 0)
// Gauss-Legendre plan covers half the table (grid_tile's GL_FOLD).
inline int plan_geom(int d, int n, int rule, int lane_axes, long long items_per_wave, long long target_waves,
                     GridGeom& G, long long merge_cap = 0, bool trade_lane_axis = true, bool gl_fold = true) {
    G.d = d;
    G.n = n;
    G.rule = rule;
    G.q = 1;
    G.nl = n;
    G.gl_fold = 0;
    int m = 0;
    if (lane_axes > 0) {
        m = lane_axes < d ? lane_axes : d;
    } else if (d == 1 && n >= 128) {
        m = 0;  // rows of 64 consecutive nodes
    } else {
        // lane axes: the smallest m whose n^m lane combinations fill >= 90% of
        // the 64-lane waves (else the best filling m); a loop axis is kept
        // whenever d > 1
        const int mmax = d > 1 ? d - 1 : 1;
        m = mmax;
        double best = -1.0;
        long long c = 1;
        for (int k = 1; k <= mmax; ++k) {
            if (c > (1LL << 31) / n) break;
            c *= n;
            const double util = (double)c / (double)(((c + kWave - 1) / kWave) * kWave);
            if (c >= 32 && util >= 0.9) {
                m = k;
                break;
            }
            if (util > best) {
                best = util;
                m = k;
            }
        }
        if (d == 1) m = 1;
        // One lane axis fewer when that lets the loop run over two short
        // axes (n^2 points per row instead of n) and the lanes still fill
        // three quarters of their waves.
        if (trade_lane_axis && merge_cap > 0 && d >= 3 && d - m == 1 && m >= 2 && n <= kMergeMaxN &&
            (long long)n * n <= merge_cap) {
            long long c1 = 1;
            for (int k = 0; k < m - 1; ++k) c1 *= n;
            const double util = (double)c1 / (double)(((c1 + kWave - 1) / kWave) * kWave);
            if (util >= 0.75) m -= 1;
        }
    }
    long long items = 1;
    if (m == 0) {
        G.m = 0;
        G.lane_count = kWave;
        G.lane_waves = 1;
        G.L = 0;
        G.gl_fold = (gl_fold && rule == TQ_RULE_GAUSS_LEGENDRE) ? 1 : 0;
        const long long nodes = G.gl_fold ? ((long long)n + 1) / 2 : (long long)n;
        items = (nodes + kWave - 1) / kWave;
    } else {
        long long lc;
        if (ipow_checked(n, m, (1LL << 31) - 64, &lc) != 0) return fail(TQ_E_ARG, "lane combinations exceed 2^31");
        G.m = m;
        G.lane_count = (int)lc;
        G.lane_waves = (int)((lc + kWave - 1) / kWave);
        G.L = d - m - 1;
        if (G.L >= 0) ipow_checked(n, d - m, 1LL << 62, &items);
        if (merge_cap > 0 && G.L >= 1 && (long long)n * n <= merge_cap) {
            G.q = 2;
            G.nl = n * n;
            G.L -= 1;
        }
    }
    G.items = items;
    G.use_table = (G.m > 0 && G.L >= 0 && (G.nl <= kLoopTable || G.q == 2)) ? 1 : 0;
    long long ipw = items_per_wave;
    if (ipw <= 0) {
        const long long work = items * (long long)G.lane_waves;
        ipw = (work + target_waves - 1) / target_waves;
        const long long min_ipw = items < 16 ? items : 16;
        if (ipw < min_ipw) ipw = min_ipw;
    }
    if (ipw < 1) ipw = 1;
    while (((items + ipw - 1) / ipw) * (long long)G.lane_waves > (1LL << 30)) ipw *= 2;
    G.ipw = ipw;
    G.chunks = (items + ipw - 1) / ipw;
    G.num_waves = G.chunks * (long long)G.lane_waves;
    G.blocks = (G.num_waves + kWavesPerBlock - 1) / kWavesPerBlock;
    return TQ_OK;
}

}  // namespace tq
