// Parameter sweeps of the Gaussian peak f(x) = exp(-sum_k a_k^2 (x_k-u_k)^2),
// the oscillatory cos(2 pi w + sum_k c_k x_k) and the product peak
// prod_k 1/(a_k^-2 + (x_k-u_k)^2) over B integrals that share rule,
// dimension and n but have their own parameters and box [lo_b, hi_b] --
// forward value and the hand-written backward.  The Gaussian peak's sums
// follow; those of the other two families are given with their kernels.
//
// Every rule maps node j of axis k as x = alpha_k r_j + beta_k with
// alpha = (hi-lo)/2, beta = (hi+lo)/2 and r_j in [-1, 1] (the Gauss-Legendre
// root, or 2j/(n-1) - 1 for Newton-Cotes), and integral = S * sum W f with
// S proportional to prod_k alpha_k.  With t = x - u and, per integral,
//   F = sum W f,  E_k = sum W f t_k,  Q_k = sum W f t_k^2,  R_k = sum W f t_k r_k,
// the gradients are
//   dI/du_k = 2 a_k^2 S E_k,    dI/da_k = -2 a_k S Q_k,
//   dI/dbeta_k = -2 a_k^2 S E_k,  dI/dalpha_k = I/alpha_k - 2 a_k^2 S R_k,
//   dI/dlo_k = (dI/dbeta_k - dI/dalpha_k)/2,  dI/dhi_k = (dI/dbeta_k + dI/dalpha_k)/2.
// The backward kernel walks the same tiles as the forward one and forms
// these sums per point in FP32 (flushed to FP64 at least every 128 points).
#include "tq_grid_tile.hpp"

namespace tq {

constexpr int kSweepMaxD = 8;

struct SweepGeom {
    GridGeom g;
    long long blocks_x;  // blocks per integral
};

// Waves per launch when items_per_wave is left to the planner: about 16
// per SIMD over the whole sweep, so each wave's fixed cost (parameters,
// loop-axis table, 1 + 3 dim block sums) is spread over many points.
constexpr long long kSweepWaves = 16384;

// The backward tile walks one loop axis.  The forward kernel may loop over
// two (tables up to kSweepMergeCap entries) with the same lane axes, so both
// plans have the same waves and the backward's workspace covers both.
constexpr long long kSweepMergeCap = 1024;

static int sweep_geom(const tq_grid_desc* gd, long long B, SweepGeom& S, long long merge_cap = 0) {
    const long long per_integral = B > 0 ? (kSweepWaves + B - 1) / B : kSweepWaves;
    // the backward's one-axis rows walk every node, so the forward does too
    int rc = plan_geom(gd->dim, gd->n, gd->rule, gd->lane_axes, gd->items_per_wave, per_integral, S.g, merge_cap,
                       false, false);
    if (rc) return rc;
    S.blocks_x = (S.g.num_waves + kWavesPerBlock - 1) / kWavesPerBlock;
    return TQ_OK;
}

// Parameters of integral b in the family's Params layout (as make_params
// builds them for one integrand): Gaussian peak a[b][d], u[b][d]; product
// peak a[b][d], u[b][d] (b_k = a_k^-2); oscillatory c[b][d] in a and w[b]
// in u (FP32 phase in revolutions, FP64 in radians).
template <typename T, int FAM>
__device__ inline void load_sweep_params(int d, int n, int rule, long long b, const T* a, const T* u, const T* dom,
                                         Params<T>* p, Axis* ax) {
    const double sl2e = 1.2011224087864498;  // sqrt(log2 e), as in make_params
    const int tid = threadIdx.x;
    if (tid < TQ_MAX_DIM) {
        T pa = T(0), pb = T(0), pu = T(0);
        if (tid < d) {
            const double ak = (double)a[b * d + tid];
            if (FAM == TQ_FAM_GAUSSIAN_PEAK) {
                const double uk = (double)u[b * d + tid];
                pa = (T)(ak * sl2e);
                pb = (T)(-ak * uk * sl2e);
                pu = (T)uk;
            } else if (FAM == TQ_FAM_PRODUCT_PEAK) {
                pa = (T)ak;
                pb = (T)(ak != 0.0 ? 1.0 / (ak * ak) : 0.0);
                pu = u[b * d + tid];
            } else {  // oscillatory
                pa = (T)(sizeof(T) == 4 ? ak / (2.0 * M_PI) : ak);
            }
            ax[tid] = make_axis(dom + b * 2 * d, tid, n, rule);
        }
        p->a[tid] = pa;
        p->b[tid] = pb;
        p->u[tid] = pu;
    }
    if (tid == 0) {
        p->c0 = T(1);
        p->c1 = p->c3 = T(0);
        p->c2 = T(0);
        if (FAM == TQ_FAM_OSCILLATORY) {
            const double w = (double)u[b];
            p->c2 = (T)(sizeof(T) == 4 ? w : 2.0 * M_PI * w);
        }
    }
}

constexpr int kSweepRowTab = 1024;  // row-axis table entries (L*n) kept in LDS
constexpr int kBwdTable = 1024;     // largest loop axis tabulated by the backward kernel

// Dynamic LDS: [row weights: L*n FP64][loop table: S nl T][row terms: L*n T]
// (every sweep family has one state slot; S = TabStride: 1 when the family
// folds the weight into the entry, else 2)
template <typename T, int FAM>
inline size_t sweep_fwd_lds(const GridGeom& g, bool table, bool rowtab) {
    size_t b = table ? (size_t)TabStride<FAM>::v * g.nl * sizeof(T) : 0;
    if (rowtab) b += (size_t)g.L * g.n * (sizeof(double) + sizeof(T));
    return b;
}
static_assert(StateSlots<TQ_FAM_GAUSSIAN_PEAK>::v == 1 && StateSlots<TQ_FAM_PRODUCT_PEAK>::v == 1 &&
                  StateSlots<TQ_FAM_OSCILLATORY>::v == 1,
              "sweep families keep one state slot");

// BW waves per block: 16 with a two-axis loop (one block per integral
// builds the n^2-entry table once), else 4.
template <typename T, int FAM, bool TABLE, bool ROWTAB, int BW>
__global__ __launch_bounds__(BW * kWave) void k_sweep_fwd(SweepGeom S, const T* a, const T* u, const T* dom,
                                                         const double* glx, const double* glw, double* partials) {
    constexpr int NT = BW * kWave;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ Params<T> s_p;
    __shared__ Axis s_axis[TQ_MAX_DIM];
    __shared__ double s_red[BW];
    double* s_rw = reinterpret_cast<double*>(smem_raw);
    T* s_tab = reinterpret_cast<T*>(smem_raw + (ROWTAB ? (size_t)S.g.L * S.g.n * sizeof(double) : 0));
    T* s_rt = s_tab + (size_t)TabStride<FAM>::v * S.g.nl;
    const GridGeom& g = S.g;
    const long long b = blockIdx.y;
    load_sweep_params<T, FAM>(g.d, g.n, g.rule, b, a, u, dom, &s_p, s_axis);
    __syncthreads();
    if (TABLE) {
        if (g.q == 2 && g.n <= 32)
            pair_table_wave<T, FAM>(s_tab, s_axis, g.L, g.n, g.rule, glx, glw, s_p,
                                    __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * kWave, NT);
        else
            for (int j = threadIdx.x; j < g.nl; j += NT)
                loop_table_entry<T, FAM>(s_tab, j, s_axis, g.L, g.q, g.n, g.rule, glx, glw, s_p);
        if (ROWTAB)
            for (int e = threadIdx.x; e < g.L * g.n; e += NT) {
                const int k = e / g.n;
                row_table_entry<T, FAM>(s_rt, s_rw, k, e - k * g.n, s_axis[k], axis_c(s_p, k), g.n, g.rule, glx,
                                        glw, s_p);
            }
        __syncthreads();
    }
    const long long wave = (long long)blockIdx.x * BW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double acc = 0.0;
    if (wave < g.num_waves) {
        const int lw = (int)(wave % g.lane_waves);
        const long long chunk = wave / g.lane_waves;
        const long long t0 = chunk * g.ipw;
        const long long t1 = (t0 + g.ipw < g.items) ? (t0 + g.ipw) : g.items;
        acc = grid_tile<T, FAM, TABLE, ROWTAB, false>(g.d, g.n, g.nl, g.m, g.L, g.rule, g.lane_count, s_p, s_axis,
                                                      s_tab, glx, glw, lw, t0, t1, s_rt, s_rw);
    }
    const double bs = block_sum(acc, s_red);
    if (threadIdx.x == 0) partials[b * gridDim.x + blockIdx.x] = bs;
}

// normalised node coordinate r in [-1, 1]
__device__ inline double node_r(int j, int n, int rule, const double* glx) {
    if (rule == TQ_RULE_GAUSS_LEGENDRE) return glx[j];
    return (n > 1) ? 2.0 * (double)j / (double)(n - 1) - 1.0 : 0.0;
}

// One loop-axis table entry of the backward kernel: exponent term
// -(a x + b)^2 + log2(w) in log2 units (the node weight folded into the
// exponent, so w f is one exp2) and the normalised node r.  The loop axis's
// sums of W f t, W f t^2 and W f t r follow from the moments of r, since
// t = x - u = alpha r + gamma on the axis (see loop_axis_affine): the loop
// keeps sum W f r and sum W f r^2, and its LDS reads move 8 bytes per point
// instead of the 16 of (term, t, t^2, t r) entries.
template <typename T>
struct alignas(2 * sizeof(T)) BwdEntry {
    T term, r;
};

// x = alpha r + beta on an axis (r the normalised node of node_r).
__device__ inline void loop_axis_affine(const Axis& ax, int n, int rule, double* alpha, double* beta) {
    if (rule == TQ_RULE_GAUSS_LEGENDRE) {
        *alpha = ax.a;
        *beta = ax.h;
    } else {
        *alpha = 0.5 * (ax.b - ax.a);
        *beta = 0.5 * (ax.b + ax.a);
    }
}

// Backward tile for D axes: flat walk of the tile's points (lane axes per
// lane, the rest decoded per row).  Only the loop axis is accumulated per
// point; row axes are accumulated per row and lane axes once per tile.
template <typename T, int D, bool TABLE>
__device__ inline void sweep_bwd_tile(const GridGeom& g, const Params<T>& p, const Axis* ax, const BwdEntry<T>* tab,
                                      const double* glx, const double* glw, int lw, long long t_begin,
                                      long long t_end, double* accF, double* accE, double* accQ, double* accR) {
    const int n = g.n, m = g.m, L = g.L, rule = g.rule;
    const int lane = threadIdx.x & (kWave - 1);
    // per-point coordinates of the current point, per axis
    T xk[D], rk[D];
    if (m == 0) {
        // single axis split: node 64 t + lane
        for (long long t = t_begin; t < t_end; ++t) {
            const long long k = t * kWave + lane;
            if (k >= n) continue;
            const int kk = (int)k;
            const T x = (T)node_of(ax[0], kk, n, rule, glx);
            const T w = (T)weight_of(rule, kk, n, glw);
            const T r = (T)node_r(kk, n, rule, glx);
            const T tt = x - p.u[0];
            const T e = fma(x, p.a[0], p.b[0]);
            const T f = tq_exp2(-e * e);
            const T wf = w * f;
            accF[0] += (double)wf;
            accE[0] += (double)(wf * tt);
            accQ[0] += (double)(wf * tt * tt);
            accR[0] += (double)(wf * tt * r);
        }
        return;
    }
    const int li = lw * kWave + lane;
    const bool active = li < g.lane_count;
    T w_lane = active ? T(1) : T(0);
    {
        int rem = active ? li : 0;
#pragma unroll
        for (int k = D - 1; k >= 0; --k) {
            xk[k] = T(0);
            rk[k] = T(0);
            if (k >= D - m) {
                const int j = rem % n;
                rem /= n;
                xk[k] = (T)node_of(ax[k], j, n, rule, glx);
                rk[k] = (T)node_r(j, n, rule, glx);
                w_lane *= (T)weight_of(rule, j, n, glw);
            }
        }
    }
    if (L < 0) {
        // every axis is a lane axis: one point per lane
        if (t_begin == 0 && t_end > 0) {
            T s = T(0);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const T e = fma(xk[k], p.a[k], p.b[k]);
                s = fma(-e, e, s);
            }
            const T wf = w_lane * tq_exp2(s);
            accF[0] += (double)wf;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const T tt = xk[k] - p.u[k];
                accE[k] += (double)(wf * tt);
                accQ[k] += (double)(wf * tt * tt);
                accR[k] += (double)(wf * tt * rk[k]);
            }
        }
        return;
    }
    long long t = t_begin;
    long long row = t / n;
    int j = (int)(t - row * n);
    int rdig[D];
    {
        long long r = row;
#pragma unroll
        for (int k = D - 1; k >= 0; --k) {
            rdig[k] = 0;
            if (k < L) {
                rdig[k] = (int)(r % n);
                r /= n;
            }
        }
    }
    T aL = T(0), bL = T(0), s_lane = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        if (k == L) {
            aL = p.a[k];
            bL = p.b[k];
        }
        if (k > L) {
            const T e = fma(xk[k], p.a[k], p.b[k]);
            s_lane = fma(-e, e, s_lane);
        }
    }
    // Along a row every axis but L is fixed, so for k != L the sums over the
    // row factor: sum W f t_k = t_k F_row, sum W f t_k^2 = t_k^2 F_row,
    // sum W f t_k r_k = t_k r_k F_row.  Lane axes (k > L) keep their node for
    // the whole tile, so theirs are t_k F_tile etc., formed after the walk.
    double F_tile = 0.0, M1 = 0.0, M2 = 0.0;  // loop-axis sums of W f r and W f r^2
    while (t < t_end) {
        double w_row = (double)w_lane;
        T s_fix = s_lane;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (k < L) {
                xk[k] = (T)node_of(ax[k], rdig[k], n, rule, glx);
                rk[k] = (T)node_r(rdig[k], n, rule, glx);
                w_row *= weight_of(rule, rdig[k], n, glw);
                const T e = fma(xk[k], p.a[k], p.b[k]);
                s_fix = fma(-e, e, s_fix);
            }
        }
        const long long left = t_end - t;
        const int jend = (left < (long long)(n - j)) ? (j + (int)left) : n;
        double Frow = 0.0;
        for (int jb = j; jb < jend; jb += kFlush) {
            const int je = (jb + kFlush < jend) ? (jb + kFlush) : jend;
            T fF = T(0), f1 = T(0), f2 = T(0);
#pragma unroll 4
            for (int jj = jb; jj < je; ++jj) {
                T wf, r;
                if (TABLE) {
                    // one add, one exp2, then F, W f r and W f r^2
                    const BwdEntry<T> en = tab[jj];
                    wf = tq_exp2(s_fix + en.term);
                    r = en.r;
                } else {
                    const T xL = (T)node_of(ax[L], jj, n, rule, glx);
                    const T e = fma(xL, aL, bL);
                    const T wL = (T)weight_of(rule, jj, n, glw);
                    wf = wL * tq_exp2(s_fix - e * e);
                    r = (T)node_r(jj, n, rule, glx);
                }
                fF += wf;
                const T wr = wf * r;
                f1 += wr;
                f2 = fma(wr, r, f2);
            }
            Frow += w_row * (double)fF;
            M1 += w_row * (double)f1;
            M2 += w_row * (double)f2;
        }
        F_tile += Frow;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (k < L) {
                const double tt = (double)(xk[k] - p.u[k]);
                const double Ft = Frow * tt;
                accE[k] += Ft;
                accQ[k] += Ft * tt;
                accR[k] += Ft * (double)rk[k];
            }
        }
        t += jend - j;
        j = jend;
        if (j == n) {
            j = 0;
#pragma unroll
            for (int k = D - 1; k >= 0; --k) {
                if (k < L) {
                    if (++rdig[k] < n) break;
                    rdig[k] = 0;
                }
            }
        }
    }
    accF[0] += F_tile;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        if (k == L) {
            // t = alpha r + gamma on the loop axis: sums of W f t, W f t^2 and
            // W f t r from F_tile = sum W f and the moments M1, M2 of r
            double alpha, beta;
            loop_axis_affine(ax[L], n, rule, &alpha, &beta);
            const double gam = beta - (double)p.u[L];
            accE[k] += alpha * M1 + gam * F_tile;
            accQ[k] += alpha * alpha * M2 + 2.0 * alpha * gam * M1 + gam * gam * F_tile;
            accR[k] += alpha * M2 + gam * M1;
        }
        if (k > L) {
            const double tt = (double)(xk[k] - p.u[k]);
            const double Ft = F_tile * tt;
            accE[k] += Ft;
            accQ[k] += Ft * tt;
            accR[k] += Ft * (double)rk[k];
        }
    }
}

// partial layout per (integral, block): [F, E_0..E_{D-1}, Q_0.., R_0..]
template <typename T, int D, bool TABLE>
__global__ __launch_bounds__(kBlock) void k_sweep_bwd(SweepGeom S, const T* a, const T* u, const T* dom,
                                                     const double* glx, const double* glw, double* partials) {
    // loop-axis table in dynamic LDS, one BwdEntry per node
    extern __shared__ __align__(16) unsigned char smem_raw[];
    BwdEntry<T>* s_tab = reinterpret_cast<BwdEntry<T>*>(smem_raw);
    __shared__ Params<T> s_p;
    __shared__ Axis s_axis[TQ_MAX_DIM];
    __shared__ double s_red[kWavesPerBlock];
    const GridGeom& g = S.g;
    const long long b = blockIdx.y;
    load_sweep_params<T, TQ_FAM_GAUSSIAN_PEAK>(g.d, g.n, g.rule, b, a, u, dom, &s_p, s_axis);
    __syncthreads();
    if (TABLE) {
        const int L = g.L;
        for (int j = threadIdx.x; j < g.n; j += kBlock) {
            const T x = (T)node_of(s_axis[L], j, g.n, g.rule, glx);
            const T e = fma(x, s_p.a[L], s_p.b[L]);
            // composite-rule and Gauss-Legendre weights are positive
            const double w = weight_of(g.rule, j, g.n, glw);
            BwdEntry<T> en;
            en.term = (T)((double)(-e * e) + log2(w));
            en.r = (T)node_r(j, g.n, g.rule, glx);
            s_tab[j] = en;
        }
        __syncthreads();
    }
    double F[1] = {0.0}, E[D], Q[D], R[D];
#pragma unroll
    for (int k = 0; k < D; ++k) E[k] = Q[k] = R[k] = 0.0;
    const long long wave = (long long)blockIdx.x * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wave < g.num_waves) {
        const int lw = (int)(wave % g.lane_waves);
        const long long chunk = wave / g.lane_waves;
        const long long t0 = chunk * g.ipw;
        const long long t1 = (t0 + g.ipw < g.items) ? (t0 + g.ipw) : g.items;
        sweep_bwd_tile<T, D, TABLE>(g, s_p, s_axis, s_tab, glx, glw, lw, t0, t1, F, E, Q, R);
    }
    const int nv = 1 + 3 * D;
    double* out = partials + (b * gridDim.x + blockIdx.x) * nv;
    double v = block_sum(F[0], s_red);
    if (threadIdx.x == 0) out[0] = v;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        v = block_sum(E[k], s_red);
        if (threadIdx.x == 0) out[1 + k] = v;
        v = block_sum(Q[k], s_red);
        if (threadIdx.x == 0) out[1 + D + k] = v;
        v = block_sum(R[k], s_red);
        if (threadIdx.x == 0) out[1 + 2 * D + k] = v;
    }
}

// ---------------------------------------------------------------------------
// Backward of the oscillatory family, f = cos(theta), theta = 2 pi w +
// sum_k c_k x_k.  Per integral
//   F = sum W cos theta,  G = sum W sin theta,  R_k = sum W sin theta r_k,
// and, with x_k = alpha_k r_k + beta_k,
//   dI/dw = -2 pi S G,  dI/dc_k = -S (alpha_k R_k + beta_k G),
//   dI/dbeta_k = -c_k S G,  dI/dalpha_k = I/alpha_k - c_k S R_k.
// Along a row only the loop axis varies, so R_k of a row axis is r_k G_row
// and of a lane axis r_k G_tile; the loop axis is accumulated per point.
// Loop-axis table entry per node: the phase term c_L x_j (in the state's
// units), w_j and w_j r_j.
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(4 * sizeof(T) < 16 ? 4 * sizeof(T) : 16) OscEntry {
    T phi, w, wr, pad;
};

// cos and sin of a phase in the family's state units (FP32: revolutions)
__device__ inline void osc_cos_sin(float s, float* c, float* sn) {
    *c = tq_cos_rev(s);
    *sn = tq_cos_rev(s - 0.25f);  // sin(2 pi s) = cos(2 pi (s - 1/4))
}
__device__ inline void osc_cos_sin(double s, double* c, double* sn) { sincos(s, sn, c); }

template <typename T, int D, bool TABLE>
__device__ inline void sweep_bwd_tile_osc(const GridGeom& g, const Params<T>& p, const Axis* ax,
                                          const OscEntry<T>* tab, const double* glx, const double* glw, int lw,
                                          long long t_begin, long long t_end, double* accF, double* accG,
                                          double* accR) {
    const int n = g.n, m = g.m, L = g.L, rule = g.rule;
    const int lane = threadIdx.x & (kWave - 1);
    if (m == 0) {
        for (long long t = t_begin; t < t_end; ++t) {
            const long long k = t * kWave + lane;
            if (k >= n) continue;
            const int kk = (int)k;
            const T x = (T)node_of(ax[0], kk, n, rule, glx);
            const T w = (T)weight_of(rule, kk, n, glw);
            const T r = (T)node_r(kk, n, rule, glx);
            T c, sn;
            osc_cos_sin(fma(p.a[0], x, p.c2), &c, &sn);
            accF[0] += (double)(w * c);
            accG[0] += (double)(w * sn);
            accR[0] += (double)(w * sn * r);
        }
        return;
    }
    T rk[D];
    const int li = lw * kWave + lane;
    const bool active = li < g.lane_count;
    double w_lane = active ? 1.0 : 0.0;
    T ph_lane = T(0);
    {
        int rem = active ? li : 0;
#pragma unroll
        for (int k = D - 1; k >= 0; --k) {
            rk[k] = T(0);
            if (k >= D - m) {
                const int j = rem % n;
                rem /= n;
                const T x = (T)node_of(ax[k], j, n, rule, glx);
                rk[k] = (T)node_r(j, n, rule, glx);
                ph_lane = fma(p.a[k], x, ph_lane);
                w_lane *= weight_of(rule, j, n, glw);
            }
        }
    }
    if (L < 0) {
        if (t_begin == 0 && t_end > 0) {
            T c, sn;
            osc_cos_sin(p.c2 + ph_lane, &c, &sn);
            accF[0] += w_lane * (double)c;
            const double gs = w_lane * (double)sn;
            accG[0] += gs;
#pragma unroll
            for (int k = 0; k < D; ++k) accR[k] += gs * (double)rk[k];
        }
        return;
    }
    long long t = t_begin;
    long long row = t / n;
    int j = (int)(t - row * n);
    int rdig[D];
    {
        long long r = row;
#pragma unroll
        for (int k = D - 1; k >= 0; --k) {
            rdig[k] = 0;
            if (k < L) {
                rdig[k] = (int)(r % n);
                r /= n;
            }
        }
    }
    T aL = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k)
        if (k == L) aL = p.a[k];
    double F_tile = 0.0, G_tile = 0.0, RL = 0.0;
    while (t < t_end) {
        double w_row = w_lane;
        T ph_fix = p.c2 + ph_lane;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (k < L) {
                const T x = (T)node_of(ax[k], rdig[k], n, rule, glx);
                rk[k] = (T)node_r(rdig[k], n, rule, glx);
                w_row *= weight_of(rule, rdig[k], n, glw);
                ph_fix = fma(p.a[k], x, ph_fix);
            }
        }
        const long long left = t_end - t;
        const int jend = (left < (long long)(n - j)) ? (j + (int)left) : n;
        double Frow = 0.0, Grow = 0.0;
        for (int jb = j; jb < jend; jb += kFlush) {
            const int je = (jb + kFlush < jend) ? (jb + kFlush) : jend;
            T fF = T(0), fG = T(0), fR = T(0);
#pragma unroll 4
            for (int jj = jb; jj < je; ++jj) {
                T c, sn, w, wr;
                if (TABLE) {
                    const OscEntry<T> en = tab[jj];
                    osc_cos_sin(ph_fix + en.phi, &c, &sn);
                    w = en.w;
                    wr = en.wr;
                } else {
                    const T xL = (T)node_of(ax[L], jj, n, rule, glx);
                    osc_cos_sin(fma(aL, xL, ph_fix), &c, &sn);
                    w = (T)weight_of(rule, jj, n, glw);
                    wr = w * (T)node_r(jj, n, rule, glx);
                }
                fF = fma(w, c, fF);
                fG = fma(w, sn, fG);
                fR = fma(wr, sn, fR);
            }
            Frow += w_row * (double)fF;
            Grow += w_row * (double)fG;
            RL += w_row * (double)fR;
        }
        F_tile += Frow;
        G_tile += Grow;
        // selects rather than stores under k-vs-L branches: those the
        // compiler merges into one store at a run-time index, which puts
        // the array in scratch memory
#pragma unroll
        for (int k = 0; k < D; ++k) accR[k] += (k < L) ? Grow * (double)rk[k] : 0.0;
        t += jend - j;
        j = jend;
        if (j == n) {
            j = 0;
#pragma unroll
            for (int k = D - 1; k >= 0; --k) {
                if (k < L) {
                    if (++rdig[k] < n) break;
                    rdig[k] = 0;
                }
            }
        }
    }
    accF[0] += F_tile;
    accG[0] += G_tile;
#pragma unroll
    for (int k = 0; k < D; ++k) accR[k] += (k == L) ? RL : ((k > L) ? G_tile * (double)rk[k] : 0.0);
}

// ---------------------------------------------------------------------------
// Backward of the product peak, f = prod_k g_k, g_k = 1/(a_k^-2 + t_k^2),
// t_k = x_k - u_k.  Per integral
//   F = sum W f,  P_k = sum W f t_k g_k,  H_k = sum W f g_k,
//   M_k = sum W f t_k g_k r_k,
// and dI/du_k = 2 S P_k,  dI/da_k = 2 a_k^-3 S H_k,  dI/dbeta_k = -2 S P_k,
// dI/dalpha_k = I/alpha_k - 2 S M_k.  A row's fixed factor (weights and g
// of the other axes) multiplies the loop axis's per-point terms, which the
// table holds with the node weight folded in: v_j = w_j g_j, v_j t_j g_j,
// v_j g_j, v_j t_j g_j r_j.
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(4 * sizeof(T) < 16 ? 4 * sizeof(T) : 16) PpEntry {
    T v, vtg, vg, vtgr;
};

template <typename T, int D, bool TABLE>
__device__ inline void sweep_bwd_tile_pp(const GridGeom& g, const Params<T>& p, const Axis* ax, const PpEntry<T>* tab,
                                         const double* glx, const double* glw, int lw, long long t_begin,
                                         long long t_end, double* accF, double* accP, double* accH, double* accM) {
    const int n = g.n, m = g.m, L = g.L, rule = g.rule;
    const int lane = threadIdx.x & (kWave - 1);
    if (m == 0) {
        for (long long t = t_begin; t < t_end; ++t) {
            const long long k = t * kWave + lane;
            if (k >= n) continue;
            const int kk = (int)k;
            const T x = (T)node_of(ax[0], kk, n, rule, glx);
            const T w = (T)weight_of(rule, kk, n, glw);
            const T r = (T)node_r(kk, n, rule, glx);
            const T tt = x - p.u[0];
            const T gk = T(1) / fma(tt, tt, p.b[0]);
            const T wf = w * gk;
            accF[0] += (double)wf;
            accP[0] += (double)(wf * tt * gk);
            accH[0] += (double)(wf * gk);
            accM[0] += (double)(wf * tt * gk * r);
        }
        return;
    }
    T tk[D], gk[D], rk[D];
    const int li = lw * kWave + lane;
    const bool active = li < g.lane_count;
    double w_lane = active ? 1.0 : 0.0;
    T g_lane = T(1);
    {
        int rem = active ? li : 0;
#pragma unroll
        for (int k = D - 1; k >= 0; --k) {
            tk[k] = gk[k] = rk[k] = T(0);
            if (k >= D - m) {
                const int j = rem % n;
                rem /= n;
                const T x = (T)node_of(ax[k], j, n, rule, glx);
                rk[k] = (T)node_r(j, n, rule, glx);
                tk[k] = x - p.u[k];
                gk[k] = T(1) / fma(tk[k], tk[k], p.b[k]);
                g_lane *= gk[k];
                w_lane *= weight_of(rule, j, n, glw);
            }
        }
    }
    if (L < 0) {
        if (t_begin == 0 && t_end > 0) {
            const double wf = w_lane * (double)g_lane;
            accF[0] += wf;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double tg = (double)(tk[k] * gk[k]);
                accP[k] += wf * tg;
                accH[k] += wf * (double)gk[k];
                accM[k] += wf * tg * (double)rk[k];
            }
        }
        return;
    }
    long long t = t_begin;
    long long row = t / n;
    int j = (int)(t - row * n);
    int rdig[D];
    {
        long long r = row;
#pragma unroll
        for (int k = D - 1; k >= 0; --k) {
            rdig[k] = 0;
            if (k < L) {
                rdig[k] = (int)(r % n);
                r /= n;
            }
        }
    }
    T uL = T(0), bL = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k)
        if (k == L) {
            uL = p.u[k];
            bL = p.b[k];
        }
    double F_tile = 0.0, PL = 0.0, HL = 0.0, ML = 0.0;
    while (t < t_end) {
        double fix = w_lane * (double)g_lane;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (k < L) {
                const T x = (T)node_of(ax[k], rdig[k], n, rule, glx);
                rk[k] = (T)node_r(rdig[k], n, rule, glx);
                tk[k] = x - p.u[k];
                gk[k] = T(1) / fma(tk[k], tk[k], p.b[k]);
                fix *= weight_of(rule, rdig[k], n, glw) * (double)gk[k];
            }
        }
        const long long left = t_end - t;
        const int jend = (left < (long long)(n - j)) ? (j + (int)left) : n;
        double Frow = 0.0;
        for (int jb = j; jb < jend; jb += kFlush) {
            const int je = (jb + kFlush < jend) ? (jb + kFlush) : jend;
            T fF = T(0), fP = T(0), fH = T(0), fM = T(0);
#pragma unroll 4
            for (int jj = jb; jj < je; ++jj) {
                if (TABLE) {
                    const PpEntry<T> en = tab[jj];
                    fF += en.v;
                    fP += en.vtg;
                    fH += en.vg;
                    fM += en.vtgr;
                } else {
                    const T xL = (T)node_of(ax[L], jj, n, rule, glx);
                    const T tL = xL - uL;
                    const T gL = T(1) / fma(tL, tL, bL);
                    const T v = (T)weight_of(rule, jj, n, glw) * gL;
                    const T vtg = v * tL * gL;
                    fF += v;
                    fP += vtg;
                    fH += v * gL;
                    fM = fma(vtg, (T)node_r(jj, n, rule, glx), fM);
                }
            }
            Frow += fix * (double)fF;
            PL += fix * (double)fP;
            HL += fix * (double)fH;
            ML += fix * (double)fM;
        }
        F_tile += Frow;
        // selects, not stores under k-vs-L branches (see the oscillatory tile)
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double fk = (k < L) ? Frow : 0.0;
            const double tg = (double)(tk[k] * gk[k]);
            accP[k] += fk * tg;
            accH[k] += fk * (double)gk[k];
            accM[k] += fk * tg * (double)rk[k];
        }
        t += jend - j;
        j = jend;
        if (j == n) {
            j = 0;
#pragma unroll
            for (int k = D - 1; k >= 0; --k) {
                if (k < L) {
                    if (++rdig[k] < n) break;
                    rdig[k] = 0;
                }
            }
        }
    }
    accF[0] += F_tile;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double fk = (k > L) ? F_tile : 0.0;
        const double tg = (double)(tk[k] * gk[k]);
        accP[k] += (k == L) ? PL : fk * tg;
        accH[k] += (k == L) ? HL : fk * (double)gk[k];
        accM[k] += (k == L) ? ML : fk * tg * (double)rk[k];
    }
}

// Backward kernels of the oscillatory family and the product peak.  Partial
// layout per (integral, block): oscillatory [F, G, R_0..R_{D-1}] (2 + D),
// product peak [F, P_0.., H_0.., M_0..] (1 + 3 D).
template <typename T, int FAM, int D, bool TABLE>
__global__ __launch_bounds__(kBlock) void k_sweep_bwd_fam(SweepGeom S, const T* a, const T* u, const T* dom,
                                                         const double* glx, const double* glw, double* partials) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ Params<T> s_p;
    __shared__ Axis s_axis[TQ_MAX_DIM];
    __shared__ double s_red[kWavesPerBlock];
    const GridGeom& g = S.g;
    const long long b = blockIdx.y;
    load_sweep_params<T, FAM>(g.d, g.n, g.rule, b, a, u, dom, &s_p, s_axis);
    __syncthreads();
    if (TABLE) {
        const int L = g.L;
        for (int j = threadIdx.x; j < g.n; j += kBlock) {
            const T x = (T)node_of(s_axis[L], j, g.n, g.rule, glx);
            const T w = (T)weight_of(g.rule, j, g.n, glw);
            const T r = (T)node_r(j, g.n, g.rule, glx);
            if constexpr (FAM == TQ_FAM_OSCILLATORY) {
                OscEntry<T> en;
                en.phi = s_p.a[L] * x;
                en.w = w;
                en.wr = w * r;
                en.pad = T(0);
                reinterpret_cast<OscEntry<T>*>(smem_raw)[j] = en;
            } else {
                const T tt = x - s_p.u[L];
                const T gk = T(1) / fma(tt, tt, s_p.b[L]);
                PpEntry<T> en;
                en.v = w * gk;
                en.vtg = en.v * tt * gk;
                en.vg = en.v * gk;
                en.vtgr = en.vtg * r;
                reinterpret_cast<PpEntry<T>*>(smem_raw)[j] = en;
            }
        }
        __syncthreads();
    }
    const long long wave = (long long)blockIdx.x * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool has_tile = wave < g.num_waves;
    const int lw = has_tile ? (int)(wave % g.lane_waves) : 0;
    const long long t0 = has_tile ? (wave / g.lane_waves) * g.ipw : 0;
    const long long t1 = has_tile ? ((t0 + g.ipw < g.items) ? (t0 + g.ipw) : g.items) : 0;
    // separate register arrays per sum (one array indexed through offsets
    // was kept in scratch memory)
    if constexpr (FAM == TQ_FAM_OSCILLATORY) {
        double F[1] = {0.0}, G[1] = {0.0}, R[D];
#pragma unroll
        for (int k = 0; k < D; ++k) R[k] = 0.0;
        if (has_tile)
            sweep_bwd_tile_osc<T, D, TABLE>(g, s_p, s_axis, reinterpret_cast<const OscEntry<T>*>(smem_raw), glx, glw,
                                            lw, t0, t1, F, G, R);
        double* out = partials + (b * gridDim.x + blockIdx.x) * (2 + D);
        double v = block_sum(F[0], s_red);
        if (threadIdx.x == 0) out[0] = v;
        v = block_sum(G[0], s_red);
        if (threadIdx.x == 0) out[1] = v;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            v = block_sum(R[k], s_red);
            if (threadIdx.x == 0) out[2 + k] = v;
        }
    } else {
        double F[1] = {0.0}, P[D], H[D], M[D];
#pragma unroll
        for (int k = 0; k < D; ++k) P[k] = H[k] = M[k] = 0.0;
        if (has_tile)
            sweep_bwd_tile_pp<T, D, TABLE>(g, s_p, s_axis, reinterpret_cast<const PpEntry<T>*>(smem_raw), glx, glw,
                                           lw, t0, t1, F, P, H, M);
        double* out = partials + (b * gridDim.x + blockIdx.x) * (1 + 3 * D);
        double v = block_sum(F[0], s_red);
        if (threadIdx.x == 0) out[0] = v;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            v = block_sum(P[k], s_red);
            if (threadIdx.x == 0) out[1 + k] = v;
            v = block_sum(H[k], s_red);
            if (threadIdx.x == 0) out[1 + D + k] = v;
            v = block_sum(M[k], s_red);
            if (threadIdx.x == 0) out[1 + 2 * D + k] = v;
        }
    }
}

template <typename T>
__device__ inline double sweep_scale(const T* dom, long long b, int d, int n, int rule) {
    double s = 1.0;
    for (int k = 0; k < d; ++k) s *= axis_scale((double)dom[(b * d + k) * 2], (double)dom[(b * d + k) * 2 + 1], n, rule);
    return s;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sweep_fwd_finalize(const double* partials, long long nbx, long long B,
                                                              const T* dom, int d, int n, int rule, T* out,
                                                              double* out64) {
    const long long b = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (b >= B) return;
    double v = 0.0;
    for (long long j = lane; j < nbx; j += kWave) v += partials[b * nbx + j];
    v = wave_sum(v);
    if (lane == 0) {
        const double r = v * sweep_scale(dom, b, d, n, rule);
        if (out) out[b] = (T)r;
        if (out64) out64[b] = r;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sweep_bwd_finalize(const double* partials, long long nbx, long long B,
                                                              int D, const T* a, const T* dom, int n, int rule,
                                                              const T* grad_out, T* ga, T* gu, T* gdom) {
    const long long b = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (b >= B) return;
    const int nv = 1 + 3 * D;
    // lanes reduce one sum each (nv <= 25 < 64), blocks in order
    double v = 0.0;
    if (lane < nv)
        for (long long j = 0; j < nbx; ++j) v += partials[(b * nbx + j) * nv + lane];
    const double S = sweep_scale(dom, b, D, n, rule);
    const double F = __shfl(v, 0, kWave);
    const double I = S * F;
    const double g = (double)grad_out[b];
    for (int k = 0; k < D; ++k) {
        const double E = __shfl(v, 1 + k, kWave);
        const double Q = __shfl(v, 1 + D + k, kWave);
        const double R = __shfl(v, 1 + 2 * D + k, kWave);
        if (lane == 0) {
            const double ak = (double)a[b * D + k];
            const double lo = (double)dom[(b * D + k) * 2], hi = (double)dom[(b * D + k) * 2 + 1];
            const double alpha = 0.5 * (hi - lo);
            const double dbeta = -2.0 * ak * ak * S * E;
            const double dalpha = I / alpha - 2.0 * ak * ak * S * R;
            if (gu) gu[b * D + k] = (T)(g * 2.0 * ak * ak * S * E);
            if (ga) ga[b * D + k] = (T)(g * -2.0 * ak * S * Q);
            if (gdom) {
                gdom[(b * D + k) * 2] = (T)(g * 0.5 * (dbeta - dalpha));
                gdom[(b * D + k) * 2 + 1] = (T)(g * 0.5 * (dbeta + dalpha));
            }
        }
    }
}

// Gradients of the oscillatory family (a = c, u = w: grad_u gets dI/dw per
// integral) and of the product peak from their block partials.
template <typename T, int FAM>
__global__ __launch_bounds__(kBlock) void k_sweep_bwd_finalize_fam(const double* partials, long long nbx, long long B,
                                                                  int D, const T* a, const T* dom, int n, int rule,
                                                                  const T* grad_out, T* ga, T* gu, T* gdom) {
    const long long b = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (b >= B) return;
    const int nv = FAM == TQ_FAM_OSCILLATORY ? 2 + D : 1 + 3 * D;
    // lanes reduce one sum each (nv <= 25 < 64), blocks in order
    double v = 0.0;
    if (lane < nv)
        for (long long j = 0; j < nbx; ++j) v += partials[(b * nbx + j) * nv + lane];
    const double S = sweep_scale(dom, b, D, n, rule);
    const double F = __shfl(v, 0, kWave);
    const double I = S * F;
    const double g = (double)grad_out[b];
    if constexpr (FAM == TQ_FAM_OSCILLATORY) {
        const double G = __shfl(v, 1, kWave);
        if (lane == 0 && gu) gu[b] = (T)(g * -2.0 * M_PI * S * G);
        for (int k = 0; k < D; ++k) {
            const double R = __shfl(v, 2 + k, kWave);
            if (lane == 0) {
                const double ck = (double)a[b * D + k];
                const double lo = (double)dom[(b * D + k) * 2], hi = (double)dom[(b * D + k) * 2 + 1];
                const double alpha = 0.5 * (hi - lo), beta = 0.5 * (hi + lo);
                const double dbeta = -ck * S * G;
                const double dalpha = I / alpha - ck * S * R;
                if (ga) ga[b * D + k] = (T)(g * -S * (alpha * R + beta * G));
                if (gdom) {
                    gdom[(b * D + k) * 2] = (T)(g * 0.5 * (dbeta - dalpha));
                    gdom[(b * D + k) * 2 + 1] = (T)(g * 0.5 * (dbeta + dalpha));
                }
            }
        }
    } else {
        for (int k = 0; k < D; ++k) {
            const double P = __shfl(v, 1 + k, kWave);
            const double H = __shfl(v, 1 + D + k, kWave);
            const double M = __shfl(v, 1 + 2 * D + k, kWave);
            if (lane == 0) {
                const double ak = (double)a[b * D + k];
                const double lo = (double)dom[(b * D + k) * 2], hi = (double)dom[(b * D + k) * 2 + 1];
                const double alpha = 0.5 * (hi - lo);
                const double dbeta = -2.0 * S * P;
                const double dalpha = I / alpha - 2.0 * S * M;
                if (gu) gu[b * D + k] = (T)(g * 2.0 * S * P);
                if (ga) ga[b * D + k] = (T)(g * 2.0 * S * H / (ak * ak * ak));
                if (gdom) {
                    gdom[(b * D + k) * 2] = (T)(g * 0.5 * (dbeta - dalpha));
                    gdom[(b * D + k) * 2 + 1] = (T)(g * 0.5 * (dbeta + dalpha));
                }
            }
        }
    }
}

static int sweep_values(int family, int D) { return family == TQ_FAM_OSCILLATORY ? 2 + D : 1 + 3 * D; }

static int check_family(int family) {
    if (family != TQ_FAM_GAUSSIAN_PEAK && family != TQ_FAM_OSCILLATORY && family != TQ_FAM_PRODUCT_PEAK)
        return fail(TQ_E_UNSUPPORTED, "sweeps support the Gaussian peak, oscillatory and product peak families");
    return TQ_OK;
}

static int check_sweep(const tq_grid_desc* g, int64_t B) {
    if (!g) return fail(TQ_E_NULL, "grid descriptor is NULL");
    if (g->dtype != TQ_F32 && g->dtype != TQ_F64) return fail(TQ_E_ARG, "dtype must be TQ_F32 or TQ_F64");
    if (g->dim < 1 || g->dim > kSweepMaxD) return fail(TQ_E_ARG, "sweep dim out of range (1..8)");
    if (B < 1 || B > 65535) return fail(TQ_E_ARG, "sweep count out of range (1..65535)");
    if (g->n < 1 || (g->rule != TQ_RULE_GAUSS_LEGENDRE && g->n < 2)) return fail(TQ_E_ARG, "n out of range");
    if (g->rule == TQ_RULE_SIMPSON && g->n % 2 != 1) return fail(TQ_E_ARG, "Simpson needs odd n");
    if (g->rule == TQ_RULE_BOOLE && (g->n - 1) % 4 != 0) return fail(TQ_E_ARG, "Boole needs n = 1 + 4k");
    if (g->rule == TQ_RULE_GAUSS_LEGENDRE && (!g->gl_nodes || !g->gl_weights))
        return fail(TQ_E_NULL, "Gauss-Legendre needs node/weight tables");
    long long P;
    if (ipow_checked(g->n, g->dim, 1LL << 62, &P) != 0) return fail(TQ_E_ARG, "n^dim too large");
    return TQ_OK;
}

template <typename T, int FAM, bool TABLE, bool ROWTAB, int BW>
static void launch_fwd_kernel(const SweepGeom& S, int64_t B, const void* a, const void* u, const void* dom,
                              const tq_grid_desc* gd, void* ws, size_t shmem, unsigned gx, hipStream_t s) {
    hipLaunchKernelGGL((k_sweep_fwd<T, FAM, TABLE, ROWTAB, BW>), dim3(gx, (unsigned)B), dim3(BW * kWave), shmem, s,
                       S, (const T*)a, (const T*)u, (const T*)dom, gd->gl_nodes, gd->gl_weights, (double*)ws);
}

template <typename T, int FAM>
static int launch_sweep_fwd(const tq_grid_desc* gd, int64_t B, const void* a, const void* u, const void* dom,
                            void* ws, size_t wsb, void* out, double* out64, hipStream_t s) {
    SweepGeom S;
    int rc = sweep_geom(gd, B, S, kSweepMergeCap);
    if (rc) return rc;
    if (!ws || wsb < (size_t)(S.blocks_x * B) * sizeof(double)) return fail(TQ_E_WORKSPACE, "sweep workspace too small");
    if (S.blocks_x > 0x7fffffff) return fail(TQ_E_ARG, "too many blocks");
    const bool wide = S.g.q == 2;
    const int bw = wide ? 16 : kWavesPerBlock;
    const unsigned gx = (unsigned)((S.g.num_waves + bw - 1) / bw);
    if (S.g.use_table && S.g.L >= 1 && (long long)S.g.L * S.g.n <= kSweepRowTab) {
        const size_t sh = sweep_fwd_lds<T, FAM>(S.g, true, true);
        if (wide) launch_fwd_kernel<T, FAM, true, true, 16>(S, B, a, u, dom, gd, ws, sh, gx, s);
        else launch_fwd_kernel<T, FAM, true, true, kWavesPerBlock>(S, B, a, u, dom, gd, ws, sh, gx, s);
    } else if (S.g.use_table) {
        const size_t sh = sweep_fwd_lds<T, FAM>(S.g, true, false);
        if (wide) launch_fwd_kernel<T, FAM, true, false, 16>(S, B, a, u, dom, gd, ws, sh, gx, s);
        else launch_fwd_kernel<T, FAM, true, false, kWavesPerBlock>(S, B, a, u, dom, gd, ws, sh, gx, s);
    } else {
        launch_fwd_kernel<T, FAM, false, false, kWavesPerBlock>(S, B, a, u, dom, gd, ws, 0, gx, s);
    }
    rc = check_launch("k_sweep_fwd");
    if (rc) return rc;
    const long long fb = (B + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL((k_sweep_fwd_finalize<T>), dim3((unsigned)fb), dim3(kBlock), 0, s, (const double*)ws,
                       (long long)gx, (long long)B, (const T*)dom, gd->dim, gd->n, gd->rule, (T*)out, out64);
    return check_launch("k_sweep_fwd_finalize");
}

template <typename T, int FAM, int D>
static int launch_bwd_d(const SweepGeom& S, int64_t B, const void* a, const void* u, const void* dom,
                        const tq_grid_desc* gd, void* ws, hipStream_t s) {
    dim3 grid((unsigned)S.blocks_x, (unsigned)B);
    if constexpr (FAM != TQ_FAM_GAUSSIAN_PEAK) {
        const size_t entry = FAM == TQ_FAM_OSCILLATORY ? sizeof(OscEntry<T>) : sizeof(PpEntry<T>);
        if (S.g.m > 0 && S.g.L >= 0 && S.g.n <= kBwdTable)
            hipLaunchKernelGGL((k_sweep_bwd_fam<T, FAM, D, true>), grid, dim3(kBlock), (size_t)S.g.n * entry, s, S,
                               (const T*)a, (const T*)u, (const T*)dom, gd->gl_nodes, gd->gl_weights, (double*)ws);
        else
            hipLaunchKernelGGL((k_sweep_bwd_fam<T, FAM, D, false>), grid, dim3(kBlock), 0, s, S, (const T*)a,
                               (const T*)u, (const T*)dom, gd->gl_nodes, gd->gl_weights, (double*)ws);
        return check_launch("k_sweep_bwd_fam");
    } else {
    if (S.g.m > 0 && S.g.L >= 0 && S.g.n <= kBwdTable)
        hipLaunchKernelGGL((k_sweep_bwd<T, D, true>), grid, dim3(kBlock), (size_t)S.g.n * sizeof(BwdEntry<T>), s, S,
                           (const T*)a, (const T*)u, (const T*)dom, gd->gl_nodes, gd->gl_weights, (double*)ws);
    else
        hipLaunchKernelGGL((k_sweep_bwd<T, D, false>), grid, dim3(kBlock), 0, s, S, (const T*)a, (const T*)u,
                           (const T*)dom, gd->gl_nodes, gd->gl_weights, (double*)ws);
    return check_launch("k_sweep_bwd");
    }
}

template <typename T, int FAM>
static int launch_sweep_bwd(const tq_grid_desc* gd, int64_t B, const void* a, const void* u, const void* dom,
                            const void* grad_out, void* ws, size_t wsb, void* ga, void* gu, void* gdom,
                            hipStream_t s) {
    SweepGeom S;
    int rc = sweep_geom(gd, B, S);
    if (rc) return rc;
    const int D = gd->dim;
    if (!ws || wsb < (size_t)(S.blocks_x * B * sweep_values(FAM, D)) * sizeof(double))
        return fail(TQ_E_WORKSPACE, "sweep workspace too small");
    switch (D) {
        case 1: rc = launch_bwd_d<T, FAM, 1>(S, B, a, u, dom, gd, ws, s); break;
        case 2: rc = launch_bwd_d<T, FAM, 2>(S, B, a, u, dom, gd, ws, s); break;
        case 3: rc = launch_bwd_d<T, FAM, 3>(S, B, a, u, dom, gd, ws, s); break;
        case 4: rc = launch_bwd_d<T, FAM, 4>(S, B, a, u, dom, gd, ws, s); break;
        case 5: rc = launch_bwd_d<T, FAM, 5>(S, B, a, u, dom, gd, ws, s); break;
        case 6: rc = launch_bwd_d<T, FAM, 6>(S, B, a, u, dom, gd, ws, s); break;
        case 7: rc = launch_bwd_d<T, FAM, 7>(S, B, a, u, dom, gd, ws, s); break;
        default: rc = launch_bwd_d<T, FAM, 8>(S, B, a, u, dom, gd, ws, s); break;
    }
    if (rc) return rc;
    const long long fb = (B + kWavesPerBlock - 1) / kWavesPerBlock;
    if constexpr (FAM == TQ_FAM_GAUSSIAN_PEAK) {
        hipLaunchKernelGGL((k_sweep_bwd_finalize<T>), dim3((unsigned)fb), dim3(kBlock), 0, s, (const double*)ws,
                           S.blocks_x, (long long)B, D, (const T*)a, (const T*)dom, gd->n, gd->rule,
                           (const T*)grad_out, (T*)ga, (T*)gu, (T*)gdom);
    } else {
        hipLaunchKernelGGL((k_sweep_bwd_finalize_fam<T, FAM>), dim3((unsigned)fb), dim3(kBlock), 0, s,
                           (const double*)ws, S.blocks_x, (long long)B, D, (const T*)a, (const T*)dom, gd->n,
                           gd->rule, (const T*)grad_out, (T*)ga, (T*)gu, (T*)gdom);
    }
    return check_launch("k_sweep_bwd_finalize");
}

// Expands the body once per sweep family with FAM bound (family checked before).
#define TQ_SWEEP_FAMILY(fam, ...)                                                                   \
    switch (fam) {                                                                                  \
        case TQ_FAM_GAUSSIAN_PEAK: { constexpr int FAM = TQ_FAM_GAUSSIAN_PEAK; __VA_ARGS__; } break; \
        case TQ_FAM_OSCILLATORY: { constexpr int FAM = TQ_FAM_OSCILLATORY; __VA_ARGS__; } break;     \
        default: { constexpr int FAM = TQ_FAM_PRODUCT_PEAK; __VA_ARGS__; } break;                    \
    }

}  // namespace tq

using namespace tq;

extern "C" {

size_t tq_sweep_workspace_size(const tq_grid_desc* g, int64_t B) {
    if (check_sweep(g, B)) return 0;
    SweepGeom S;
    if (sweep_geom(g, B, S)) return 0;
    // the largest partial layout of the sweep families (1 + 3 dim values)
    return (size_t)(S.blocks_x * B * (1 + 3 * g->dim)) * sizeof(double);
}

int tq_sweep_forward(const tq_grid_desc* g, int32_t family, int64_t B, const void* a, const void* u,
                     const void* domains, void* workspace, size_t workspace_bytes, void* out, double* out64,
                     void* stream) {
    int rc = check_sweep(g, B);
    if (!rc) rc = check_family(family);
    if (rc) return rc;
    if (!a || !u || !domains) return fail(TQ_E_NULL, "a/u/domains is NULL");
    if (!out && !out64) return fail(TQ_E_NULL, "no output pointer");
    TQ_SWEEP_FAMILY(family, {
        if (g->dtype == TQ_F32)
            return launch_sweep_fwd<float, FAM>(g, B, a, u, domains, workspace, workspace_bytes, out, out64,
                                                as_stream(stream));
        return launch_sweep_fwd<double, FAM>(g, B, a, u, domains, workspace, workspace_bytes, out, out64,
                                             as_stream(stream));
    });
    return TQ_OK;
}

int tq_sweep_backward(const tq_grid_desc* g, int32_t family, int64_t B, const void* a, const void* u,
                      const void* domains, const void* grad_out, void* workspace, size_t workspace_bytes,
                      void* grad_a, void* grad_u, void* grad_domains, void* stream) {
    int rc = check_sweep(g, B);
    if (!rc) rc = check_family(family);
    if (rc) return rc;
    if (!a || !u || !domains || !grad_out) return fail(TQ_E_NULL, "a/u/domains/grad_out is NULL");
    TQ_SWEEP_FAMILY(family, {
        if (g->dtype == TQ_F32)
            return launch_sweep_bwd<float, FAM>(g, B, a, u, domains, grad_out, workspace, workspace_bytes, grad_a,
                                                grad_u, grad_domains, as_stream(stream));
        return launch_sweep_bwd<double, FAM>(g, B, a, u, domains, grad_out, workspace, workspace_bytes, grad_a,
                                             grad_u, grad_domains, as_stream(stream));
    });
    return TQ_OK;
}

int tq_sweep_gaussian_forward(const tq_grid_desc* g, int64_t B, const void* a, const void* u, const void* domains,
                              void* workspace, size_t workspace_bytes, void* out, double* out64, void* stream) {
    return tq_sweep_forward(g, TQ_FAM_GAUSSIAN_PEAK, B, a, u, domains, workspace, workspace_bytes, out, out64,
                            stream);
}

int tq_sweep_gaussian_backward(const tq_grid_desc* g, int64_t B, const void* a, const void* u, const void* domains,
                               const void* grad_out, void* workspace, size_t workspace_bytes, void* grad_a,
                               void* grad_u, void* grad_domains, void* stream) {
    return tq_sweep_backward(g, TQ_FAM_GAUSSIAN_PEAK, B, a, u, domains, grad_out, workspace, workspace_bytes, grad_a,
                             grad_u, grad_domains, stream);
}

}  // extern "C"
