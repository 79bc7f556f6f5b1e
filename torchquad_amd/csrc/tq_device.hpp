// Device-side building blocks shared by every kernel of the library:
// deterministic wave/block reductions, the Philox-4x32-10 counter-based
// generator, and the built-in integrand families written as axis folds.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/torchquad_amd.h"

namespace tq {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

// ---------------------------------------------------------------------------
// Reductions.  Butterfly over the 64-lane wavefront: every lane ends with the
// same value and the summation order is fixed, so results are bit-identical
// from run to run.
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block sum in wave order; the value is valid in thread 0 only.
__device__ inline double block_sum(double v, double* smem /* [kWavesPerBlock] */) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    if (lane == 0) smem[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += smem[w];
    }
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Ten rounds of two 32x32->64 multiplies with the
// round key bumped by the Weyl constants between rounds.
// ---------------------------------------------------------------------------
struct u32x4 {
    uint32_t v[4];
};

__device__ __host__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.v[0];
        const uint64_t p1 = (uint64_t)M1 * c.v[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        u32x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// The library's sample stream: sample s, block b of four 32-bit words uses
// counter (lo32(s), hi32(s), b, 0) under key (lo32(seed), hi32(seed)).
__device__ inline u32x4 philox_block(uint64_t seed, uint64_t s, uint32_t b) {
    u32x4 c;
    c.v[0] = (uint32_t)s;
    c.v[1] = (uint32_t)(s >> 32);
    c.v[2] = b;
    c.v[3] = 0u;
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// 32-bit word -> float in [0,1) with 24 random bits.
__device__ inline float u01f(uint32_t r) { return (float)(r >> 8) * 0x1.0p-24f; }
// two words -> double in [0,1) with 53 random bits.
__device__ inline double u01d(uint32_t a, uint32_t b) {
    return (double)(((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6)) * 0x1.0p-53;
}

// Fills x[0..d) with the sample-s coordinates in [0,1): fp32 uses word k for
// coordinate k, fp64 words 2k and 2k+1.
template <typename T>
struct Uniform;

template <>
struct Uniform<float> {
    template <int MAXD>
    __device__ static inline void draw(uint64_t seed, uint64_t s, int d, float* u) {
#pragma unroll
        for (int b = 0; b < (MAXD + 3) / 4; ++b) {
            if (4 * b < d) {
                const u32x4 r = philox_block(seed, s, (uint32_t)b);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (4 * b + i < MAXD) u[4 * b + i] = u01f(r.v[i]);
            }
        }
    }
};

template <>
struct Uniform<double> {
    template <int MAXD>
    __device__ static inline void draw(uint64_t seed, uint64_t s, int d, double* u) {
#pragma unroll
        for (int b = 0; b < (MAXD + 1) / 2; ++b) {
            if (2 * b < d) {
                const u32x4 r = philox_block(seed, s, (uint32_t)b);
                u[2 * b] = u01d(r.v[0], r.v[1]);
                if (2 * b + 1 < MAXD) u[2 * b + 1] = u01d(r.v[2], r.v[3]);
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Math wrappers (device math library calls only).
// ---------------------------------------------------------------------------
__device__ inline float tq_exp(float x) { return __expf(x); }
__device__ inline double tq_exp(double x) { return exp(x); }
// 2^x: one v_exp_f32 in FP32 (callers fold log2(e) into their constants)
__device__ inline float tq_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ inline double tq_exp2(double x) { return exp2(x); }
__device__ inline float tq_cos(float x) { return cosf(x); }
__device__ inline double tq_cos(double x) { return cos(x); }
__device__ inline float tq_sin(float x) { return sinf(x); }
__device__ inline double tq_sin(double x) { return sin(x); }
__device__ inline float tq_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// cos(2 pi s) for s in revolutions (FP32).  r = s - rint(s) lies in
// [-1/2, 1/2] (exact for |s| < 2^23); q = |r| is reflected into [0, 1/4]
// with cos(2 pi q) = -cos(2 pi (1/2 - q)); then an even polynomial in q^2
// whose coefficients were fitted by least squares on 400 Chebyshev nodes of
// q^2 in [0, 1/16] against mpmath (max abs error 2.2e-10 before FP32
// rounding, 1.6e-7 after).
__device__ inline float tq_cos_rev(float s) {
    const float r = s - rintf(s);
    float q = fabsf(r);
    const bool flip = q > 0.25f;
    q = flip ? 0.5f - q : q;
    const float t = q * q;
    float p = -24.981836135953497f;
    p = fmaf(p, t, 60.14395233715792f);
    p = fmaf(p, t, -85.4535681952404f);
    p = fmaf(p, t, 64.93934655290128f);
    p = fmaf(p, t, -19.739208548315304f);
    p = fmaf(p, t, 0.9999999997799042f);
    return flip ? -p : p;
}
// oscillatory finish: FP32 state in revolutions, FP64 state in radians
__device__ inline float tq_osc_cos(float s) { return tq_cos_rev(s); }
__device__ inline double tq_osc_cos(double s) { return cos(s); }
__device__ inline double tq_rcp(double x) { return 1.0 / x; }

// ---------------------------------------------------------------------------
// Integrand families as associative folds over axes.
//
// f(x) = c0 * finish(fold(...fold(start, 0, x_0)..., d-1, x_{d-1})).
// `start` carries the family's constant term and is used exactly once per
// point; `zero` is the identity of `merge`, which combines two partial folds
// over disjoint axis sets.  Kernels fold the axes that stay fixed over many
// points once and fold the varying axis per point; every point still gets
// its own finish() (the transcendental) and its own weighted accumulate.
// ---------------------------------------------------------------------------
template <typename T>
struct Params {
    T c0, c1, c2, c3;
    T a[TQ_MAX_DIM];
    T u[TQ_MAX_DIM];
    T b[TQ_MAX_DIM];  // family-specific per-axis derived constant
};

template <typename T>
struct State {
    T s0, s1;
};

// Per-axis family constants, loaded once per axis outside the hot loops.
template <typename T>
struct AxisC {
    T a, u, b;
};

template <typename T>
__device__ inline AxisC<T> axis_c(const Params<T>& p, int k) {
    return AxisC<T>{p.a[k], p.u[k], p.b[k]};
}

template <int FAM, typename T>
struct Family;

// exp(-sum a_k^2 (x_k-u_k)^2) = 2^(-sum t_k^2) with t_k = fma(x, a', b'),
// a' = a sqrt(log2 e), b' = -a u sqrt(log2 e) (Params prepared on the host).
template <typename T>
struct Family<TQ_FAM_GAUSSIAN_PEAK, T> {
    __device__ static inline State<T> start(const Params<T>&) { return {T(0), T(0)}; }
    __device__ static inline State<T> zero() { return {T(0), T(0)}; }
    __device__ static inline State<T> fold_c(State<T> s, const AxisC<T>& c, T x, const Params<T>&) {
        const T t = fma(x, c.a, c.b);
        s.s0 = fma(-t, t, s.s0);
        return s;
    }
    __device__ static inline State<T> merge(State<T> a, State<T> b) { return {a.s0 + b.s0, T(0)}; }
    __device__ static inline T finish(State<T> s, const Params<T>&) { return tq_exp2(s.s0); }
    // the weight inside the state: w 2^s = 2^(s + log2 w); the key of a
    // product of weights is the sum of their keys
    __device__ static inline double weight_key(double w) { return log2(w); }
    __device__ static inline double combine_keys(double a, double b) { return a + b; }
    __device__ static inline T fold_key(T s0, double key) { return (T)((double)s0 + key); }
};

// cos(2*pi*c1 + sum a_k x_k);  start carries the phase 2*pi*c1 (in c2).  In
// FP32 the state is in revolutions (host divides a_k by 2*pi, c2 = c1).
template <typename T>
struct Family<TQ_FAM_OSCILLATORY, T> {
    __device__ static inline State<T> start(const Params<T>& p) { return {p.c2, T(0)}; }
    __device__ static inline State<T> zero() { return {T(0), T(0)}; }
    __device__ static inline State<T> fold_c(State<T> s, const AxisC<T>& c, T x, const Params<T>&) {
        s.s0 = fma(c.a, x, s.s0);
        return s;
    }
    __device__ static inline State<T> merge(State<T> a, State<T> b) { return {a.s0 + b.s0, T(0)}; }
    __device__ static inline T finish(State<T> s, const Params<T>&) { return tq_osc_cos(s.s0); }
};

// prod 1/(a_k^-2 + (x_k-u_k)^2);  b_k = a_k^-2; the fold keeps the product of
// the denominators and finish takes one reciprocal.
template <typename T>
struct Family<TQ_FAM_PRODUCT_PEAK, T> {
    __device__ static inline State<T> start(const Params<T>&) { return {T(1), T(0)}; }
    __device__ static inline State<T> zero() { return {T(1), T(0)}; }
    __device__ static inline State<T> fold_c(State<T> s, const AxisC<T>& c, T x, const Params<T>&) {
        const T t = x - c.u;
        s.s0 *= fma(t, t, c.b);
        return s;
    }
    __device__ static inline State<T> merge(State<T> a, State<T> b) { return {a.s0 * b.s0, T(0)}; }
    __device__ static inline T finish(State<T> s, const Params<T>&) { return tq_rcp(s.s0); }
    // the weight inside the state: w / s = 1 / (s (1 / w))
    __device__ static inline double weight_key(double w) { return 1.0 / w; }
    __device__ static inline double combine_keys(double a, double b) { return a * b; }
    __device__ static inline T fold_key(T s0, double key) { return (T)((double)s0 * key); }
};

// exp(c1 + sum a_k x_k) = 2^(c1' + sum a'_k x_k), a' = a log2 e, c1' = c1 log2 e
template <typename T>
struct Family<TQ_FAM_EXP_LINEAR, T> {
    __device__ static inline State<T> start(const Params<T>& p) { return {p.c1, T(0)}; }
    __device__ static inline State<T> zero() { return {T(0), T(0)}; }
    __device__ static inline State<T> fold_c(State<T> s, const AxisC<T>& c, T x, const Params<T>&) {
        s.s0 = fma(c.a, x, s.s0);
        return s;
    }
    __device__ static inline State<T> merge(State<T> a, State<T> b) { return {a.s0 + b.s0, T(0)}; }
    __device__ static inline T finish(State<T> s, const Params<T>&) { return tq_exp2(s.s0); }
    __device__ static inline double weight_key(double w) { return log2(w); }
    __device__ static inline double combine_keys(double a, double b) { return a + b; }
    __device__ static inline T fold_key(T s0, double key) { return (T)((double)s0 + key); }
};

// sum sin(a_k x_k + u_k)
template <typename T>
struct Family<TQ_FAM_SIN_SUM, T> {
    __device__ static inline State<T> start(const Params<T>&) { return {T(0), T(0)}; }
    __device__ static inline State<T> zero() { return {T(0), T(0)}; }
    __device__ static inline State<T> fold_c(State<T> s, const AxisC<T>& c, T x, const Params<T>&) {
        s.s0 += tq_sin(fma(c.a, x, c.u));
        return s;
    }
    __device__ static inline State<T> merge(State<T> a, State<T> b) { return {a.s0 + b.s0, T(0)}; }
    __device__ static inline T finish(State<T> s, const Params<T>&) { return s.s0; }
};

// exp(-c3 sum (x_k-c1)^2) + exp(-c3 sum (x_k-c2)^2); Params c3 holds c3 log2 e
template <typename T>
struct Family<TQ_FAM_DOUBLE_GAUSSIAN, T> {
    __device__ static inline State<T> start(const Params<T>&) { return {T(0), T(0)}; }
    __device__ static inline State<T> zero() { return {T(0), T(0)}; }
    __device__ static inline State<T> fold_c(State<T> s, const AxisC<T>&, T x, const Params<T>& p) {
        const T t1 = x - p.c1, t2 = x - p.c2;
        s.s0 = fma(t1, t1, s.s0);
        s.s1 = fma(t2, t2, s.s1);
        return s;
    }
    __device__ static inline State<T> merge(State<T> a, State<T> b) {
        return {a.s0 + b.s0, a.s1 + b.s1};
    }
    __device__ static inline T finish(State<T> s, const Params<T>& p) {
        return tq_exp2(-p.c3 * s.s0) + tq_exp2(-p.c3 * s.s1);
    }
};

template <typename T>
struct Family<TQ_FAM_CONSTANT, T> {
    __device__ static inline State<T> start(const Params<T>&) { return {T(0), T(0)}; }
    __device__ static inline State<T> zero() { return {T(0), T(0)}; }
    __device__ static inline State<T> fold_c(State<T> s, const AxisC<T>&, T, const Params<T>&) { return s; }
    __device__ static inline State<T> merge(State<T> a, State<T>) { return a; }
    __device__ static inline T finish(State<T>, const Params<T>&) { return T(1); }
};

template <int FAM, typename T>
__device__ inline State<T> fold(State<T> s, int k, T x, const Params<T>& p) {
    return Family<FAM, T>::fold_c(s, axis_c(p, k), x, p);
}

// Full evaluation at one point (VEGAS, Monte Carlo): fold every axis.
template <int FAM, typename T, int MAXD>
__device__ inline T eval_point(const T* x, int d, const Params<T>& p) {
    using F = Family<FAM, T>;
    State<T> s = F::start(p);
#pragma unroll
    for (int k = 0; k < MAXD; ++k)
        if (k < d) s = F::fold_c(s, axis_c(p, k), x[k], p);
    return F::finish(s, p);
}

}  // namespace tq
