/*
 * ORACLE — test infrastructure only (never linked into the product library).
 *
 * Plain C FP64 checkers for the integration hot path:
 *   oracle_philox4x32_10  Philox-4x32-10 from Salmon et al., SC'11 (pinned by
 *                         the Random123 known-answer vectors in tests/golden).
 *   oracle_uniform_*      the library's documented sample stream
 *                         (include/torchquad_amd.h, tq_uniform).
 *   oracle_grid_sum       brute-force tensor-product rule: every point decoded
 *                         from its flat index, full per-axis weights
 *                         (optionally also sum |w f|, the scale of the
 *                         rounding error of any summation order)
 *                         (trapezoid.py:28-37, simpson.py:30-45, boole.py:30-48,
 *                         gaussian.py:44-69,132-144 written as weight vectors),
 *                         full integrand evaluation per point, FP64 sum.
 *   oracle_mc_sum         Monte Carlo sums on the same stream (monte_carlo.py
 *                         :60-106) with points formed exactly like the fp32 /
 *                         fp64 kernels and the integrand in FP64.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    int round;
    for (round = 0; round < 10; ++round) {
        uint64_t prod0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t prod1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(prod0 >> 32), lo0 = (uint32_t)(prod0 & 0xffffffffu);
        uint32_t hi1 = (uint32_t)(prod1 >> 32), lo1 = (uint32_t)(prod1 & 0xffffffffu);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 = k0 + 0x9E3779B9u;
        k1 = k1 + 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static void stream_block(uint64_t seed, uint64_t sample, uint32_t block, uint32_t out[4]) {
    uint32_t ctr[4] = {(uint32_t)(sample & 0xffffffffu), (uint32_t)(sample >> 32), block, 0u};
    uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    oracle_philox4x32_10(ctr, key, out);
}

static float word_to_f32(uint32_t r) { return (float)(r >> 8) * (1.0f / 16777216.0f); }
static double words_to_f64(uint32_t a, uint32_t b) {
    uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return (double)m * (1.0 / 9007199254740992.0);
}

void oracle_uniform_f32(uint64_t seed, uint64_t offset, int64_t rows, int32_t cols, float* out) {
    int64_t r;
    int32_t c;
    for (r = 0; r < rows; ++r) {
        for (c = 0; c < cols; ++c) {
            uint32_t w[4];
            stream_block(seed, offset + (uint64_t)r, (uint32_t)(c / 4), w);
            out[r * cols + c] = word_to_f32(w[c % 4]);
        }
    }
}

void oracle_uniform_f64(uint64_t seed, uint64_t offset, int64_t rows, int32_t cols, double* out) {
    int64_t r;
    int32_t c;
    for (r = 0; r < rows; ++r) {
        for (c = 0; c < cols; ++c) {
            uint32_t w[4];
            stream_block(seed, offset + (uint64_t)r, (uint32_t)(c / 2), w);
            out[r * cols + c] = (c % 2) ? words_to_f64(w[2], w[3]) : words_to_f64(w[0], w[1]);
        }
    }
}

/* ---- integrand families, FP64 ------------------------------------------ */
enum { GAUSSIAN_PEAK = 0, OSCILLATORY = 1, PRODUCT_PEAK = 2, EXP_LINEAR = 3, SIN_SUM = 4,
       DOUBLE_GAUSSIAN = 5, CONSTANT = 6 };

typedef struct {
    int family, dim;
    double c0, c1, c2, c3;
    const double* a;
    const double* u;
} fam_t;

static double fam_eval(const fam_t* F, const double* x) {
    int k;
    double s = 0.0, s2 = 0.0, p = 1.0;
    switch (F->family) {
        case GAUSSIAN_PEAK:
            for (k = 0; k < F->dim; ++k) {
                double t = F->a[k] * (x[k] - F->u[k]);
                s += t * t;
            }
            return F->c0 * exp(-s);
        case OSCILLATORY:
            s = 2.0 * M_PI * F->c1;
            for (k = 0; k < F->dim; ++k) s += F->a[k] * x[k];
            return F->c0 * cos(s);
        case PRODUCT_PEAK:
            for (k = 0; k < F->dim; ++k) {
                double t = x[k] - F->u[k];
                p *= 1.0 / (1.0 / (F->a[k] * F->a[k]) + t * t);
            }
            return F->c0 * p;
        case EXP_LINEAR:
            s = F->c1;
            for (k = 0; k < F->dim; ++k) s += F->a[k] * x[k];
            return F->c0 * exp(s);
        case SIN_SUM:
            for (k = 0; k < F->dim; ++k) s += sin(F->a[k] * x[k] + F->u[k]);
            return F->c0 * s;
        case DOUBLE_GAUSSIAN:
            for (k = 0; k < F->dim; ++k) {
                double t1 = x[k] - F->c1, t2 = x[k] - F->c2;
                s += t1 * t1;
                s2 += t2 * t2;
            }
            return F->c0 * (exp(-F->c3 * s) + exp(-F->c3 * s2));
        default:
            return F->c0;
    }
}

/* ---- tensor-product rule, brute force ---------------------------------- */
enum { TRAPEZOID = 0, SIMPSON = 1, BOOLE = 2, GAUSS_LEGENDRE = 3 };

static void axis_rule(int rule, double a, double b, int n, const double* glx, const double* glw,
                      double* x, double* w) {
    int j;
    if (rule == GAUSS_LEGENDRE) {
        for (j = 0; j < n; ++j) {
            x[j] = (b - a) / 2.0 * glx[j] + (a + b) / 2.0;
            w[j] = glw[j] * (b - a) / 2.0;
        }
        return;
    }
    double h = (b - a) / (double)(n - 1);
    for (j = 0; j < n; ++j) {
        x[j] = (j == n - 1) ? b : a + (double)j * h;
        int end = (j == 0 || j == n - 1);
        double c;
        if (rule == TRAPEZOID) c = end ? 0.5 : 1.0;
        else if (rule == SIMPSON) c = end ? 1.0 / 3.0 : ((j % 2) ? 4.0 / 3.0 : 2.0 / 3.0);
        else c = end ? 7.0 / 22.5 : ((j % 2) ? 32.0 / 22.5 : ((j % 4 == 2) ? 12.0 / 22.5 : 14.0 / 22.5));
        w[j] = c * h;
    }
}

double oracle_grid_sum(int rule, int family, int dim, int n, const double* domain, double c0, double c1,
                       double c2, double c3, const double* a, const double* u, const double* glx,
                       const double* glw, double* abs_sum) {
    fam_t F = {family, dim, c0, c1, c2, c3, a, u};
    double* X = (double*)malloc(sizeof(double) * dim * n);
    double* W = (double*)malloc(sizeof(double) * dim * n);
    int* dig = (int*)calloc(dim, sizeof(int));
    double* x = (double*)malloc(sizeof(double) * dim);
    int k;
    for (k = 0; k < dim; ++k) axis_rule(rule, domain[2 * k], domain[2 * k + 1], n, glx, glw, X + k * n, W + k * n);
    double total = 0.0, comp = 0.0, atotal = 0.0;
    for (;;) {
        double w = 1.0;
        for (k = 0; k < dim; ++k) {
            x[k] = X[k * n + dig[k]];
            w *= W[k * n + dig[k]];
        }
        /* Kahan-compensated FP64 sum */
        double wf = w * fam_eval(&F, x);
        double term = wf - comp;
        double t = total + term;
        atotal += fabs(wf);
        comp = (t - total) - term;
        total = t;
        for (k = dim - 1; k >= 0; --k) {
            if (++dig[k] < n) break;
            dig[k] = 0;
        }
        if (k < 0) break;
    }
    free(X);
    free(W);
    free(dig);
    free(x);
    if (abs_sum) *abs_sum = atotal;
    return total;
}

/* ---- Monte Carlo on the library's stream -------------------------------- */
void oracle_mc_sum(int dtype, int family, int dim, int64_t N, uint64_t seed, uint64_t offset,
                   const double* domain, double c0, double c1, double c2, double c3, const double* a,
                   const double* u, double* sum, double* sum2, double* sum_abs) {
    fam_t F = {family, dim, 1.0, c1, c2, c3, a, u};
    double x[64];
    double s1 = 0.0, s2 = 0.0, sa = 0.0;
    int64_t i;
    int k;
    (void)c0;
    for (i = 0; i < N; ++i) {
        if (dtype == 0) {
            float uu[64];
            oracle_uniform_f32(seed, offset + (uint64_t)i, 1, dim, uu);
            for (k = 0; k < dim; ++k) {
                float lo = (float)domain[2 * k], hi = (float)domain[2 * k + 1];
                x[k] = (double)fmaf(uu[k], hi - lo, lo);
            }
        } else {
            double uu[64];
            oracle_uniform_f64(seed, offset + (uint64_t)i, 1, dim, uu);
            for (k = 0; k < dim; ++k) {
                double lo = domain[2 * k], hi = domain[2 * k + 1];
                x[k] = fma(uu[k], hi - lo, lo);
            }
        }
        double f = fam_eval(&F, x);
        s1 += f;
        s2 += f * f;
        sa += fabs(f);
    }
    *sum = s1;
    *sum2 = s2;
    if (sum_abs) *sum_abs = sa;
}
