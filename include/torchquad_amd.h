/*
 * torchquad_amd — C ABI of the MI355X (gfx950) integration hot path.
 *
 * Every entry point takes device pointers that the caller owns and
 * allocated, plus a hipStream_t (passed as void* so that this header needs
 * no HIP include).  Nothing here allocates device memory except the opaque
 * VEGAS state handles, which say so.  Scratch memory is passed in as a
 * workspace whose size the matching *_workspace_size() call reports.
 *
 * Status codes: 0 = ok, < 0 = invalid argument (TQ_E_*), > 0 = HIP error
 * code.  tq_last_error() returns a thread-local message for the last
 * failing call on the calling thread.
 *
 * The reference (esa/torchquad 0.5.0) is pure Python with no native
 * boundary; each entry point below cites the reference function whose
 * device work it replaces.
 */
#ifndef TORCHQUAD_AMD_H
#define TORCHQUAD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TQ_MAX_DIM 16

/* status codes */
#define TQ_OK 0
#define TQ_E_ARG (-1)        /* invalid argument value                 */
#define TQ_E_NULL (-2)       /* required pointer is NULL               */
#define TQ_E_WORKSPACE (-3)  /* workspace smaller than required        */
#define TQ_E_UNSUPPORTED (-4)/* combination not implemented            */

/* element types */
#define TQ_F32 0
#define TQ_F64 1

/* quadrature rules (reference: trapezoid.py, simpson.py, boole.py, gaussian.py) */
#define TQ_RULE_TRAPEZOID 0
#define TQ_RULE_SIMPSON 1
#define TQ_RULE_BOOLE 2
#define TQ_RULE_GAUSS_LEGENDRE 3

/* built-in integrand families evaluated inside the fused kernels */
#define TQ_FAM_GAUSSIAN_PEAK 0   /* c0*exp(-sum a_k^2 (x_k-u_k)^2)                  */
#define TQ_FAM_OSCILLATORY 1     /* c0*cos(2*pi*c1 + sum a_k x_k)                    */
#define TQ_FAM_PRODUCT_PEAK 2    /* c0*prod 1/(a_k^-2 + (x_k-u_k)^2)                 */
#define TQ_FAM_EXP_LINEAR 3      /* c0*exp(c1 + sum a_k x_k)                         */
#define TQ_FAM_SIN_SUM 4         /* c0*sum sin(a_k x_k + u_k)                        */
#define TQ_FAM_DOUBLE_GAUSSIAN 5 /* c0*sum_{c in {c1,c2}} exp(-c3*sum (x_k-c)^2)     */
#define TQ_FAM_CONSTANT 6        /* c0                                               */
#define TQ_NUM_FAMILIES 7

/* Integrand descriptor: family id, dimension and parameters (host values). */
typedef struct tq_integrand {
    int32_t family;
    int32_t dim;
    double c0, c1, c2, c3;
    double a[TQ_MAX_DIM];
    double u[TQ_MAX_DIM];
} tq_integrand;

/* Tensor-product grid description.  n points per axis on every axis (the
 * reference's grids have the same n on each axis, integration_grid.py:64). */
typedef struct tq_grid_desc {
    int32_t dtype;          /* TQ_F32 / TQ_F64: arithmetic type of points and f  */
    int32_t rule;           /* TQ_RULE_*                                          */
    int32_t dim;            /* 1..TQ_MAX_DIM (fused) ; any >=1 for points/reduce  */
    int32_t n;              /* points per axis                                    */
    int32_t lane_axes;      /* tuning: axes enumerated by lanes (0 = automatic)   */
    int32_t items_per_wave; /* tuning: loop items per wave (0 = automatic)        */
    const double* gl_nodes;   /* device [n] Gauss-Legendre nodes on [-1,1]      */
    const double* gl_weights; /* device [n] Gauss-Legendre weights               */
} tq_grid_desc;

const char* tq_last_error(void);
int tq_version(void);
int tq_device_count(void);

/* ---- grid rules ---------------------------------------------------------
 * Fused tensor-product rule: nodes, per-axis weights, the built-in
 * integrand and the weighted reduction in one pass; points never touch HBM.
 * Replaces IntegrationGrid.__init__ (integration_grid.py:28-103),
 * BaseIntegrator.evaluate_integrand (base_integrator.py:43-91) and the
 * _apply_composite_rule collapse (trapezoid.py:28-37, simpson.py:30-45,
 * boole.py:30-48, gaussian.py:132-144) for built-in integrands.
 * domain: device [dim][2] of dtype.  out: device scalar of dtype.
 * out_f64 (optional, may be NULL): device scalar receiving the FP64 value. */
size_t tq_grid_workspace_size(const tq_grid_desc* g);
int tq_grid_integrate(const tq_grid_desc* g, const tq_integrand* f, const void* domain,
                      void* workspace, size_t workspace_bytes, void* out, double* out_f64,
                      void* stream);

/* Point writer for user integrands: writes the [n^dim, dim] grid in ij
 * order (last axis fastest), as IntegrationGrid.points
 * (integration_grid.py:98-99). */
int tq_grid_points(const tq_grid_desc* g, const void* domain, void* points, void* stream);

/* h_d = x_1 - x_0 per axis (integration_grid.py:90-93), dtype, [dim]. */
int tq_grid_h(const tq_grid_desc* g, const void* domain, void* h_out, void* stream);

/* Weighted reduction of user integrand values f[n^dim, S] (contiguous) to
 * out[S] with the rule's tensor-product weights: the composite collapse of
 * GridIntegrator.calculate_result (grid_integrator.py:57-91).
 * hs (optional, device [dim] of dtype): Newton-Cotes mesh widths to scale
 * with, as the reference's _apply_composite_rule does; NULL derives them
 * from the domain.  unit_weights != 0 sums with weight 1 per point (Gaussian
 * rules whose values were already multiplied by the weights,
 * base_integrator.py:77-89), scaling by (b-a)/2 per axis. */
size_t tq_grid_reduce_workspace_size(const tq_grid_desc* g, int64_t S);
int tq_grid_reduce(const tq_grid_desc* g, const void* domain, const void* hs, int32_t unit_weights,
                   const void* fvals, int64_t S, void* workspace, size_t workspace_bytes, void* out,
                   void* stream);

/* Backward of tq_grid_reduce: grad_f[p, s] = grad_out[s] * scale * W(p),
 * the tensor-product weight of point p (same hs/unit_weights meaning). */
int tq_grid_reduce_backward(const tq_grid_desc* g, const void* domain, const void* hs,
                            int32_t unit_weights, const void* grad_out, int64_t S, void* grad_f,
                            void* stream);

/* Backward of tq_grid_points: grad_domain[k, 0/1] = sum_p grad_points[p, k]
 * * d x_k(p) / d(a_k, b_k) (the derivative of linspace / of the
 * Gauss-Legendre affine map, utils.py:24-58, gaussian.py:159-162). */
size_t tq_grid_points_backward_workspace_size(const tq_grid_desc* g);
int tq_grid_points_backward(const tq_grid_desc* g, const void* domain, const void* grad_points,
                            void* workspace, size_t workspace_bytes, void* grad_domain, void* stream);

/* Gauss-Legendre nodes (ascending) and weights on [-1,1] for any n >= 1,
 * computed on the device (Newton on the Legendre recurrence from Tricomi's
 * estimates).  Replaces numpy.polynomial.legendre.leggauss, the reference's
 * root function (gaussian.py:25, :110-129), whose dense n x n eigenproblem
 * does not scale to the n ~ 1e6 axes of large one-dimensional rules. */
int tq_gauss_legendre(int32_t n, double* nodes, double* weights, void* stream);

/* ---- Monte Carlo --------------------------------------------------------
 * Counter-based sampling (Philox-4x32-10, keyed by seed; counter = sample
 * index, block, stream offset) fused with the integrand and a block
 * reduction of (sum f, sum f^2).  Replaces MonteCarlo.calculate_sample_points
 * (monte_carlo.py:84-106), RNG.uniform (rng.py:121-125) and
 * MonteCarlo.calculate_result (monte_carlo.py:60-82). */
typedef struct tq_mc_desc {
    int32_t dtype;
    int32_t dim;
    int64_t N;               /* number of samples                               */
    uint64_t seed;           /* Philox key                                      */
    uint64_t offset;         /* first Philox counter (stream position)          */
    int32_t samples_per_thread; /* tuning (0 = automatic)                       */
    int32_t reserved;
} tq_mc_desc;

size_t tq_mc_workspace_size(const tq_mc_desc* m);
/* out: device scalar (dtype) = volume * mean f.  stats (optional): device
 * double[3] = {sum f, sum f^2, N} for the standard error. */
int tq_mc_integrate(const tq_mc_desc* m, const tq_integrand* f, const void* domain,
                    void* workspace, size_t workspace_bytes, void* out, double* stats,
                    void* stream);
/* Sample writer for user integrands: points[N, dim] = a + U*(b-a), the same
 * stream that tq_mc_integrate draws. */
int tq_mc_points(const tq_mc_desc* m, const void* domain, void* points, void* stream);
/* volume * sum_i f[i, s] / N for user integrand values f[N, S]. */
size_t tq_mc_reduce_workspace_size(int64_t N, int64_t S);
int tq_mc_reduce(int32_t dtype, const void* fvals, int64_t N, int64_t S, const void* domain,
                 int32_t dim, void* workspace, size_t workspace_bytes, void* out, void* stream);

/* Uniform [0,1) numbers from the same Philox stream (RNG.uniform,
 * rng.py:121-125): out[r, c] (row-major [rows, cols]) is coordinate c of
 * sample offset+r, i.e. exactly the numbers tq_mc_points scales. */
int tq_uniform(uint64_t seed, uint64_t offset, int32_t dtype, int64_t rows, int32_t cols,
               void* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TORCHQUAD_AMD_H */
