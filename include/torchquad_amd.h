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
#define TQ_I32 2 /* tq_add_at_indices only */
#define TQ_I64 3 /* tq_add_at_indices only */

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
 * out_f64 (optional, may be NULL): device scalar receiving the FP64 value.
 * workspace: tq_grid_workspace_size() bytes; its first 4352 bytes hold
 * completion counters (the last workgroup to finish sums the partials, so one
 * launch does the whole integral) that must be zero before the first call
 * with that workspace -- clear them once after allocating -- and that every
 * call leaves zero. */
size_t tq_grid_workspace_size(const tq_grid_desc* g);
int tq_grid_integrate(const tq_grid_desc* g, const tq_integrand* f, const void* domain,
                      void* workspace, size_t workspace_bytes, void* out, double* out_f64,
                      void* stream);

/* tq_grid_integrate without a workspace, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The
 * completion counter and block partials live in a static device array of
 * 1 MiB per device (zero from load; every call leaves the counter zero), each device's
 * address looked up once on that device's first call; the call allocates
 * nothing.  Grids with tq_grid_workspace_size() > 1 MiB (more than 130528
 * workgroups) return TQ_E_WORKSPACE.
 * Results equal tq_grid_integrate's with a workspace of
 * tq_grid_workspace_size() bytes.  Every call on a device shares that array:
 * calls that overlap on one device (different streams, or host threads)
 * race and must be serialised by the caller. */
int tq_grid_integrate_noalloc(const tq_grid_desc* g, const tq_integrand* f, const void* domain, void* out,
                              double* out_f64, void* stream);

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
/* tq_mc_integrate with the first Philox counter read on the device from
 * *offset_dev (m->offset is ignored), so that a captured HIP graph of the
 * call draws the samples of whatever stream position the caller stores there
 * before each replay: MonteCarlo.get_jit_compiled_integrate's compiled
 * function draws new points on every call (monte_carlo.py:108-225). */
int tq_mc_integrate_dev_offset(const tq_mc_desc* m, const tq_integrand* f, const void* domain,
                               const uint64_t* offset_dev, void* workspace, size_t workspace_bytes, void* out,
                               double* stats, void* stream);
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

/* ---- batched form ------------------------------------------------------
 * Many independent tensor-product integrals (mixed rule, family, dim, n,
 * domain, parameters) in one launch per LDS class.  The reference has no batched
 * integrator; its nearest relatives are vectorised integrands
 * (base_integrator.py:77-89) and the tutorial's Batch1DIntegrator
 * (docs tutorial.rst:800-888).  Items carry the per-axis n already rounded
 * by the rule (simpson.py:54-81, boole.py:56-84, integration_grid.py:64). */
typedef struct tq_batch_item {
    int32_t rule;
    int32_t family;
    int32_t dim;
    int32_t n;               /* points per axis                                   */
    int64_t gl_offset;       /* first node of this n-point rule in the GL tables  */
    double domain[2 * TQ_MAX_DIM];
    double c0, c1, c2, c3;
    double a[TQ_MAX_DIM];
    double u[TQ_MAX_DIM];
} tq_batch_item;

typedef struct tq_batch_tile {
    int32_t item;            /* integral index                                    */
    int32_t lane_wave;       /* first 64-lane group of the integral's lane axes   */
    int32_t lane_wave_end;   /* one past the last 64-lane group of the tile       */
    int32_t reserved;
    int64_t t0, t1;          /* loop items [t0, t1) of each of those groups       */
} tq_batch_tile;

/* Host-side plan sizes: bytes of the device descriptor array and tile count
 * for tiles of about tile_points points. */
int tq_batch_plan_sizes(const tq_batch_item* items, int64_t count, int32_t dtype, int64_t tile_points,
                        int64_t* desc_bytes, int64_t* n_tiles);
/* Host-side plan: descriptors (desc_out, host memory of desc_bytes), tiles
 * ordered by kernel class and, within a class, longest integral first, each
 * integral's tiles contiguous; tile_begin[i] / tile_count[i] for integral i.
 * class_offsets (optional, [5]): tiles [class_offsets[c], class_offsets[c+1])
 * belong to class c -- 0: one-axis integrals; multi-axis ones with 1: loop
 * and row-axis tables within 4 KB of LDS per wave (or no loop axis), 2:
 * within 8 KB, 3: larger (general kernel). */
int tq_batch_plan(const tq_batch_item* items, int64_t count, int32_t dtype, int64_t tile_points, void* desc_out,
                  tq_batch_tile* tiles_out, int64_t* tile_begin, int64_t* tile_count, int64_t* class_offsets);
/* Runs a plan whose arrays were copied to the device: one kernel per
 * non-empty class, one wave per tile in the plan's order (longest integrals
 * first, so the hardware dispatcher hands out the large tiles first), then
 * out[i] (dtype) / out64[i] = integral i.  With more than one non-empty
 * class the classes run concurrently on two library-owned streams forked
 * from `stream` and joined back into it with events (under graph capture
 * the run becomes a forked subgraph).  tile_partials: [n_tiles];
 * class_offsets: host [5] from tq_batch_plan, or NULL to run every tile on
 * a general kernel that carries both the one-axis and multi-axis code.  gl_nodes / gl_weights: every Gauss-Legendre item's
 * table at its gl_offset, gl_len entries each; gl_pairs32 (optional, F32
 * batches only, else ignored): the same tables as interleaved FP32 (node,
 * weight) pairs, pair k at [2 (gl_offset + k)], read by one-axis items
 * instead of the FP64 tables.  A Gauss-Legendre item whose gl_offset + n
 * exceeds gl_len (or with NULL tables) reads no table and its result is
 * NaN.  The run needs no workspace beyond tile_partials and no clearing, so
 * it can be captured in a HIP graph and replayed. */
int tq_batch_run(int32_t dtype, const void* desc, int64_t count, const tq_batch_tile* tiles, int64_t n_tiles,
                 const int64_t* tile_begin, const int64_t* tile_count, const double* gl_nodes,
                 const double* gl_weights, const float* gl_pairs32, int64_t gl_len, double* tile_partials,
                 void* out, double* out64, const int64_t* class_offsets, void* stream);

/* ---- differentiation ----------------------------------------------------
 * Parameter sweep of the Gaussian peak f = exp(-sum a_k^2 (x_k-u_k)^2) over B
 * integrals sharing rule / dim (1..8) / n with their own a[B][dim],
 * u[B][dim] and domains[B][dim][2] (dtype), and its hand-written backward:
 * grad_a, grad_u, grad_domains (any may be NULL) = grad_out[b] * d I_b / d.
 * The reference differentiates with torch autograd through the materialised
 * grid (gaussian.py:44-162, utils.py:24-58; tests/gradient_test.py). */
size_t tq_sweep_workspace_size(const tq_grid_desc* g, int64_t B);
int tq_sweep_gaussian_forward(const tq_grid_desc* g, int64_t B, const void* a, const void* u, const void* domains,
                              void* workspace, size_t workspace_bytes, void* out, double* out64, void* stream);
int tq_sweep_gaussian_backward(const tq_grid_desc* g, int64_t B, const void* a, const void* u, const void* domains,
                               const void* grad_out, void* workspace, size_t workspace_bytes, void* grad_a,
                               void* grad_u, void* grad_domains, void* stream);
/* The same sweep for other families (family: TQ_FAM_GAUSSIAN_PEAK,
 * TQ_FAM_OSCILLATORY or TQ_FAM_PRODUCT_PEAK; others TQ_E_UNSUPPORTED), with
 * the workspace of tq_sweep_workspace_size.  Parameters as the family's
 * tq_integrand: Gaussian and product peak a[B][dim], u[B][dim] (grad_a,
 * grad_u alike); oscillatory cos(2 pi w + sum c_k x_k): c[B][dim] in a and
 * w[B] in u, so grad_a = dI/dc [B][dim] and grad_u = dI/dw [B].  The
 * reference differentiates these integrands with torch autograd through the
 * materialised grid (utils.py:24-58, tests/gradient_test.py:162-259). */
int tq_sweep_forward(const tq_grid_desc* g, int32_t family, int64_t B, const void* a, const void* u,
                     const void* domains, void* workspace, size_t workspace_bytes, void* out, double* out64,
                     void* stream);
int tq_sweep_backward(const tq_grid_desc* g, int32_t family, int64_t B, const void* a, const void* u,
                      const void* domains, const void* grad_out, void* workspace, size_t workspace_bytes,
                      void* grad_a, void* grad_u, void* grad_domains, void* stream);

/* ---- measurement --------------------------------------------------------
 * No counterpart in the reference: a measurement hook for the benchmark,
 * which states the fused grid kernel's rate against the device's own rate
 * for the kernel's inner loop.  kind 0: independent FP32 FMA chains, 32
 * instructions per lane and trip; 1: FP32 exp2 chains, likewise; 2: the
 * grid loop's body with (term, weight) table entries (state + LDS table
 * term, exp2, weighted FMA), 256 points per lane and trip; 3: the body with
 * the weight folded into 4-byte entries (state + term, exp2, add).  Launches
 * blocks x 256 threads for iters trips; out: device float[blocks * 256]. */
int tq_probe_rates(int32_t kind, int64_t blocks, int32_t iters, float* out, void* stream);

/* ---- VEGAS ----------------------------------------------------------------
 * Device-resident VEGAS+ state (reference: vegas_map.py, vegas_stratification.py,
 * vegas.py).  Every buffer is caller-owned device memory; the library only
 * reads/writes it.  Map and stratification state are FP64; samples and the
 * integrand use sample_dtype.  Capacities: jf/x/jac/ids hold max_samples
 * samples, the partial buffers hold max_blocks sample blocks of
 * TQ_VEGAS_BLOCK_SAMPLES samples each. */
#define TQ_VEGAS_BLOCK_SAMPLES 1024
#define TQ_VEGAS_MAX_ITERS 4096

typedef struct tq_vegas_state {
    int32_t dim;
    int32_t n_intervals;     /* map bins per axis (VEGASMap N_intervals)            */
    int32_t n_strat;         /* stratification steps per axis                       */
    int32_t sample_dtype;    /* TQ_F32 / TQ_F64                                     */
    int64_t n_cubes;         /* n_strat^dim                                         */
    int64_t max_samples;     /* capacity of per-sample buffers                      */
    int64_t max_blocks;      /* capacity of per-block partial buffers               */
    double v_cubes;          /* n_strat^-dim                                        */
    double alpha;            /* map smoothing exponent (vegas_map.py:15)            */
    double beta;             /* stratification damping (vegas_stratification.py:13) */
    double* x_edges;         /* [dim][n_intervals+1]                                */
    double* dx_edges;        /* [dim][n_intervals]                                  */
    double* weights;         /* [dim][n_intervals]  accumulated (f J)^2             */
    double* counts;          /* [dim][n_intervals]  accumulated sample counts       */
    double* dh;              /* [n_cubes] normalised damped counts                  */
    double* jf_cube;         /* [n_cubes] JF                                        */
    double* jf2_cube;        /* [n_cubes] JF2                                       */
    int64_t* nh;             /* [n_cubes] samples per cube                          */
    int64_t* offsets;        /* [n_cubes+1] exclusive scan of nh                    */
    int64_t* chunk_cube;     /* [max_blocks+1] cube of the first sample of a block  */
    double* jf;              /* [max_samples] per-sample f*J                        */
    double* hist_w;          /* [max(max_blocks*dim*n_intervals, 5*dim*(n_intervals+1))]
                                block partials, then update_map scratch              */
    uint32_t* hist_c;        /* [max_blocks][dim][n_intervals] block partials       */
    double* block_sums;      /* [max(max_blocks, ceil(n_cubes/256))][4] scratch     */
    int64_t* scan_sums;      /* [ceil(n_cubes/2048)+1] scratch                      */
    double* iter_stats;      /* [TQ_VEGAS_MAX_ITERS][4] = result, sigma2, evals, dh sum */
    int32_t* flags;          /* [4] = map not updated, infinite edges, unrepaired, spare */
} tq_vegas_state;

/* Resets map (uniform edges), stratification (dh = 1/n_cubes) and flags
 * (VEGASMap.__init__ vegas_map.py:15-42, VEGASStratification.__init__
 * vegas_stratification.py:13-44). */
int tq_vegas_init(const tq_vegas_state* st, void* stream);

/* get_NH (vegas_stratification.py:92-103): dh <- dh / dh_sum when
 * normalize_it >= 0 (the update_DH normalisation, :86-90, of iteration
 * normalize_it), nh = max(2, floor(dh * n_exp)), then offsets, the block
 * cube index and iter_stats[it].evals = sum nh.  nh_in != NULL uses the
 * given counts instead (step API). */
int tq_vegas_nh(const tq_vegas_state* st, int64_t n_exp, int32_t normalize_it, int32_t it,
                const int64_t* nh_in, void* stream);

/* One fused VEGAS iteration body after tq_vegas_nh (vegas.py:268-315):
 * stratified samples y (get_Y, vegas_stratification.py:140-165), map
 * (get_X/get_Jac, vegas_map.py:44-74), the built-in integrand on
 * x*(b-a)+a times the volume (vegas.py:104-112), per-sample jf, map
 * histograms (accumulate_weight, vegas_map.py:99-111), then the per-cube
 * sums, ih, sigma^2 and dh (vegas_stratification.py:46-90) into
 * iter_stats[it].  Sample s uses Philox counter offset + s.
 * warmup != 0: n_warm plain uniform samples y = u*0.999999 instead
 * (vegas.py:211-266), no stratification. */
int tq_vegas_iterate(const tq_vegas_state* st, const tq_integrand* f, const void* domain, uint64_t seed,
                     uint64_t offset, int32_t it, int32_t use_grid_improve, int32_t warmup, int64_t n_warm,
                     void* stream);

/* Split form for user integrands: write the samples' domain points
 * x[total, dim] (sample_dtype), Jacobians jac[total] (FP64), bin ids
 * ids[total, dim] (int32) and optionally y[total, dim].  u_in (optional,
 * [total, dim] sample_dtype) supplies the uniforms instead of the Philox
 * stream, for caller-provided RNG objects (vegas.py:98-102) ... */
int tq_vegas_points(const tq_vegas_state* st, const void* domain, uint64_t seed, uint64_t offset,
                    int32_t warmup, int64_t n_warm, const void* u_in, void* x_out, double* jac_out,
                    int32_t* ids_out, void* y_out, void* stream);
/* ... then consume the user's values f[total] (sample_dtype, already
 * evaluated at x): jf = f * volume * jac, histograms, per-cube sums. */
int tq_vegas_consume(const tq_vegas_state* st, const void* fvals, const void* domain, const double* jac,
                     const int32_t* ids, int32_t it, int32_t use_grid_improve, int32_t warmup, int64_t n_warm,
                     void* stream);

/* update_map (vegas_map.py:113-261): reduce the histogram partials of the
 * last n_samples samples (NULL count -> iter_stats[it].evals), smooth,
 * redistribute edges.  Flags count the "cannot update" and "infinite
 * edges" cases. */
int tq_vegas_update_map(const tq_vegas_state* st, int64_t n_samples, int32_t it, void* stream);

/* Map primitives on given y[n, dim] (sample_dtype), for VEGASMap's public
 * methods: ids (_get_interval_ID :76-85), offsets (_get_interval_offset
 * :87-97), x (get_X :44-58), jac (get_Jac :60-74).  Any output may be NULL. */
int tq_vegas_map_eval(const tq_vegas_state* st, const void* y, int64_t n, int64_t* ids_out, void* off_out,
                      void* x_out, void* jac_out, void* stream);
/* accumulate_weight (vegas_map.py:99-111) for given y[n, dim] and
 * jf2[n] (sample_dtype): adds into weights/counts. */
int tq_vegas_map_accumulate(const tq_vegas_state* st, const void* y, const void* jf2, int64_t n,
                            void* stream);
/* update_map from the accumulated weights/counts (no partial reduction),
 * then zero them (_reset_weight :174-183). */
int tq_vegas_map_update_accumulated(const tq_vegas_state* st, void* stream);
/* _smooth_map (vegas_map.py:113-172) on given weights/counts [dim][n]
 * (FP64): out [dim][n], scratch [dim][2n]; *not_updatable (device int)
 * becomes 1 when a weight sum is 0 (the reference then returns None). */
int tq_vegas_smooth(int32_t dim, int32_t n, const double* weights, const double* counts, double alpha,
                    double* out, double* scratch, int32_t* not_updatable, void* stream);

/* get_Y (vegas_stratification.py:140-165) for the nh/offsets of the last
 * tq_vegas_nh: y[total, dim] (sample_dtype), Philox counter offset + s, or
 * the caller's uniforms u_in[total, dim] when given. */
int tq_vegas_strat_y(const tq_vegas_state* st, uint64_t seed, uint64_t offset, const void* u_in, void* y_out,
                     void* stream);
/* accumulate_weight (vegas_stratification.py:46-70) for given per-sample
 * weights w[total] (FP64) in the offsets of the last tq_vegas_nh: JF, JF2
 * and iter_stats[it] result/sigma^2. */
int tq_vegas_strat_accumulate(const tq_vegas_state* st, const double* w, int32_t it, void* stream);
/* update_DH (vegas_stratification.py:72-90) from JF/JF2/nh except the
 * normalisation, which the next tq_vegas_nh or tq_vegas_normalize_dh
 * applies with iter_stats[it].dh_sum. */
int tq_vegas_update_dh(const tq_vegas_state* st, int32_t it, void* stream);
/* dh <- dh / iter_stats[it].dh_sum when the sum is non-zero. */
int tq_vegas_normalize_dh(const tq_vegas_state* st, int32_t it, void* stream);

/* ------------------------------------------------------------------------
 * Scatter-add: target[indices[i]] += source[i] for i < count, in place.
 * Replaces _add_at_indices (torchquad/integration/utils.py:61-107; its torch
 * branch is target.scatter_add_(0, indices, source), :75-76).  dtype:
 * TQ_F32, TQ_F64, TQ_I32 or TQ_I64 for target and source; indices int64.
 * Each wave sorts its 64 (index, value) pairs (skipped when they are
 * already ascending), sums runs of equal indices and adds each run with one
 * atomic.  Integer results are exact; float results depend on the order in
 * which waves reach a shared entry, as with scatter_add_ on a GPU.
 * Indices outside [0, target_len) are skipped; if status is not NULL the
 * kernel then sets *status (device int32, zero on entry) to 1.
 * ------------------------------------------------------------------------ */
int tq_add_at_indices(int32_t dtype, void* target, int64_t target_len, const int64_t* indices,
                      const void* source, int64_t count, int32_t* status, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TORCHQUAD_AMD_H */
