// Gauss-Legendre nodes and weights on [-1, 1] for any n, on the device.
//
// Near the ends (n sin(theta) < 50) and for n < 200: Newton iteration in
// theta (x = cos theta) on P_n from the three-term (Bonnet) recurrence
//     j P_j = (2j-1) x P_{j-1} - (j-1) P_{j-2},
// run on z = 1 - x = 2 sin^2(theta/2) and the differences D_j = P_j - P_{j-1},
//     D_j = ((j-1) D_{j-1} - (2j-1) z P_{j-1}) / j,
// so that nodes next to x = 1, where x itself rounds to a few ulp, still get
// their angle and weight to full relative precision.  Started from
// Tricomi's estimate
//     x_k ~ (1 - 1/(8 n^2) + 1/(8 n^3)) cos(pi (k - 1/4) / (n + 1/2)),
// with dP_n/dtheta = n (x P_n - P_{n-1}) / sin(theta) = n (D_n - z P_n) / sin(theta)
// and weights w_k = 2 / (1 - x_k^2) / P_n'(x_k)^2 = 2 / (dP_n/dtheta)^2.
// For n >= 200 the boundary nodes (n sin(theta) < 50, about 16 per end),
// started from Gatteschi's Bessel-zero estimate, use the same recurrence,
// spread over a workgroup: the step (D_{j-1}, P_{j-1})
// -> (D_j, P_j) is linear, so each of the 256 (1024 for n >= 16384) lanes
// forms the 2x2 product of its contiguous range of j, and an ordered tree
// (wave shuffles, then LDS across waves) multiplies the products --
// O(n/lanes + log) steps instead of O(n) on one lane.
// Interior nodes of large n: Newton in theta on the Stieltjes asymptotic
// expansion used by Hale & Townsend (SIAM J. Sci. Comput. 35(2), 2013)
//     P_n(cos t) = C_n sum_m h_m cos(a_m) / (2 sin t)^(m+1/2),
//     a_m = (n+m+1/2) t - (m+1/2) pi/2,  h_0 = 1,
//     h_m = h_{m-1} (m-1/2)^2 / (m (n+m+1/2)),
//     C_n = (4/pi) prod_{j=1..n} j/(j+1/2),
// truncated after kAsymTerms terms (relative error < 1e-17 for
// n sin t >= 50), with w = 2 / (dP_n/dt)^2.  O(1) work per interior node.
// Output is in ascending order like numpy.polynomial.legendre.leggauss,
// the reference's root function (gaussian.py:25).
#include "tq_host.hpp"

namespace tq {

// P_n(cos t) and dP_n/dt from the difference form of the recurrence.
__device__ inline void legendre_theta(int n, double t, double* pn, double* dpdt) {
    const double sh = sin(0.5 * t);
    const double z = 2.0 * sh * sh;
    double p = 1.0 - z, d = -z;  // P_1, D_1 = P_1 - P_0
    for (int j = 2; j <= n; ++j) {
        d = ((j - 1.0) * d - (2.0 * j - 1.0) * z * p) / (double)j;
        p += d;
    }
    *pn = p;
    *dpdt = (double)n * (d - z * p) / sin(t);
}

// P_n(cos t) and dP_n/dt by the workgroup-parallel product of the
// recurrence's transfer matrices.  Uniform in t across the workgroup;
// s_mat holds 4 doubles per wave.
// lanes per boundary node: 256, or 1024 for long recurrences (n >= 16384)
constexpr int kBoundaryLongN = 16384;

template <int BLOCK>
__device__ inline void legendre_theta_block(int n, double t, double* s_mat, double* pn, double* dpdt) {
    constexpr int kBoundaryWaves = BLOCK / kWave;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid >> 6;
    const double sh = sin(0.5 * t);
    const double z = 2.0 * sh * sh;
    const long long steps = (long long)n - 1;  // j = 2..n
    const long long c = (steps + BLOCK - 1) / BLOCK;
    const long long jlo = 2 + (long long)tid * c;
    const long long jhi = (jlo + c - 1 < (long long)n) ? (jlo + c - 1) : (long long)n;
    double p00 = 1.0, p01 = 0.0, p10 = 0.0, p11 = 1.0;
    for (long long j = jlo; j <= jhi; ++j) {
        const double jd = (double)j;
        // 1/j: hardware reciprocal refined by two Newton steps
        double r = __builtin_amdgcn_rcp(jd);
        r = fma(r, fma(-jd, r, 1.0), r);
        r = fma(r, fma(-jd, r, 1.0), r);
        const double a = 1.0 - r, bz = (2.0 - r) * z;
        // [d; p] <- [[a, -bz], [a, 1 - bz]] [d; p]:  d' = a d - bz p,  p' = p + d'
        const double n00 = fma(a, p00, -bz * p10);
        const double n01 = fma(a, p01, -bz * p11);
        p10 += n00;
        p11 += n01;
        p00 = n00;
        p01 = n01;
    }
    // ordered product within the wave: lane l <- P(l + s) * P(l)
    for (int s = 1; s < kWave; s <<= 1) {
        const double q00 = __shfl_down(p00, s, kWave), q01 = __shfl_down(p01, s, kWave);
        const double q10 = __shfl_down(p10, s, kWave), q11 = __shfl_down(p11, s, kWave);
        if ((lane & (2 * s - 1)) == 0) {
            const double m00 = fma(q00, p00, q01 * p10), m01 = fma(q00, p01, q01 * p11);
            const double m10 = fma(q10, p00, q11 * p10), m11 = fma(q10, p01, q11 * p11);
            p00 = m00;
            p01 = m01;
            p10 = m10;
            p11 = m11;
        }
    }
    if (lane == 0) {
        s_mat[4 * wave + 0] = p00;
        s_mat[4 * wave + 1] = p01;
        s_mat[4 * wave + 2] = p10;
        s_mat[4 * wave + 3] = p11;
    }
    __syncthreads();
    double t00 = s_mat[0], t01 = s_mat[1], t10 = s_mat[2], t11 = s_mat[3];
    for (int w = 1; w < kBoundaryWaves; ++w) {
        const double q00 = s_mat[4 * w], q01 = s_mat[4 * w + 1], q10 = s_mat[4 * w + 2], q11 = s_mat[4 * w + 3];
        const double m00 = fma(q00, t00, q01 * t10), m01 = fma(q00, t01, q01 * t11);
        const double m10 = fma(q10, t00, q11 * t10), m11 = fma(q10, t01, q11 * t11);
        t00 = m00;
        t01 = m01;
        t10 = m10;
        t11 = m11;
    }
    __syncthreads();  // s_mat is rewritten by the next call
    // [d_1; p_1] = [-z; 1 - z]
    const double d = fma(t00, -z, t01 * (1.0 - z));
    const double p = fma(t10, -z, t11 * (1.0 - z));
    *pn = p;
    *dpdt = (double)n * (d - z * p) / sin(t);
}

constexpr int kAsymTerms = 14;
constexpr int kAsymMinN = 200;
constexpr double kAsymMinNSin = 50.0;

// f(t) = sum_m h_m cos(a_m) / (2 sin t)^(m+1/2) and f'(t).
__device__ inline void stieltjes(int n, double t, double* f, double* fp) {
    const double s = sin(t), c = cos(t);
    const double inv2s = 0.5 / s;
    const double nn = (double)n;
    double pw = sqrt(inv2s);
    double h = 1.0;
    const double a0 = (nn + 0.5) * t - 0.25 * M_PI;
    double ca = cos(a0), sa = sin(a0);
    double F = 0.0, FP = 0.0;
    for (int m = 0; m < kAsymTerms; ++m) {
        if (m > 0) {
            h *= ((m - 0.5) * (m - 0.5)) / ((double)m * (nn + m + 0.5));
            pw *= inv2s;
            // a_m = a_{m-1} + t - pi/2
            const double ca2 = ca * s + sa * c;
            const double sa2 = sa * s - ca * c;
            ca = ca2;
            sa = sa2;
        }
        const double term = h * pw;
        F += term * ca;
        FP += term * (-(nn + m + 0.5) * sa - (m + 0.5) * ca * (c / s));
    }
    *f = F;
    *fp = FP;
}

// Tricomi's x estimate moved to theta to first order
__device__ inline double tricomi_theta(int n, int k) {
    const double nn = (double)n;
    const double t0 = M_PI * ((double)k + 0.75) / (nn + 0.5);
    const double eps = 1.0 / (8.0 * nn * nn) - 1.0 / (8.0 * nn * nn * nn);
    return (n == 1) ? 0.5 * M_PI : t0 + eps * cos(t0) / sin(t0);
}

// Zeros j_{0,k}, k = 1..24, of the Bessel function J_0 (A&S table 9.5).
__constant__ double kBesselJ0Zeros[24] = {
    2.4048255576957724, 5.5200781102863106, 8.6537279129110125, 11.791534439014281, 14.930917708487787,
    18.071063967910924, 21.211636629879258, 24.352471530749302, 27.493479132040253, 30.634606468431976,
    33.775820213573567, 36.917098353664045, 40.05842576462824,  43.19979171317673,  46.341188371661815,
    49.482609897397815, 52.624051841114998, 55.765510755019982, 58.90698392608094,  62.048469190227166,
    65.189964800206866, 68.331469329856802, 71.472981603593738, 74.614500643701831};

// j_{0,k} (k >= 1): the table, then McMahon's expansion (A&S 9.5.12) with
// beta = (k - 1/4) pi, good to 1e-14 relative from k = 20 on.
__device__ inline double bessel_j0_zero(int k) {
    if (k <= 24) return kBesselJ0Zeros[k - 1];
    const double b = ((double)k - 0.25) * M_PI, e = 8.0 * b, e2 = e * e;
    return b + 1.0 / e - 124.0 / (3.0 * e * e2) + 120928.0 / (15.0 * e * e2 * e2);
}

// Gatteschi's estimate of the angle of the k-th node from x = 1 (0-based):
// theta_k ~ j_{0,k+1} / sqrt((n + 1/2)^2 + 1/12), relative error about
// j^2 / (720 nu^4) -- near 1e-9 at n = 200 and the 16th node, far below at
// larger n -- so Newton needs one step where Tricomi's interior estimate
// needed three or four.
__device__ inline double gatteschi_theta(int n, int k) {
    const double rho = (double)n + 0.5;
    return bessel_j0_zero(k + 1) / sqrt(rho * rho + 1.0 / 12.0);
}

// Boundary nodes k < nb of n >= kAsymMinN: one workgroup per node.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_gl_boundary(int n, double* xs, double* ws) {
    __shared__ double s_mat[4 * (BLOCK / kWave)];
    const int k = blockIdx.x;
    double t = gatteschi_theta(n, k);
    double pn, dp;
    for (int it = 0; it < 16; ++it) {
        legendre_theta_block<BLOCK>(n, t, s_mat, &pn, &dp);
        const double dt = pn / dp;
        t -= dt;
        if (fabs(dt) <= 1e-11 * t) break;
    }
    legendre_theta_block<BLOCK>(n, t, s_mat, &pn, &dp);
    if (threadIdx.x == 0) {
        const double x = cos(t);
        const double w = 2.0 / (dp * dp);
        xs[n - 1 - k] = x;
        ws[n - 1 - k] = w;
        xs[k] = -x;
        ws[k] = w;
    }
}

// Nodes k in [k_first, half): asymptotic expansion for n >= kAsymMinN (the
// boundary kernel covers k < k_first), the per-lane recurrence for small n.
__global__ __launch_bounds__(kBlock) void k_gl_nodes(int n, int k_first, double cn, double* xs, double* ws) {
    const int k = k_first + blockIdx.x * blockDim.x + threadIdx.x;
    const int half = (n + 1) / 2;
    if (k >= half) return;
    const double nn = (double)n;
    const double t0 = M_PI * ((double)k + 0.75) / (nn + 0.5);
    if (n >= kAsymMinN) {
        double t = t0, f, fp;
        for (int it = 0; it < 6; ++it) {
            stieltjes(n, t, &f, &fp);
            const double dt = f / fp;
            t -= dt;
            if (fabs(dt) <= 1e-17 * t) break;
        }
        stieltjes(n, t, &f, &fp);
        double x = cos(t);
        const double dpdt = cn * fp;
        const double w = 2.0 / (dpdt * dpdt);
        if ((n & 1) && k == half - 1) x = 0.0;
        xs[n - 1 - k] = x;
        ws[n - 1 - k] = w;
        xs[k] = -x;
        ws[k] = w;
        return;
    }
    double t = tricomi_theta(n, k);
    double pn, dp;
    // quadratic convergence: once a step is below 1e-11 t the next error is
    // below rounding, and the recurrence's own rounding keeps steps from
    // shrinking to an ulp, so stop there
    for (int it = 0; it < 16; ++it) {
        legendre_theta(n, t, &pn, &dp);
        const double dt = pn / dp;
        t -= dt;
        if (fabs(dt) <= 1e-11 * t) break;
    }
    legendre_theta(n, t, &pn, &dp);
    double x = cos(t);
    const double w = 2.0 / (dp * dp);
    if ((n & 1) && k == half - 1) x = 0.0;
    xs[n - 1 - k] = x;
    ws[n - 1 - k] = w;
    xs[k] = -x;
    ws[k] = w;
}

}  // namespace tq

using namespace tq;

extern "C" int tq_gauss_legendre(int32_t n, double* nodes, double* weights, void* stream) {
    if (n < 1) return fail(TQ_E_ARG, "n must be >= 1");
    if (!nodes || !weights) return fail(TQ_E_NULL, "nodes/weights is NULL");
    const int half = (n + 1) / 2;
    // C_n = (4/pi) prod_{j=1..n} j/(j+1/2) = (2/sqrt(pi)) Gamma(n+1)/Gamma(n+3/2).
    // The log of the ratio from Stirling's series, -ln(n)/2 + sum_k c_k n^-k
    // with c_k = (-1)^(k+1) (B_{k+1}(1) - B_{k+1}(3/2)) / (k (k+1)) (Bernoulli
    // polynomials); eight terms leave a relative error below 1e-23 from
    // n = 200 on, in O(1) host work instead of n extended-precision steps.
    long double cn = 0.0L;
    int nb = 0;  // boundary nodes (n sin(theta_k) < 50) of the workgroup kernel
    if (n >= kAsymMinN) {
        static const long double kC[8] = {-3.0L / 8.0L,  1.0L / 8.0L,   -3.0L / 64.0L,     1.0L / 64.0L,
                                          -3.0L / 640.0L, 1.0L / 384.0L, -33.0L / 14336.0L, 1.0L / 2048.0L};
        const long double inv = 1.0L / (long double)n;
        long double series = 0.0L;
        for (int k = 7; k >= 0; --k) series = (series + kC[k]) * inv;
        cn = 2.0L / sqrtl(3.14159265358979323846264338327950288L) * expl(series - 0.5L * logl((long double)n));
        while (nb < half && (double)n * sin(M_PI * (nb + 0.75) / (n + 0.5)) < kAsymMinNSin) ++nb;
        if (n >= kBoundaryLongN)
            hipLaunchKernelGGL((k_gl_boundary<1024>), dim3(nb), dim3(1024), 0, as_stream(stream), (int)n, nodes, weights);
        else
            hipLaunchKernelGGL((k_gl_boundary<kBlock>), dim3(nb), dim3(kBlock), 0, as_stream(stream), (int)n, nodes,
                               weights);
        int rc = check_launch("k_gl_boundary");
        if (rc) return rc;
    }
    const int rest = half - nb;
    if (rest > 0) {
        const int blocks = (rest + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_gl_nodes, dim3(blocks), dim3(kBlock), 0, as_stream(stream), (int)n, nb, (double)cn,
                           nodes, weights);
        return check_launch("k_gl_nodes");
    }
    return TQ_OK;
}
