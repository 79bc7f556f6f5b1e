// Gauss-Legendre nodes and weights on [-1, 1] for any n, computed on the
// device in O(n) per node: Newton iteration on P_n from the three-term
// (Bonnet) recurrence j P_j = (2j-1) x P_{j-1} - (j-1) P_{j-2}, started
// from Tricomi's asymptotic node estimate
//     x_k ~ (1 - 1/(8 n^2) + 1/(8 n^3)) cos(pi (k - 1/4) / (n + 1/2)),
// with P_n'(x) = n (x P_n - P_{n-1}) / (x^2 - 1) and weights
//     w_k = 2 / ((1 - x_k^2) P_n'(x_k)^2).
// Output is in ascending order like numpy.polynomial.legendre.leggauss,
// the reference's root function (gaussian.py:25).
#include "tq_host.hpp"

namespace tq {

__device__ inline void legendre_pair(int n, double x, double* pn, double* pn1) {
    double p0 = 1.0, p1 = x;
    for (int j = 2; j <= n; ++j) {
        const double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / (double)j;
        p0 = p1;
        p1 = p2;
    }
    *pn = p1;
    *pn1 = p0;
}

__global__ __launch_bounds__(kBlock) void k_gl_nodes(int n, double* xs, double* ws) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int half = (n + 1) / 2;
    if (k >= half) return;
    const double nn = (double)n;
    double x = (1.0 - 1.0 / (8.0 * nn * nn) + 1.0 / (8.0 * nn * nn * nn)) *
               cos(M_PI * ((double)k + 0.75) / (nn + 0.5));
    double pn, pn1, dp;
    for (int it = 0; it < 16; ++it) {
        legendre_pair(n, x, &pn, &pn1);
        dp = nn * (x * pn - pn1) / (x * x - 1.0);
        const double dx = pn / dp;
        x -= dx;
        if (fabs(dx) <= 1e-16) break;
    }
    legendre_pair(n, x, &pn, &pn1);
    dp = nn * (x * pn - pn1) / (x * x - 1.0);
    const double w = 2.0 / ((1.0 - x * x) * dp * dp);
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
    const int blocks = (half + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_gl_nodes, dim3(blocks), dim3(kBlock), 0, as_stream(stream), (int)n, nodes, weights);
    return check_launch("k_gl_nodes");
}
