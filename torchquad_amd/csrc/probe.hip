// Measurement hooks: the attainable rate, on the device at hand, of the
// fused grid kernel's inner point loop and of the instructions it is made
// of.  bench.py states the grid kernel's rate against these.
#include "tq_host.hpp"

namespace tq {

constexpr int kProbeTable = 256;  // loop-table entries of the point-loop probe

// KIND 0: eight independent v_fma_f32 chains per lane; 1: eight v_exp_f32
// chains; 2: the grid loop's body per point with an (term, weight) table --
// the lane's state plus an LDS table term (wave-uniform address), exp2, FMA
// with the table weight into the lane's partial; 3: the body with the weight
// folded into the term (4-byte entries, exp2, add), as the kernel runs it for
// the Gaussian peak.  Each trip is 32 instructions (0, 1) or 256 points (2, 3).
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_probe(float* out, int iters) {
    __shared__ float2 tab[kProbeTable];
    __shared__ float tab1[kProbeTable];
    for (int j = threadIdx.x; j < kProbeTable; j += kBlock) {
        tab[j] = make_float2(-0.001f * (float)j, 1.0f / (1.0f + (float)j));
        tab1[j] = -0.001f * (float)j - log2f(1.0f + (float)j);
    }
    __syncthreads();
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = 0.1f + 0.01f * (float)((threadIdx.x + k) & 31);
    const float s_rl = -0.01f * (float)(threadIdx.x & 63);
    float acc = 0.0f;
    for (int i = 0; i < iters; ++i) {
        if (KIND == 2) {
#pragma unroll 16
            for (int j = 0; j < kProbeTable; ++j) {
                const float2 e = tab[j];
                acc = fmaf(e.y, __builtin_amdgcn_exp2f(s_rl + e.x), acc);
            }
        } else if (KIND == 3) {
#pragma unroll 16
            for (int j = 0; j < kProbeTable; ++j) acc += __builtin_amdgcn_exp2f(s_rl + tab1[j]);
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    a[k] = KIND == 0 ? fmaf(a[k], 0.999f, 0.0005f) : __builtin_amdgcn_exp2f(a[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += a[k];
    out[(long long)blockIdx.x * kBlock + threadIdx.x] = acc;
}

}  // namespace tq

using namespace tq;

extern "C" int tq_probe_rates(int32_t kind, int64_t blocks, int32_t iters, float* out, void* stream) {
    if (!out) return fail(TQ_E_NULL, "out is NULL");
    if (kind < 0 || kind > 3) return fail(TQ_E_ARG, "kind must be 0, 1, 2 or 3");
    if (blocks < 1 || blocks > (1 << 20) || iters < 1 || iters > (1 << 20))
        return fail(TQ_E_ARG, "blocks and iters must be in 1..2^20");
    const hipStream_t s = as_stream(stream);
    if (kind == 0) hipLaunchKernelGGL(k_probe<0>, dim3((unsigned)blocks), dim3(kBlock), 0, s, out, iters);
    else if (kind == 1) hipLaunchKernelGGL(k_probe<1>, dim3((unsigned)blocks), dim3(kBlock), 0, s, out, iters);
    else if (kind == 2) hipLaunchKernelGGL(k_probe<2>, dim3((unsigned)blocks), dim3(kBlock), 0, s, out, iters);
    else hipLaunchKernelGGL(k_probe<3>, dim3((unsigned)blocks), dim3(kBlock), 0, s, out, iters);
    return check_launch("k_probe");
}
