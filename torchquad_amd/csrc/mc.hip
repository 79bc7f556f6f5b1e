// Monte Carlo: counter-based sampling fused with the integrand and a block
// reduction of (sum f, sum f^2).  No sample ever touches HBM on the fused
// path; the point writer produces the identical stream for user integrands.
#include "tq_host.hpp"

namespace tq {

constexpr int kMcFlush = 32;

template <typename T>
struct McArgs {
    int d;
    long long N;
    long long spt;  // samples per thread
    uint64_t seed, offset;
    const uint64_t* offset_dev;  // when set, the first counter is read from here
    const T* domain;
    double* partials;  // [blocks][2]
    Params<T> p;
};

// MAXD: compile-time bound on the dimension (4, 8 or 16), so the unrolled
// per-coordinate work is not predicated over 16 coordinates for d <= 8.
template <typename T, int FAM, int MAXD>
__global__ __launch_bounds__(kBlock) void k_mc(McArgs<T> A) {
    __shared__ double s_red[kWavesPerBlock];
    const int d = A.d;
    T lo[MAXD], sz[MAXD];
#pragma unroll
    for (int k = 0; k < MAXD; ++k) {
        if (k < d) {
            lo[k] = A.domain[2 * k];
            sz[k] = A.domain[2 * k + 1] - A.domain[2 * k];
        } else {
            lo[k] = T(0);
            sz[k] = T(0);
        }
    }
    const uint64_t base = A.offset_dev ? *A.offset_dev : A.offset;
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long s0 = gt * A.spt;
    const long long s1 = (s0 + A.spt < A.N) ? (s0 + A.spt) : A.N;
    double acc = 0.0, acc2 = 0.0;
    for (long long sb = s0; sb < s1; sb += kMcFlush) {
        const long long se = (sb + kMcFlush < s1) ? (sb + kMcFlush) : s1;
        T fa = T(0), fa2 = T(0);
        for (long long s = sb; s < se; ++s) {
            T x[MAXD];
            Uniform<T>::template draw<MAXD>(A.seed, base + (uint64_t)s, d, x);
#pragma unroll
            for (int k = 0; k < MAXD; ++k)
                if (k < d) x[k] = fma(x[k], sz[k], lo[k]);
            const T f = eval_point<FAM, T, MAXD>(x, d, A.p);
            fa += f;
            fa2 = fma(f, f, fa2);
        }
        acc += (double)fa;
        acc2 += (double)fa2;
    }
    const double b1 = block_sum(acc, s_red);
    const double b2 = block_sum(acc2, s_red);
    if (threadIdx.x == 0) {
        A.partials[2 * blockIdx.x] = b1;
        A.partials[2 * blockIdx.x + 1] = b2;
    }
}

template <typename T>
__global__ __launch_bounds__(1024) void k_mc_finalize(const double* partials, long long nb, const T* domain,
                                                       int d, long long N, double c0, T* out, double* stats) {
    __shared__ double s_red[1024 / kWave];
    double v1 = 0.0, v2 = 0.0;
    for (long long i = threadIdx.x; i < nb; i += blockDim.x) {
        v1 += partials[2 * i];
        v2 += partials[2 * i + 1];
    }
    v1 = block_sum(v1, s_red);
    v2 = block_sum(v2, s_red);
    if (threadIdx.x == 0) {
        double vol = 1.0;
        for (int k = 0; k < d; ++k) vol *= (double)domain[2 * k + 1] - (double)domain[2 * k];
        const double r = c0 * vol * v1 / (double)N;
        if (out) *out = (T)r;
        if (stats) {
            stats[0] = c0 * vol * v1;
            stats[1] = c0 * c0 * vol * vol * v2;
            stats[2] = (double)N;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mc_points(const T* domain, int d, long long N, uint64_t seed,
                                                      uint64_t offset, T* pts) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < N; s += stride) {
        T u[TQ_MAX_DIM];
        Uniform<T>::template draw<TQ_MAX_DIM>(seed, offset + (uint64_t)s, d, u);
#pragma unroll
        for (int k = 0; k < TQ_MAX_DIM; ++k) {
            if (k < d) {
                const T lo = domain[2 * k];
                const T sz = domain[2 * k + 1] - lo;
                pts[s * d + k] = fma(u[k], sz, lo);
            }
        }
    }
}

// Generic-width uniform stream: row r, column c.
template <typename T>
__device__ inline T uniform_at(uint64_t seed, uint64_t row, int c);
template <>
__device__ inline float uniform_at<float>(uint64_t seed, uint64_t row, int c) {
    const u32x4 r = philox_block(seed, row, (uint32_t)(c >> 2));
    return u01f(r.v[c & 3]);
}
template <>
__device__ inline double uniform_at<double>(uint64_t seed, uint64_t row, int c) {
    const u32x4 r = philox_block(seed, row, (uint32_t)(c >> 1));
    return (c & 1) ? u01d(r.v[2], r.v[3]) : u01d(r.v[0], r.v[1]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_uniform(uint64_t seed, uint64_t offset, long long rows, int cols,
                                                    T* out) {
    const long long total = rows * (long long)cols;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const long long r = e / cols;
        const int c = (int)(e - r * cols);
        out[e] = uniform_at<T>(seed, offset + (uint64_t)r, c);
    }
}

// Column sums of f[N, S].  Narrow S: each thread keeps one fixed column
// (block size is a multiple of S).  Wide S: one thread per column, rows
// split over blockIdx.y.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_colsum_narrow(const T* fv, long long N, long long S,
                                                          double* partials /* [blocks][S] */) {
    __shared__ double s_acc[kBlock];
    const long long bd = blockDim.x;
    const long long total = N * S;
    const long long stride = (long long)gridDim.x * bd;
    double acc = 0.0;
    for (long long e = (long long)blockIdx.x * bd + threadIdx.x; e < total; e += stride) acc += (double)fv[e];
    s_acc[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < S) {
        double v = 0.0;
        for (long long i = threadIdx.x; i < bd; i += S) v += s_acc[i];
        partials[(long long)blockIdx.x * S + threadIdx.x] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_colsum_wide(const T* fv, long long N, long long S, long long rpb,
                                                        double* partials /* [gridDim.y][S] */) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const long long r0 = (long long)blockIdx.y * rpb;
    const long long r1 = (r0 + rpb < N) ? (r0 + rpb) : N;
    double acc = 0.0;
    for (long long r = r0; r < r1; ++r) acc += (double)fv[r * S + s];
    partials[(long long)blockIdx.y * S + s] = acc;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_colsum_finalize(const double* partials, long long nb, long long S,
                                                            const T* domain, int d, long long N, T* out) {
    __shared__ double s_red[kWavesPerBlock];
    const long long s = blockIdx.x;
    double v = 0.0;
    for (long long b = threadIdx.x; b < nb; b += blockDim.x) v += partials[b * S + s];
    v = block_sum(v, s_red);
    if (threadIdx.x == 0) {
        double vol = 1.0;
        for (int k = 0; k < d; ++k) vol *= (double)domain[2 * k + 1] - (double)domain[2 * k];
        out[s] = (T)(vol * v / (double)N);
    }
}

static long long mc_spt(const tq_mc_desc* m) {
    if (m->samples_per_thread > 0) return m->samples_per_thread;
    // aim for ~512K threads (8 waves per SIMD), at least 16 samples each
    long long spt = (m->N + (1LL << 19) - 1) >> 19;
    if (spt < 16) spt = 16;
    return spt;
}

static long long mc_blocks(const tq_mc_desc* m) {
    const long long spt = mc_spt(m);
    const long long threads = (m->N + spt - 1) / spt;
    long long b = (threads + kBlock - 1) / kBlock;
    return b < 1 ? 1 : b;
}

static int check_mc(const tq_mc_desc* m) {
    if (!m) return fail(TQ_E_NULL, "mc descriptor is NULL");
    if (m->dtype != TQ_F32 && m->dtype != TQ_F64) return fail(TQ_E_ARG, "dtype must be TQ_F32 or TQ_F64");
    if (m->dim < 1 || m->dim > TQ_MAX_DIM) return fail(TQ_E_ARG, "dim out of range (1..16)");
    if (m->N < 1) return fail(TQ_E_ARG, "N must be >= 1");
    if (mc_blocks(m) > 0x7fffffffLL) return fail(TQ_E_ARG, "N too large for one launch; raise samples_per_thread");
    return TQ_OK;
}

static void colsum_shape(long long N, long long S, long long* nb, int* bd, long long* rpb) {
    if (S < 64) {
        *bd = (int)(kBlock - (kBlock % S));
        long long b = (N * S + (long long)(*bd) * 16 - 1) / ((long long)(*bd) * 16);
        if (b > 4096) b = 4096;
        if (b < 1) b = 1;
        *nb = b;
        *rpb = 0;
    } else {
        *bd = kBlock;
        long long ry = (N + 255) / 256;
        if (ry > 1024) ry = 1024;
        if (ry < 1) ry = 1;
        *rpb = (N + ry - 1) / ry;
        *nb = (N + *rpb - 1) / *rpb;
    }
}

template <typename T>
static int launch_mc(const tq_mc_desc* m, const tq_integrand* f, const void* domain, void* ws, size_t wsb,
                     void* out, double* stats, hipStream_t st, const uint64_t* offset_dev = nullptr) {
    const long long nb = mc_blocks(m);
    if (!ws || wsb < (size_t)nb * 2 * sizeof(double)) return fail(TQ_E_WORKSPACE, "mc workspace too small");
    McArgs<T> A;
    A.d = m->dim;
    A.N = m->N;
    A.spt = mc_spt(m);
    A.seed = m->seed;
    A.offset = m->offset;
    A.offset_dev = offset_dev;
    A.domain = (const T*)domain;
    A.partials = (double*)ws;
    A.p = make_params<T>(*f);
    TQ_DISPATCH_FAMILY(f->family, {
        if (m->dim <= 4)
            hipLaunchKernelGGL((k_mc<T, FAM, 4>), dim3((unsigned)nb), dim3(kBlock), 0, st, A);
        else if (m->dim <= 8)
            hipLaunchKernelGGL((k_mc<T, FAM, 8>), dim3((unsigned)nb), dim3(kBlock), 0, st, A);
        else
            hipLaunchKernelGGL((k_mc<T, FAM, TQ_MAX_DIM>), dim3((unsigned)nb), dim3(kBlock), 0, st, A);
    });
    int rc = check_launch("k_mc");
    if (rc) return rc;
    hipLaunchKernelGGL((k_mc_finalize<T>), dim3(1), dim3(1024), 0, st, (const double*)ws, nb, (const T*)domain,
                       m->dim, m->N, f->c0, (T*)out, stats);
    return check_launch("k_mc_finalize");
}

template <typename T>
static int launch_colsum(const void* fv, long long N, long long S, const void* domain, int d, void* ws,
                         size_t wsb, void* out, hipStream_t st) {
    long long nb, rpb;
    int bd;
    colsum_shape(N, S, &nb, &bd, &rpb);
    if (!ws || wsb < (size_t)(nb * S) * sizeof(double)) return fail(TQ_E_WORKSPACE, "reduce workspace too small");
    if (S < 64) {
        hipLaunchKernelGGL((k_colsum_narrow<T>), dim3((unsigned)nb), dim3(bd), 0, st, (const T*)fv, N, S,
                           (double*)ws);
    } else {
        dim3 grid((unsigned)((S + kBlock - 1) / kBlock), (unsigned)nb);
        hipLaunchKernelGGL((k_colsum_wide<T>), grid, dim3(kBlock), 0, st, (const T*)fv, N, S, rpb, (double*)ws);
    }
    int rc = check_launch("k_colsum");
    if (rc) return rc;
    hipLaunchKernelGGL((k_colsum_finalize<T>), dim3((unsigned)S), dim3(kBlock), 0, st, (const double*)ws, nb, S,
                       (const T*)domain, d, N, (T*)out);
    return check_launch("k_colsum_finalize");
}

}  // namespace tq

using namespace tq;

extern "C" {

size_t tq_mc_workspace_size(const tq_mc_desc* m) {
    if (check_mc(m)) return 0;
    return (size_t)mc_blocks(m) * 2 * sizeof(double);
}

int tq_mc_integrate(const tq_mc_desc* m, const tq_integrand* f, const void* domain, void* workspace,
                    size_t workspace_bytes, void* out, double* stats, void* stream) {
    int rc = check_mc(m);
    if (rc) return rc;
    rc = check_integrand(f, m->dim);
    if (rc) return rc;
    if (!domain) return fail(TQ_E_NULL, "domain is NULL");
    if (!out && !stats) return fail(TQ_E_NULL, "no output pointer");
    if (m->dtype == TQ_F32)
        return launch_mc<float>(m, f, domain, workspace, workspace_bytes, out, stats, as_stream(stream));
    return launch_mc<double>(m, f, domain, workspace, workspace_bytes, out, stats, as_stream(stream));
}

int tq_mc_integrate_dev_offset(const tq_mc_desc* m, const tq_integrand* f, const void* domain,
                               const uint64_t* offset_dev, void* workspace, size_t workspace_bytes, void* out,
                               double* stats, void* stream) {
    int rc = check_mc(m);
    if (rc) return rc;
    rc = check_integrand(f, m->dim);
    if (rc) return rc;
    if (!domain || !offset_dev) return fail(TQ_E_NULL, "domain/offset_dev is NULL");
    if (!out && !stats) return fail(TQ_E_NULL, "no output pointer");
    if (m->dtype == TQ_F32)
        return launch_mc<float>(m, f, domain, workspace, workspace_bytes, out, stats, as_stream(stream), offset_dev);
    return launch_mc<double>(m, f, domain, workspace, workspace_bytes, out, stats, as_stream(stream), offset_dev);
}

int tq_mc_points(const tq_mc_desc* m, const void* domain, void* points, void* stream) {
    int rc = check_mc(m);
    if (rc) return rc;
    if (!domain || !points) return fail(TQ_E_NULL, "domain/points is NULL");
    long long blocks = (m->N + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    if (m->dtype == TQ_F32)
        hipLaunchKernelGGL((k_mc_points<float>), dim3((unsigned)blocks), dim3(kBlock), 0, as_stream(stream),
                           (const float*)domain, m->dim, m->N, m->seed, m->offset, (float*)points);
    else
        hipLaunchKernelGGL((k_mc_points<double>), dim3((unsigned)blocks), dim3(kBlock), 0, as_stream(stream),
                           (const double*)domain, m->dim, m->N, m->seed, m->offset, (double*)points);
    return check_launch("k_mc_points");
}

size_t tq_mc_reduce_workspace_size(int64_t N, int64_t S) {
    if (N < 1 || S < 1) return 0;
    long long nb, rpb;
    int bd;
    colsum_shape(N, S, &nb, &bd, &rpb);
    return (size_t)(nb * S) * sizeof(double);
}

int tq_mc_reduce(int32_t dtype, const void* fvals, int64_t N, int64_t S, const void* domain, int32_t dim,
                 void* workspace, size_t workspace_bytes, void* out, void* stream) {
    if (dtype != TQ_F32 && dtype != TQ_F64) return fail(TQ_E_ARG, "dtype must be TQ_F32 or TQ_F64");
    if (N < 1 || S < 1 || S > 65535) return fail(TQ_E_ARG, "N/S out of range");
    if (dim < 1) return fail(TQ_E_ARG, "dim must be >= 1");
    if (!fvals || !domain || !out) return fail(TQ_E_NULL, "fvals/domain/out is NULL");
    if (dtype == TQ_F32)
        return launch_colsum<float>(fvals, N, S, domain, dim, workspace, workspace_bytes, out, as_stream(stream));
    return launch_colsum<double>(fvals, N, S, domain, dim, workspace, workspace_bytes, out, as_stream(stream));
}

int tq_uniform(uint64_t seed, uint64_t offset, int32_t dtype, int64_t rows, int32_t cols, void* out,
               void* stream) {
    if (dtype != TQ_F32 && dtype != TQ_F64) return fail(TQ_E_ARG, "dtype must be TQ_F32 or TQ_F64");
    if (rows < 0 || cols < 1) return fail(TQ_E_ARG, "rows/cols out of range");
    if (rows == 0) return TQ_OK;
    if (!out) return fail(TQ_E_NULL, "out is NULL");
    long long blocks = (rows * cols + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    if (dtype == TQ_F32)
        hipLaunchKernelGGL((k_uniform<float>), dim3((unsigned)blocks), dim3(kBlock), 0, as_stream(stream), seed,
                           offset, (long long)rows, (int)cols, (float*)out);
    else
        hipLaunchKernelGGL((k_uniform<double>), dim3((unsigned)blocks), dim3(kBlock), 0, as_stream(stream), seed,
                           offset, (long long)rows, (int)cols, (double*)out);
    return check_launch("k_uniform");
}

}  // extern "C"
