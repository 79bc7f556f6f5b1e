// Scatter-add of a value vector into a target vector at given indices, the
// operation VEGAS uses for its map and stratification histograms
// (torchquad/integration/utils.py:61-107).  Each wave takes 64 consecutive
// (index, value) pairs, sorts them by index with a bitonic network over the
// lanes unless they already ascend, sums every run of equal indices with a
// segmented scan, and the last lane of each run adds the run's sum with one
// atomic: a wave that hits one entry 64 times issues one atomic, not 64.
#include "tq_host.hpp"

namespace tq {

template <typename T>
__device__ inline void atomic_add_to(T* p, T v) {
    atomicAdd(p, v);
}
template <>
__device__ inline void atomic_add_to<long long>(long long* p, long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

// keys are (index << 6 | lane); an invalid lane carries the largest index
constexpr unsigned long long kNoIndex = (1ULL << 57) - 1;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_add_at(T* target, long long len, const long long* idx, const T* src,
                                                  long long count, int* status) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long stride = (long long)gridDim.x * kBlock;
    bool bad = false;
    for (long long base = (long long)blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); base < count;
         base += stride) {
        const long long i = base + lane;
        unsigned long long id = kNoIndex;
        T v = T(0);
        if (i < count) {
            const long long j = idx[i];
            if (j >= 0 && j < len) {
                id = (unsigned long long)j;
                v = src[i];
            } else {
                bad = true;
            }
        }
        unsigned long long key = (id << 6) | (unsigned long long)lane;
        const unsigned long long below = __shfl_up(key, 1, kWave);
        if (__ballot(lane > 0 && below > key)) {
            for (int size = 2; size <= kWave; size <<= 1) {
                for (int st = size >> 1; st > 0; st >>= 1) {
                    const unsigned long long ok = __shfl_xor(key, st, kWave);
                    const T ov = __shfl_xor(v, st, kWave);
                    const bool up = (lane & size) == 0;
                    const bool lower = (lane & st) == 0;
                    if ((lower == up) ? (ok < key) : (ok > key)) {
                        key = ok;
                        v = ov;
                    }
                }
            }
        }
        const unsigned long long run = key >> 6;
        T sum = v;
        for (int off = 1; off < kWave; off <<= 1) {
            const T ps = __shfl_up(sum, off, kWave);
            const unsigned long long pr = __shfl_up(run, off, kWave);
            if (lane >= off && pr == run) sum += ps;
        }
        const unsigned long long next = __shfl_down(run, 1, kWave);
        if ((lane == kWave - 1 || next != run) && run != kNoIndex) atomic_add_to(&target[run], sum);
    }
    if (bad && status) atomicOr(status, 1);
}

template <typename T>
static int launch_add_at(void* target, int64_t len, const int64_t* idx, const void* src, int64_t count,
                         int32_t* status, hipStream_t s) {
    long long blocks = (count + kBlock - 1) / kBlock;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL((k_add_at<T>), dim3((unsigned)blocks), dim3(kBlock), 0, s, (T*)target, (long long)len,
                       (const long long*)idx, (const T*)src, (long long)count, (int*)status);
    return check_launch("k_add_at");
}

}  // namespace tq

using namespace tq;

extern "C" int tq_add_at_indices(int32_t dtype, void* target, int64_t target_len, const int64_t* indices,
                                 const void* source, int64_t count, int32_t* status, void* stream) {
    if (count < 0 || target_len < 0) return fail(TQ_E_ARG, "count and target_len must be >= 0");
    if (count == 0) return TQ_OK;
    if (!target || !indices || !source) return fail(TQ_E_NULL, "target/indices/source is NULL");
    if (target_len >= (int64_t)kNoIndex) return fail(TQ_E_ARG, "target_len must be < 2^57 - 1");
    const hipStream_t s = as_stream(stream);
    switch (dtype) {
        case TQ_F32: return launch_add_at<float>(target, target_len, indices, source, count, status, s);
        case TQ_F64: return launch_add_at<double>(target, target_len, indices, source, count, status, s);
        case TQ_I32: return launch_add_at<int>(target, target_len, indices, source, count, status, s);
        case TQ_I64: return launch_add_at<long long>(target, target_len, indices, source, count, status, s);
        default: return fail(TQ_E_ARG, "dtype must be TQ_F32, TQ_F64, TQ_I32 or TQ_I64");
    }
}
