// Host-side helpers for the C ABI: thread-local error text, argument checks,
// parameter conversion and family dispatch.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "tq_device.hpp"

namespace tq {

void set_error(const char* fmt, ...);

inline int fail(int code, const char* msg) {
    set_error("%s", msg);
    return code;
}

inline int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return (int)e;
    }
    return TQ_OK;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr double kLog2e = 1.4426950408889634074;

// Family-specific derived constants computed in double precision before
// rounding to T: Params::b, the oscillatory phase (radians in FP64,
// revolutions in FP32), and log2(e) folded into the exponential families so
// that their finish is one 2^x.
template <typename T>
inline Params<T> make_params(const tq_integrand& f) {
    Params<T> p{};
    p.c0 = (T)f.c0;
    p.c1 = (T)f.c1;
    p.c2 = (T)f.c2;
    p.c3 = (T)f.c3;
    const double sl2e = 1.2011224087864498;  // sqrt(log2 e)
    for (int k = 0; k < TQ_MAX_DIM; ++k) {
        double a = f.a[k], b = 0.0;
        switch (f.family) {
            case TQ_FAM_GAUSSIAN_PEAK:
                b = -f.a[k] * f.u[k] * sl2e;
                a = f.a[k] * sl2e;
                break;
            case TQ_FAM_PRODUCT_PEAK: b = (f.a[k] != 0.0) ? 1.0 / (f.a[k] * f.a[k]) : 0.0; break;
            case TQ_FAM_EXP_LINEAR: a = f.a[k] * kLog2e; break;
            case TQ_FAM_OSCILLATORY:
                if (sizeof(T) == 4) a = f.a[k] / (2.0 * M_PI);  // FP32 phase in revolutions
                break;
            default: b = 0.0;
        }
        p.a[k] = (T)a;
        p.u[k] = (T)f.u[k];
        p.b[k] = (T)b;
    }
    if (f.family == TQ_FAM_OSCILLATORY) p.c2 = (T)(sizeof(T) == 4 ? f.c1 : 2.0 * M_PI * f.c1);
    if (f.family == TQ_FAM_EXP_LINEAR) p.c1 = (T)(f.c1 * kLog2e);
    if (f.family == TQ_FAM_DOUBLE_GAUSSIAN) p.c3 = (T)(f.c3 * kLog2e);
    return p;
}

inline int check_integrand(const tq_integrand* f, int dim) {
    if (!f) return fail(TQ_E_NULL, "integrand descriptor is NULL");
    if (f->family < 0 || f->family >= TQ_NUM_FAMILIES) return fail(TQ_E_ARG, "unknown integrand family");
    if (f->dim != dim) return fail(TQ_E_ARG, "integrand dim does not match the grid/sample dim");
    if (dim < 1 || dim > TQ_MAX_DIM) return fail(TQ_E_ARG, "dim out of range for fused kernels (1..16)");
    return TQ_OK;
}

}  // namespace tq

// Expands BODY once per family with FAM bound to the family id constant.
#define TQ_DISPATCH_FAMILY(fam, BODY)                                         \
    switch (fam) {                                                            \
        case TQ_FAM_GAUSSIAN_PEAK: { constexpr int FAM = TQ_FAM_GAUSSIAN_PEAK; BODY; } break;     \
        case TQ_FAM_OSCILLATORY: { constexpr int FAM = TQ_FAM_OSCILLATORY; BODY; } break;         \
        case TQ_FAM_PRODUCT_PEAK: { constexpr int FAM = TQ_FAM_PRODUCT_PEAK; BODY; } break;       \
        case TQ_FAM_EXP_LINEAR: { constexpr int FAM = TQ_FAM_EXP_LINEAR; BODY; } break;           \
        case TQ_FAM_SIN_SUM: { constexpr int FAM = TQ_FAM_SIN_SUM; BODY; } break;                 \
        case TQ_FAM_DOUBLE_GAUSSIAN: { constexpr int FAM = TQ_FAM_DOUBLE_GAUSSIAN; BODY; } break; \
        case TQ_FAM_CONSTANT: { constexpr int FAM = TQ_FAM_CONSTANT; BODY; } break;               \
        default: return tq::fail(TQ_E_ARG, "unknown integrand family");                          \
    }
