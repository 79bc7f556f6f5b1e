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

// Family-specific derived per-axis constants (Params::b, c2) computed in
// double precision before rounding to T.
template <typename T>
inline Params<T> make_params(const tq_integrand& f) {
    Params<T> p{};
    p.c0 = (T)f.c0;
    p.c1 = (T)f.c1;
    p.c2 = (T)f.c2;
    p.c3 = (T)f.c3;
    for (int k = 0; k < TQ_MAX_DIM; ++k) {
        p.a[k] = (T)f.a[k];
        p.u[k] = (T)f.u[k];
        double b = 0.0;
        switch (f.family) {
            case TQ_FAM_GAUSSIAN_PEAK: b = -f.a[k] * f.u[k]; break;
            case TQ_FAM_PRODUCT_PEAK: b = (f.a[k] != 0.0) ? 1.0 / (f.a[k] * f.a[k]) : 0.0; break;
            default: b = 0.0;
        }
        p.b[k] = (T)b;
    }
    if (f.family == TQ_FAM_OSCILLATORY) p.c2 = (T)(2.0 * M_PI * f.c1);
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
