// Edit-buffer arithmetic: the edit value of main_3d.cu:239-244 and the epsilon-checked update of
// applyDeltaBuffer1 main_3d.cu:1122-1164.
//
// double fields: the reference's arithmetic, operation for operation.
// float fields : values are stored as float32; the edit value is evaluated in double from the
//   float32 inputs, the candidate g + delta is rounded to float32 and the epsilon check is done
//   on the rounded value (so |g' - f| <= eps holds for the stored value), and the snap value is
//   the smallest float32 x with f - x <= eps (the float32 form of f - eps).
// Both CPU and GPU builds use -ffp-contract=off, so every operation rounds once.
#pragma once
#include <cstdint>
#include <cstring>
#include "stencil.h"

namespace dmtz {

DMTZ_HD float next_up_f32(float x) {
    if (x == 0.0f) return 1.40129846e-45f;  // smallest positive subnormal
    uint32_t b;
    memcpy(&b, &x, 4);
    b = (x > 0.0f) ? b + 1u : b - 1u;
    float r;
    memcpy(&r, &b, 4);
    return r;
}

DMTZ_HD double next_up_f64(double x) {
    if (x == 0.0) return 4.9406564584124654e-324;  // smallest positive subnormal
    uint64_t b;
    memcpy(&b, &x, 8);
    b = (x > 0.0) ? b + 1u : b - 1u;
    double r;
    memcpy(&r, &b, 8);
    return r;
}

// The next representable value above x (x finite).
DMTZ_HD float next_up(float x) { return next_up_f32(x); }
DMTZ_HD double next_up(double x) { return next_up_f64(x); }

template <class T>
struct EditArith;

template <>
struct EditArith<double> {
    DMTZ_HD static double delta0(double f, double g, double eps) { return (f - eps - g) / 2.0; }
    DMTZ_HD static bool step(double f, double& g, double d, double eps) {
        double c = g + d;
        double r = f - c;
        if ((r < 0 ? -r : r) <= eps) {
            g = c;
            return true;
        }
        return false;
    }
    DMTZ_HD static double snap(double f, double eps) { return f - eps; }
};

template <>
struct EditArith<float> {
    DMTZ_HD static double delta0(float f, float g, double eps) { return ((double)f - eps - (double)g) / 2.0; }
    DMTZ_HD static bool step(float f, float& g, double d, double eps) {
        float c = (float)((double)g + d);
        double r = (double)f - (double)c;
        if ((r < 0 ? -r : r) <= eps) {
            g = c;
            return true;
        }
        return false;
    }
    DMTZ_HD static float snap(float f, double eps) {
        float x = (float)((double)f - eps);
        if ((double)f - (double)x > eps) x = next_up_f32(x);
        return x;
    }
};

// 2^-k * eps for the counter-driven edit type (main_3d.cu:241): exact scaling by a power of two.
DMTZ_HD double halving_delta(double eps, int k) {
    double s = 1.0;
    for (int i = 0; i < k; i++) s *= 0.5;
    return -(s * eps);
}

DMTZ_HD double fabs_d(double x) { return x < 0 ? -x : x; }

}  // namespace dmtz
