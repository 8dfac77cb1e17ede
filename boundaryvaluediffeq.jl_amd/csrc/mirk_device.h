// mirk_device.h -- device building blocks of the MIRK collocation path (gfx950).
//
//  * MIRK4 / MIRK6 tableaus as compile-time constants
//    (MIRK/mirk_tableaus.jl:62-87, 120-152; paths relative to the reference
//    root, MIRK/ = lib/BoundaryValueDiffEqMIRK/src/)
//  * continuous-extension weights w(tau), w'(tau) (MIRK/interpolation.jl:481-575)
//  * forward-mode dual numbers, used to get the per-interval Jacobian blocks
//    in the same pass as the residual (the reference differentiates the whole
//    residual with sparse ForwardDiff, MIRK/mirk.jl:697-862, 927-1002)
//  * the right-hand sides of the supported problem families, written once and
//    instantiated for double and for dual numbers.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/bvp_mirk.h"

namespace bvpdev {

// ----------------------------------------------------------------- tableaus
template <int METHOD> struct Tab;

template <> struct Tab<4> {
    static constexpr int s = 3, sstar = 4, order = 4;
    __host__ __device__ static constexpr double c(int r) { return r == 0 ? 0.0 : r == 1 ? 1.0 : 0.5; }
    __host__ __device__ static constexpr double v(int r) { return r == 0 ? 0.0 : r == 1 ? 1.0 : 0.5; }
    __host__ __device__ static constexpr double b(int r) { return r == 2 ? 2.0 / 3 : 1.0 / 6; }
    __host__ __device__ static constexpr double x(int r, int j) {
        return (r == 2 && j == 0) ? 1.0 / 8 : (r == 2 && j == 1) ? -1.0 / 8 : 0.0;
    }
    // interpolation stages (one): c* = 3/4, v* = 27/32, x* = [3/64, -9/64, 0 | 0]
    __host__ __device__ static constexpr double cs(int) { return 3.0 / 4; }
    __host__ __device__ static constexpr double vs(int) { return 27.0 / 32; }
    __host__ __device__ static constexpr double xs(int, int j) {
        return j == 0 ? 3.0 / 64 : j == 1 ? -9.0 / 64 : 0.0;
    }
    __host__ __device__ static constexpr double tau() { return 0.226; }
};

template <> struct Tab<6> {
    static constexpr int s = 5, sstar = 9, order = 6;
    __host__ __device__ static constexpr double c(int r) {
        return r == 0 ? 0.0 : r == 1 ? 1.0 : r == 2 ? 0.25 : r == 3 ? 0.75 : 0.5;
    }
    __host__ __device__ static constexpr double v(int r) {
        return r == 0 ? 0.0 : r == 1 ? 1.0 : r == 2 ? 5.0 / 32 : r == 3 ? 27.0 / 32 : 0.5;
    }
    __host__ __device__ static constexpr double b(int r) {
        return r <= 1 ? 7.0 / 90 : r <= 3 ? 16.0 / 45 : 2.0 / 15;
    }
    __host__ __device__ static constexpr double x(int r, int j) {
        return r == 2 ? (j == 0 ? 9.0 / 64 : j == 1 ? -3.0 / 64 : 0.0)
             : r == 3 ? (j == 0 ? 3.0 / 64 : j == 1 ? -9.0 / 64 : 0.0)
             : r == 4 ? (j == 0 ? -5.0 / 24 : j == 1 ? 5.0 / 24 : j == 2 ? 2.0 / 3 : j == 3 ? -2.0 / 3 : 0.0)
             : 0.0;
    }
    __host__ __device__ static constexpr double cs(int r) {
        return r == 0 ? 7.0 / 16 : r == 1 ? 3.0 / 8 : r == 2 ? 9.0 / 16 : 1.0 / 8;
    }
    __host__ __device__ static constexpr double vs(int r) { return cs(r); }
    __host__ __device__ static constexpr double xs(int r, int j) {
        return r == 0 ? (j == 0 ? 1547.0 / 32768 : j == 1 ? -1225.0 / 32768 : j == 2 ? 749.0 / 4096
                         : j == 3 ? -287.0 / 2048 : j == 4 ? -861.0 / 16384 : 0.0)
             : r == 1 ? (j == 0 ? 83.0 / 1536 : j == 1 ? -13.0 / 384 : j == 2 ? 283.0 / 1536
                         : j == 3 ? -167.0 / 1536 : j == 4 ? -49.0 / 512 : 0.0)
             : r == 2 ? (j == 0 ? 1225.0 / 32768 : j == 1 ? -1547.0 / 32768 : j == 2 ? 287.0 / 2048
                         : j == 3 ? -749.0 / 4096 : j == 4 ? 861.0 / 16384 : 0.0)
             : (j == 0 ? 233.0 / 3456 : j == 1 ? -19.0 / 1152 : j == 5 ? -5.0 / 72 : j == 6 ? 7.0 / 72
                : j == 7 ? -17.0 / 216 : 0.0);
    }
    __host__ __device__ static constexpr double tau() { return 0.7156; }
};

// Continuous-extension weights (MIRK/interpolation.jl:481-499, 528-575).
template <int METHOD>
__host__ __device__ inline void interp_weights(double tau, double *w, double *wp);

template <>
__host__ __device__ inline void interp_weights<4>(double tau, double *w, double *wp) {
    const double t2 = tau * tau, tm1 = tau - 1.0, t4m3 = tau * 4.0 - 3.0, t2m1 = tau * 2.0 - 1.0;
    w[0] = -tau * (2.0 * tau - 3.0) * (2.0 * t2 - 3.0 * tau + 2.0) / 6.0;
    w[1] = t2 * (12.0 * t2 - 20.0 * tau + 9.0) / 6.0;
    w[2] = 2.0 * t2 * (6.0 * t2 - 14.0 * tau + 9.0) / 3.0;
    w[3] = -16.0 * t2 * tm1 * tm1 / 3.0;
    if (wp) {
        wp[0] = -tm1 * t4m3 * t2m1 / 3.0;
        wp[1] = tau * t2m1 * t4m3;
        wp[2] = 4.0 * tau * t4m3 * tm1;
        wp[3] = -32.0 * tau * t2m1 * tm1 / 3.0;
    }
}

template <>
__host__ __device__ inline void interp_weights<6>(double tau, double *w, double *wp) {
    const double t = tau, t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t, t6 = t5 * t;
    w[0] = t - 28607.0 / 7434 * t2 - 166210.0 / 33453 * t3 + 334780.0 / 11151 * t4 -
           1911296.0 / 55755 * t5 + 406528.0 / 33453 * t6;
    w[1] = 777.0 / 590 * t2 - 2534158.0 / 234171 * t3 + 2088580.0 / 78057 * t4 -
           10479104.0 / 390285 * t5 + 11328512.0 / 1170855 * t6;
    w[2] = -1008.0 / 59 * t2 + 222176.0 / 1593 * t3 - 180032.0 / 531 * t4 + 876544.0 / 2655 * t5 -
           180224.0 / 1593 * t6;
    w[3] = w[2];  // stages 3 and 4 share one polynomial in the reference's table
    w[4] = -378.0 / 59 * t2 + 27772.0 / 531 * t3 - 22504.0 / 177 * t4 + 109568.0 / 885 * t5 -
           22528.0 / 531 * t6;
    w[5] = -95232.0 / 413 * t2 + 62384128.0 / 33453 * t3 - 49429504.0 / 11151 * t4 +
           46759936.0 / 11151 * t5 - 46661632.0 / 33453 * t6;
    w[6] = 896.0 / 5 * t2 - 4352.0 / 3 * t3 + 3456.0 * t4 - 16384.0 / 5 * t5 + 16384.0 / 15 * t6;
    w[7] = 50176.0 / 531 * t2 - 179554304.0 / 234171 * t3 + 143363072.0 / 78057 * t4 -
           136675328.0 / 78057 * t5 + 137363456.0 / 234171 * t6;
    w[8] = 16384.0 / 441 * t3 - 16384.0 / 147 * t4 + 16384.0 / 147 * t5 - 16384.0 / 441 * t6;
    if (wp) {
        wp[0] = 1 - 28607.0 / 3717 * t - 166210.0 / 11151 * t2 + 1339120.0 / 11151 * t3 -
                1911296.0 / 11151 * t4 + 813056.0 / 11151 * t5;
        wp[1] = 777.0 / 295 * t - 2534158.0 / 78057 * t2 + 8354320.0 / 78057 * t3 -
                10479104.0 / 78057 * t4 + 22657024.0 / 390285 * t5;
        wp[2] = -2016.0 / 59 * t + 222176.0 / 531 * t2 - 720128.0 / 531 * t3 + 876544.0 / 531 * t4 -
                360448.0 / 531 * t5;
        wp[3] = wp[2];
        wp[4] = -756.0 / 59 * t + 27772.0 / 177 * t2 - 90016.0 / 177 * t3 + 109568.0 / 177 * t4 -
                45056.0 / 177 * t5;
        wp[5] = -190464.0 / 413 * t + 62384128.0 / 11151 * t2 - 197718016.0 / 11151 * t3 +
                233799680.0 / 11151 * t4 - 93323264.0 / 11151 * t5;
        wp[6] = 1792.0 / 5 * t - 4352.0 * t2 + 13824.0 * t3 - 16384.0 * t4 + 32768.0 / 5 * t5;
        wp[7] = 100352.0 / 531 * t - 179554304.0 / 78057 * t2 + 573452288.0 / 78057 * t3 -
                683376640.0 / 78057 * t4 + 274726912.0 / 78057 * t5;
        wp[8] = 16384.0 / 147 * t2 - 65536.0 / 147 * t3 + 81920.0 / 147 * t4 - 32768.0 / 147 * t5;
    }
}

// ----------------------------------------------------------------- dual numbers
// v + sum_k d[k] eps_k, first order in each eps_k.
template <int ND> struct Dual {
    double v;
    double d[ND];
};

template <int ND>
__device__ inline Dual<ND> dconst(double a) {
    Dual<ND> r;
    r.v = a;
#pragma unroll
    for (int k = 0; k < ND; ++k) r.d[k] = 0.0;
    return r;
}
template <int ND>
__device__ inline Dual<ND> operator+(const Dual<ND> &a, const Dual<ND> &b) {
    Dual<ND> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < ND; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int ND>
__device__ inline Dual<ND> operator-(const Dual<ND> &a, const Dual<ND> &b) {
    Dual<ND> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < ND; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int ND>
__device__ inline Dual<ND> operator-(const Dual<ND> &a) {
    Dual<ND> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < ND; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int ND>
__device__ inline Dual<ND> operator*(const Dual<ND> &a, const Dual<ND> &b) {
    Dual<ND> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < ND; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int ND>
__device__ inline Dual<ND> operator*(double a, const Dual<ND> &b) {
    Dual<ND> r;
    r.v = a * b.v;
#pragma unroll
    for (int k = 0; k < ND; ++k) r.d[k] = a * b.d[k];
    return r;
}
template <int ND>
__device__ inline Dual<ND> operator/(const Dual<ND> &a, double b) {
    Dual<ND> r;
    r.v = a.v / b;
#pragma unroll
    for (int k = 0; k < ND; ++k) r.d[k] = a.d[k] / b;
    return r;
}
template <int ND>
__device__ inline Dual<ND> operator+(const Dual<ND> &a, double b) {
    Dual<ND> r = a;
    r.v += b;
    return r;
}
template <int ND>
__device__ inline Dual<ND> scale_by(const Dual<ND> &a, double dv) {  // chain rule: g(a) with g' = dv
    Dual<ND> r;
#pragma unroll
    for (int k = 0; k < ND; ++k) r.d[k] = dv * a.d[k];
    return r;
}
template <int ND>
__device__ inline Dual<ND> dexp(const Dual<ND> &a) {
    const double e = exp(a.v);
    Dual<ND> r = scale_by(a, e);
    r.v = e;
    return r;
}
template <int ND>
__device__ inline Dual<ND> dsin(const Dual<ND> &a) {
    double s, c;
    sincos(a.v, &s, &c);
    Dual<ND> r = scale_by(a, c);
    r.v = s;
    return r;
}
template <int ND>
__device__ inline Dual<ND> dtanh(const Dual<ND> &a) {
    const double t = tanh(a.v);
    Dual<ND> r = scale_by(a, 1.0 - t * t);
    r.v = t;
    return r;
}

// scalar helpers so that one template body serves double and Dual
__device__ inline double dexp(double a) { return exp(a); }
__device__ inline double dsin(double a) { return sin(a); }
__device__ inline double dtanh(double a) { return tanh(a); }
template <typename T> __device__ inline T make_const(double a);
template <> __device__ inline double make_const<double>(double a) { return a; }
template <typename T> struct ConstOf { __device__ static T mk(double a) { return dconst<sizeof(T) / sizeof(double) - 1>(a); } };
template <> struct ConstOf<double> { __device__ static double mk(double a) { return a; } };
__device__ inline double value_of(double a) { return a; }
template <int ND> __device__ inline double value_of(const Dual<ND> &a) { return a.v; }

// ----------------------------------------------------------------- families
// Parameters are preloaded into registers: p[0..MAXP-1].
template <int M> struct Params {
    static constexpr int MAXP = 1 + M * M;
    double p[MAXP];
};

// f(y, p, t) for the supported families (include/bvp_mirk.h, BVP_FAMILY_*).
template <int M, typename T>
__device__ inline void f_eval(int fam, const Params<M> &P, double t, const T *y, T *out) {
    if constexpr (M == 2) {
        if (fam == BVP_FAMILY_BRATU) {
            out[0] = y[1];
            out[1] = -P.p[0] * dexp(y[0]);
            return;
        }
        if (fam == BVP_FAMILY_PENDULUM) {
            out[0] = y[1];
            out[1] = -P.p[0] * dsin(y[0]);
            return;
        }
        if (fam == BVP_FAMILY_SINGPERT) {
            out[0] = y[1];
            out[1] = (y[0] - t * y[1]) / P.p[0];
            return;
        }
        if (fam == BVP_FAMILY_TURNING) {
            const double pi = 3.14159265358979323846;
            double sp, cp;
            sincos(pi * t, &sp, &cp);
            out[0] = y[1];
            out[1] = (-t / P.p[0]) * y[1] + (-pi * pi * cp - pi * t / P.p[0] * sp);
            return;
        }
    }
    if constexpr (M == 1) {
        if (fam == BVP_FAMILY_SCALAR_TU) {
            out[0] = t * y[0];
            return;
        }
    }
    if constexpr (M == 6) {
        if (fam == BVP_FAMILY_SWIRL) {
            out[0] = y[1];
            out[1] = (y[0] * y[3] - y[2] * y[1]) / P.p[0];
            out[2] = y[3];
            out[3] = y[4];
            out[4] = y[5];
            out[5] = (-y[2] * y[5] - y[0] * y[1]) / P.p[0];
            return;
        }
    }
    // BVP_FAMILY_LINEAR_TANH: f = A y + c tanh(y)
    {
        const double c = P.p[0];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            T acc = P.p[1 + i * M] * y[0];
#pragma unroll
            for (int j = 1; j < M; ++j) acc = acc + P.p[1 + i * M + j] * y[j];
            out[i] = acc + c * dtanh(y[i]);
        }
    }
}

template <int M>
__host__ __device__ inline bool family_supported(int fam) {
    if (fam == BVP_FAMILY_LINEAR_TANH) return true;
    if (M == 2) return fam == BVP_FAMILY_BRATU || fam == BVP_FAMILY_PENDULUM || fam == BVP_FAMILY_SINGPERT ||
                       fam == BVP_FAMILY_TURNING;
    if (M == 1) return fam == BVP_FAMILY_SCALAR_TU;
    if (M == 6) return fam == BVP_FAMILY_SWIRL;
    return false;
}

}  // namespace bvpdev
