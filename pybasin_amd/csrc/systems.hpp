// Device right-hand sides for the per-lane Dormand-Prince kernel.
//
// Each functor has a compile-time state size S and parameter count NP and
// evaluates dy = f(t, y; p) entirely in registers.  The arithmetic follows
// the reference's torch RHS term by term (same operand order; the compiler
// may contract a multiply-add into one FMA), except Duffing's v', whose three
// terms are regrouped to shorten its dependency chain (same terms, another
// rounding order).
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

// Largest |x| for which the branch-free float sine and cosine below are
// used; the step code checks its arguments against it and recomputes a
// step with the library functions when one exceeds it.
constexpr float kFastTrigMax = 131072.0f;

// sin and cos of a float argument |x| <= kFastTrigMax, branch-free:
// Cody-Waite reduction by pi/2 in three FMA steps (pi/2 split into three
// floats, each product formed exactly inside its FMA; quadrant k = rint(x *
// 2/pi)), then sin and cos of the remainder |r| <= 0.8 by polynomials of
// degree 7 and 8 fitted on [-0.8, 0.8] by Lawson-weighted least squares
// (near-minimax; sin weighted by its relative error) and rounded to float:
// within 7.3e-8 absolute of sin x for |x| <= 2^17 (one unit in the last
// place of values near 1), checked on 4.5e6 sampled floats including the
// neighbours of multiples of pi/2.  About a third of the instructions of
// the library sinf, and no branch: on a wave running alone on its SIMD a
// branch on a fresh comparison costs about as much as ten arithmetic
// instructions.
__device__ __forceinline__ void cw_reduce(float x, float &r, int &q) {
    const float k = rintf(x * 6.36619747e-01f);
    r = fmaf(-k, 1.57079637e+00f, x);
    r = fmaf(-k, -4.37113883e-08f, r);
    r = fmaf(-k, -1.71512451e-15f, r);
    q = (int)k;
}
__device__ __forceinline__ float cw_sin_poly(float r, float r2) {
    const float ps = fmaf(r2, fmaf(r2, -1.95031665e-04f, 8.33207183e-03f), -1.66666538e-01f);
    return fmaf(r2 * r, ps, r);
}
__device__ __forceinline__ float cw_cos_poly(float r2) {
    const float pc = fmaf(r2, fmaf(r2, fmaf(r2, 2.43752329e-05f, -1.38866028e-03f), 4.16666195e-02f),
                          -5.0e-01f);
    return fmaf(r2, pc, 1.0f);
}
// sin alone of a float argument |x| <= kFastTrigMax, branch-free, in 13
// instructions (the quadrant form above takes 20): Cody-Waite reduction by
// pi in three FMA steps, k = the integer nearest x / pi formed in the low
// mantissa bits of t = x / pi + 1.5 * 2^23 (|x / pi| < 2^22), so the
// parity of k is the last bit of t; sin x = (-1)^k sin r with |r| <= pi/2,
// the sign applied to r (sin is odd); sin r = r + r^3 P(r^2) with P of
// degree 3 fitted on [0, pi/2] by Lawson-weighted least squares on the
// relative error and rounded to float: within 1.25e-7 of sin x (1.74 units
// in the last place of the result) for |x| <= 2^17 on 7.7e6 sampled floats,
// the neighbours of multiples of pi/2 included.
__device__ __forceinline__ float sin_fast(float x) {
    const float t = fmaf(x, 3.18309873e-01f, 1.25829120e+07f);
    const float k = t - 1.25829120e+07f;
    float r = fmaf(-k, 3.14159274e+00f, x);
    r = fmaf(-k, -8.74227766e-08f, r);
    r = fmaf(-k, -3.43024902e-15f, r);
    r = __int_as_float(__float_as_int(r) ^ (__float_as_int(t) << 31));
    const float r2 = r * r;
    const float p = fmaf(fmaf(fmaf(2.60559003e-06f, r2, -1.98095210e-04f), r2, 8.33306555e-03f), r2,
                         -1.66666597e-01f);
    return fmaf(r2 * r, p, r);
}
__device__ __forceinline__ void sincos_fast(float x, float &s, float &c) {
    float r;
    int q;
    cw_reduce(x, r, q);
    const float r2 = r * r;
    const float sv = cw_sin_poly(r, r2), cv = cw_cos_poly(r2);
    // quadrant q: sin = sv, cv, -sv, -cv; cos = cv, -sv, -cv, sv
    const float a = (q & 1) ? cv : sv, b = (q & 1) ? sv : cv;
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}
// sin of any float: the branch-free form inside its range, the library
// sinf beyond
__device__ __forceinline__ float sin_cw(float x) {
    if (!(fabsf(x) <= kFastTrigMax)) return sinf(x);
    return sin_fast(x);
}

template <class T> __device__ __forceinline__ T dsin(T x);
template <class T> __device__ __forceinline__ T dcos(T x);
template <> __device__ __forceinline__ float dsin<float>(float x) { return sin_cw(x); }
template <> __device__ __forceinline__ double dsin<double>(double x) { return sin(x); }
template <> __device__ __forceinline__ float dcos<float>(float x) { return cosf(x); }
template <> __device__ __forceinline__ double dcos<double>(double x) { return cos(x); }

// sin and cos of an argument |d| <= 1.  float: s = d + d^3 P(d^2) and
// c = 1 - d^2 Q(d^2) with P of degree 2 and Q of degree 3, coefficients fitted
// on [-1, 1] by Lawson-weighted least squares (near-minimax; sin weighted by
// its relative error), rounded to float and then moved by a few units in the
// last place to minimise the error of this float evaluation: within 1.3
// ulp of sin and cos on 8e7 sampled floats of [-1, 1].  double: the
// Taylor series to d^19 / 19! and d^18 / 18! (truncation below 1e-17).
template <class T> __device__ __forceinline__ void sincos_unit(T d, T &s, T &c);
template <> __device__ __forceinline__ void sincos_unit<float>(float d, float &s, float &c) {
    const float d2 = d * d;
    s = d + d * d2 * (-1.6666615e-01f + d2 * (8.330241e-03f + d2 * -1.9313331e-04f));
    c = 1.0f - d2 * (4.9999997e-01f + d2 * (-4.1666474e-02f + d2 * (1.3883339e-03f + d2 * -2.4138277e-05f)));
}
template <> __device__ __forceinline__ void sincos_unit<double>(double d, double &s, double &c) {
    const double d2 = d * d;
    double ps = 1.0 / 121645100408832000.0;  // 1/19!
    ps = -1.0 / 355687428096000.0 + d2 * ps;
    ps = 1.0 / 1307674368000.0 + d2 * ps;
    ps = -1.0 / 6227020800.0 + d2 * ps;
    ps = 1.0 / 39916800.0 + d2 * ps;
    ps = -1.0 / 362880.0 + d2 * ps;
    ps = 1.0 / 5040.0 + d2 * ps;
    ps = -1.0 / 120.0 + d2 * ps;
    ps = 1.0 / 6.0 + d2 * ps;
    s = d - d * d2 * ps;
    double pc = 1.0 / 6402373705728000.0;  // 1/18!
    pc = -1.0 / 20922789888000.0 + d2 * pc;
    pc = 1.0 / 87178291200.0 + d2 * pc;
    pc = -1.0 / 479001600.0 + d2 * pc;
    pc = 1.0 / 3628800.0 + d2 * pc;
    pc = -1.0 / 40320.0 + d2 * pc;
    pc = 1.0 / 720.0 + d2 * pc;
    pc = -1.0 / 24.0 + d2 * pc;
    pc = 0.5 + d2 * pc;
    c = 1.0 - d2 * pc;
}
template <class T> __device__ __forceinline__ void dsincos(T x, T &s, T &c);
template <> __device__ __forceinline__ void dsincos<float>(float x, float &s, float &c) { sincosf(x, &s, &c); }
template <> __device__ __forceinline__ void dsincos<double>(double x, double &s, double &c) { sincos(x, &s, &c); }

// Per-step time terms.  A system whose right-hand side depends on t gets the
// step's stage times (t0 + c_i h for c = 1/5, 3/10, 4/5, 8/9, in the state
// dtype), the sine and cosine of t0 carried from the previous step, the end
// time tB = (T)(t + dt) and the c = 1 stage time t1 (the previous
// representable value of tB); it returns the sine and cosine of tB (carried
// to the next step after an accept) and fills one auxiliary value per stage.
// The others ignore them.  time_terms is the form for any argument;
// time_terms_fast (float) takes the end time's sine and cosine from the
// branch-free form and returns false when |tB| is outside its range (the
// step code then integrates the lane again with time_terms).
struct NoTimeTerms {
    static constexpr bool kTime = false;
    static constexpr bool kFastRhs = false;
    template <class T>
    __device__ __forceinline__ static void time_terms(int32_t, T, const T (&)[4], T, T, T, T, T &sB,
                                                      T &cB, T (&aux)[5]) {
        sB = cB = T(0);
#pragma unroll
        for (int i = 0; i < 5; i++) aux[i] = T(0);
    }
    template <class T>
    __device__ __forceinline__ static bool time_terms_fast(T, const T (&)[4], T, T, T, T, T &sB, T &cB,
                                                           T (&aux)[5]) {
        sB = cB = T(0);
#pragma unroll
        for (int i = 0; i < 5; i++) aux[i] = T(0);
        return true;
    }
};

// tests/unit/solvers/test_solver.py:16-18 -- dy/dt = p0 * y
template <int S_> struct Decay : NoTimeTerms {
    static constexpr int S = S_, NP = 1;
    static constexpr int FLOPS = S_;
    template <class T>
    __device__ __forceinline__ static void f(const T (&p)[NP], T, const T (&y)[S], T (&dy)[S]) {
#pragma unroll
        for (int i = 0; i < S; i++) dy[i] = p[0] * y[i];
    }
};

// y' = M y with M = [[p0, p1], [p2, p3]] (harmonic / damped oscillators in tests)
struct Linear2 : NoTimeTerms {
    static constexpr int S = 2, NP = 4;
    static constexpr int FLOPS = 6;
    template <class T>
    __device__ __forceinline__ static void f(const T (&p)[NP], T, const T (&y)[S], T (&dy)[S]) {
        dy[0] = p[0] * y[0] + p[1] * y[1];
        dy[1] = p[2] * y[0] + p[3] * y[1];
    }
};

// case_studies/pendulum/pendulum_ode.py:32-40
//   theta' = omega;  omega' = -alpha*omega + T - K*sin(theta);  p = [alpha, T, K]
struct Pendulum : NoTimeTerms {
    static constexpr int S = 2, NP = 3;
    static constexpr int FLOPS = 5;
    static constexpr bool kFastRhs = true;
    template <class T>
    __device__ __forceinline__ static void f(const T (&p)[NP], T, const T (&y)[S], T (&dy)[S]) {
        dy[0] = y[1];
        dy[1] = -p[0] * y[1] + p[1] - p[2] * dsin<T>(y[0]);
    }
    // float, branch-free: the same terms with the sine of |theta| <=
    // kFastTrigMax; amax collects |theta| for the step's range check
    __device__ __forceinline__ static void f_fast(const float (&p)[NP], const float (&y)[S],
                                                  float (&dy)[S], float &amax) {
        dy[0] = y[1];
        dy[1] = -p[0] * y[1] + p[1] - p[2] * sin_fast(y[0]);
        amax = fmaxf(amax, fabsf(y[0]));
    }
};

// case_studies/duffing_oscillator/duffing_ode.py:35-42
//   x' = v;  v' = -delta*v - k3*x^3 + A*cos(t);  p = [delta, k3, A]
// Within a step the stage times t_i (values of the state dtype) differ from
// t_0 by d_i = t_i - t_0, which the subtraction forms exactly once
// t_i <= 2 t_0 (Sterbenz; every step after the first few); cos(t_i) is then
// cos(t_0) cos(d_i) - sin(t_0) sin(d_i): one range-reduced sincos per step
// instead of one cos per stage, for the same value of cos(t_i).
struct Duffing {
    static constexpr int S = 2, NP = 3;
    static constexpr int FLOPS = 8;
    static constexpr bool kTime = true;
    template <class T>
    __device__ __forceinline__ static void f(const T (&p)[NP], T t, const T (&y)[S], T (&dy)[S]) {
        f_aux(p, dcos<T>(t), y, dy);
    }
    template <class T>
    __device__ __forceinline__ static void f_aux(const T (&p)[NP], T cos_t, const T (&y)[S], T (&dy)[S]) {
        // -delta v - k3 x^3 + A cos t grouped so that three dependent
        // operations follow the stage input (k3 x, x^2 and delta v at once)
        dy[0] = y[1];
        dy[1] = fma(-(p[1] * y[0]), y[0] * y[0], fma(p[2], cos_t, -(p[0] * y[1])));
    }
    static constexpr bool kFastRhs = false;
    template <class T>
    __device__ __forceinline__ static void time_terms(int32_t steps, T t0, const T (&ts)[4], T s0,
                                                      T c0, T tB, T t1, T &sB, T &cB, T (&aux)[5]) {
        // sin / cos of the end time rotated from the step's start by
        // d = tB - t0 (exact, as the stage offsets), with a fresh
        // range-reduced sincos every 8th attempted step of the lane, so at
        // most 7 rotations accumulate into (s0, c0)
        const T dB = tB - t0;
        if ((steps & 7) == 0 || !(fabs(dB) <= T(1))) {
            dsincos<T>(tB, sB, cB);
        } else {
            T sd, cd;
            sincos_unit<T>(dB, sd, cd);
            sB = fma(s0, cd, c0 * sd);
            cB = fma(c0, cd, -(s0 * sd));
        }
        // cos(t0 + d_i) = cos t0 cos d_i - sin t0 sin d_i with d_i = t_i - t0
        // exact (Sterbenz) and sin/cos of d_i by one short polynomial.  The
        // c = 8/9 stage has the largest offset, so one test per lane decides;
        // nothing depends on the other lanes of the wave, so a trajectory's
        // rounding does not depend on which ICs share its wave.  The c = 1
        // stage time t1 = tB - e (one unit in the last place below tB) takes
        // cos t1 from (sin tB, cos tB) by the series of cos e and sin e while
        // e <= 2^-6; a larger e (|t| beyond 2^17 in float) takes the direct
        // cosine with the other stages.
        const T dmax = fabs(ts[3] - t0), e = tB - t1;
        if (dmax <= T(1) && e <= T(0.015625)) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                T sd, cd;
                sincos_unit<T>(ts[i] - t0, sd, cd);
                aux[i] = c0 * cd - s0 * sd;
            }
            aux[4] = cos_prev(sB, cB, e);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) aux[i] = dcos<T>(ts[i]);
            aux[4] = dcos<T>(t1);
        }
    }
    // cos(tB - e) from sin tB, cos tB for 0 <= e <= 2^-6: cos e = 1 - e^2/2 and
    // sin e = e - e^3/6 (the next terms are below 3e-9 relative); for e below
    // 1e-8 both round to 1 and e exactly
    template <class T>
    __device__ __forceinline__ static T cos_prev(T sB, T cB, T e) {
        const T e2 = e * e;
        return fma(sB, fma(e2 * e, T(-1.0 / 6.0), e), cB * fma(e2, T(-0.5), T(1)));
    }
    // float: a fresh branch-free sincos of tB every step (no rotation carried
    // from step to step, so no refresh branch), the stage terms as in
    // time_terms; false when |tB| > kFastTrigMax
    __device__ __forceinline__ static bool time_terms_fast(float t0, const float (&ts)[4], float s0,
                                                           float c0, float tB, float t1, float &sB,
                                                           float &cB, float (&aux)[5]) {
        sincos_fast(tB, sB, cB);
        const float dmax = fabsf(ts[3] - t0);
        if (dmax <= 1.0f) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float sd, cd;
                sincos_unit<float>(ts[i] - t0, sd, cd);
                aux[i] = c0 * cd - s0 * sd;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) aux[i] = cosf(ts[i]);
        }
        // |tB| <= 2^17 bounds e = tB - t1 by 2^-6
        aux[4] = cos_prev(sB, cB, tB - t1);
        return fabsf(tB) <= kFastTrigMax;
    }
};

// right-hand side with the per-step time terms (time-independent systems
// ignore them); kFast (float) takes a system's branch-free form where it
// has one, which adds its argument magnitude to amax
template <class Sys, class T, bool kFast>
__device__ __forceinline__ void rhs_aux(const T (&p)[Sys::NP], T aux, const T (&y)[Sys::S],
                                        T (&dy)[Sys::S], T &amax) {
    if constexpr (Sys::kTime)
        Sys::f_aux(p, aux, y, dy);
    else if constexpr (kFast && Sys::kFastRhs)
        Sys::f_fast(p, y, dy, amax);
    else
        Sys::f(p, T(0), y, dy);
}

// case_studies/lorenz/lorenz_ode.py:40-64, including the freeze mask:
// the derivative is multiplied by (max|y| < 200).
struct Lorenz : NoTimeTerms {
    static constexpr int S = 3, NP = 3;
    static constexpr int FLOPS = 9;
    template <class T>
    __device__ __forceinline__ static void f(const T (&p)[NP], T, const T (&y)[S], T (&dy)[S]) {
        T m = fmax(fmax(fabs(y[0]), fabs(y[1])), fabs(y[2]));
        T mask = (m < T(200)) ? T(1) : T(0);
        dy[0] = (p[0] * (y[1] - y[0])) * mask;
        dy[1] = (p[1] * y[0] - y[0] * y[2] - y[1]) * mask;
        dy[2] = (y[0] * y[1] - p[2] * y[2]) * mask;
    }
};

}  // namespace bh
