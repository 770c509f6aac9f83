// dopri5_lane.hip -- one trajectory per GPU lane, adaptive Dormand-Prince 5(4).
//
// Replaces TorchDiffEqSolver._integrate (src/pybasin/solvers/torchdiffeq_solver.py:97-134)
// for systems with a device functor.  Per lane, the kernel applies the step
// rules of torchdiffeq 0.2.5's dopri5 (the solver the reference calls), with
// the error norm taken over the lane's own S components instead of the whole
// batch tensor:
//   * tableau, Shampine error weights and DPS_C_MID midpoint as restated in
//     src/zigode/src/solver/dopri5_generic.zig:16-56;
//   * stage arithmetic in the state dtype T (stage input y + dt sum_j a_ij k_j;
//     torchdiffeq forms sum_j (a_ij dt) k_j, which differs in rounding only);
//     time-like values (t, dt, save times) in double; stage times cast to T,
//     and the c = 1 stages evaluated at the previous representable value of
//     (T)(t + dt) (torchdiffeq's Perturb.PREV);
//   * error ratio RMS(err / (atol + rtol * max(|y0|, |y1|))) <= 1 accepts;
//     next dt = dt * clamp(0.9 * ratio^(-1/5), dfac, 10) with dfac = 1 on
//     accept and 0.2 on reject (dopri5_generic.zig:228-235);
//   * initial dt by Hairer-Norsett-Wanner II.4 (dopri5_generic.zig:270-321);
//   * no clamp at the last save time: the lane steps past it and every save
//     time in (t0, t1] of an accepted step is evaluated from the 4th-order
//     quartic through y0, y1, f0, f1 and the DPS_C_MID midpoint, in power
//     form with x = (s - t0) / (t1 - t0) formed in double.
//
// Save values are consumed in registers: written to HBM only when a
// trajectory buffer is given, and folded into streaming statistics
// (shifted sums in double, min, max) over the feature window
// [steady_idx, n_save) -- the fused form of filter_time + mean / std / min /
// max (src/pybasin/feature_extractors/feature_extractor.py:102-130,
// src/pybasin/ts_torch/calculators/torch_features_statistical.py:39-99).
//
// Orbit peaks (optional, kPeaks): strict local maxima with support 1 of each
// state over the same window, i.e. the find_peak_mask / extract_peak_values
// pair (src/pybasin/ts_torch/calculators/torch_features_pattern.py:64-150)
// behind extract_orbit_data (src/pybasin/utils.py:254-303), found from the
// last two saved values as each new one arrives.
//
// Work distribution: a resident set of lanes; when a lane's trajectory ends
// (all saves emitted, divergence stop, or failure) it takes the next index
// from a device counter, so uneven lengths do not idle the lane.  Finished
// lanes of a wave refill together once a quorum of them waits (flags bits
// 8-14), which trades idle lanes for lanes that stay in step.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/basinhip.h"
#include "systems.hpp"

namespace bh {

// refill quorum when the flags leave it at 0 (chosen by measurement on the
// case-study workloads)
constexpr int kDefaultRefillQuorum = 64;

struct LaneArgs {
    const void *y0;
    const double *params;
    int64_t p_stride;
    int64_t n;
    const double *save_t;
    int32_t n_save;
    int32_t steady_idx;
    double t0;
    double rtol, atol;
    int64_t max_steps;
    uint32_t flags;
    double stop_thr;
    void *traj;
    double *feats;
    void *fin;
    int32_t *status;
    int64_t *nsteps;
    unsigned long long *counter;
    // orbit peaks over the feature window (kernels instantiated with kPeaks)
    void *peak_vals;       // (peak_cap, n, S), values of strict local maxima in time order
    int32_t *peak_counts;  // (n, S), number of maxima (may exceed peak_cap)
    int32_t peak_cap;
    // two-phase launches: step budget of the first phase and the compact
    // buffer of suspended lane states with their trajectory indices
    int32_t phase_iters;
    void *state;
    int64_t *state_ic;
};

// Dormand-Prince 5(4) coefficients (Dormand & Prince 1980; Shampine's error
// weights c_sol - c_hat; DPS_C_MID midpoint weights).
__device__ constexpr double A21 = 1.0 / 5.0;
__device__ constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
__device__ constexpr double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
__device__ constexpr double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0,
                            A53 = 64448.0 / 6561.0, A54 = -212.0 / 729.0;
__device__ constexpr double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0,
                            A64 = 49.0 / 176.0, A65 = -5103.0 / 18656.0;
__device__ constexpr double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0,
                            B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
__device__ constexpr double C2 = 1.0 / 5.0, C3 = 3.0 / 10.0, C4 = 4.0 / 5.0, C5 = 8.0 / 9.0;
__device__ constexpr double E1 = 35.0 / 384.0 - 1951.0 / 21600.0;
__device__ constexpr double E3 = 500.0 / 1113.0 - 22642.0 / 50085.0;
__device__ constexpr double E4 = 125.0 / 192.0 - 451.0 / 720.0;
__device__ constexpr double E5 = -2187.0 / 6784.0 - -12231.0 / 42400.0;
__device__ constexpr double E6 = 11.0 / 84.0 - 649.0 / 6300.0;
__device__ constexpr double E7 = -1.0 / 60.0;
__device__ constexpr double M1 = 6025192743.0 / 30085553152.0 / 2.0;
__device__ constexpr double M3 = 51252292925.0 / 65400821598.0 / 2.0;
__device__ constexpr double M4 = -2691868925.0 / 45128329728.0 / 2.0;
__device__ constexpr double M5 = 187940372067.0 / 1594534317056.0 / 2.0;
__device__ constexpr double M6 = -1776094331.0 / 19743644256.0 / 2.0;
__device__ constexpr double M7 = 11237099.0 / 235043384.0 / 2.0;

// previous representable value (towards -inf) of a finite x
template <class T> __device__ __forceinline__ T prev_value(T x);
template <> __device__ __forceinline__ float prev_value<float>(float x) {
    const int i = __float_as_int(x);
    if (x == 0.0f) return -__int_as_float(1);
    return __int_as_float(x > 0.0f ? i - 1 : i + 1);
}
template <> __device__ __forceinline__ double prev_value<double>(double x) {
    const long long i = __double_as_longlong(x);
    if (x == 0.0) return -__longlong_as_double(1);
    return __longlong_as_double(x > 0.0 ? i - 1 : i + 1);
}

// reciprocal: hardware v_rcp_f32 (1 ulp) in single precision; the ratio only
// steers the step size, so a 1-ulp quotient is as good as a rounded one
template <class T> __device__ __forceinline__ T fast_rcp(T x);
template <> __device__ __forceinline__ float fast_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double fast_rcp<double>(double x) { return 1.0 / x; }

template <class T> __device__ __forceinline__ T tpow(T x, T e);
template <> __device__ __forceinline__ float tpow<float>(float x, float e) { return powf(x, e); }
template <> __device__ __forceinline__ double tpow<double>(double x, double e) { return pow(x, e); }

// Controller factor 0.9 * ratio^(-1/5) clamped to [dfac, 10], taken from the
// squared ratio r2 as 0.9 * r2^(-1/10) with the hardware log2 / exp2 in
// single precision (relative error ~1e-7 in the next dt, far below the
// step-size granularity that matters; ratios below the normal range give a
// factor >= 10).
__device__ __forceinline__ double next_dt_sq(double dt, float r2, bool below_one) {
    // r2 == 0 gives log2 = -inf and a factor clamped to 10
    float f = 0.9f * __builtin_amdgcn_exp2f(-0.1f * __builtin_amdgcn_logf(r2));
    f = fminf(10.0f, fmaxf(f, below_one ? 1.0f : 0.2f));
    return dt * (double)f;
}

// previous two window values and the maxima count per state
template <int S, class T, bool kOn> struct PeakState {
    T p1[S], p2[S];
    int32_t count[S];
};
template <int S, class T> struct PeakState<S, T, false> {};

template <class Sys, class T, bool kPeaks = false>
struct Lane {
    static constexpr int S = Sys::S;
    static constexpr bool kP = kPeaks;
    PeakState<Sys::S, T, kPeaks> pk;
    T p[Sys::NP];
    T y[S], k1[S];
    T s0, c0;  // sin / cos of (T)t for systems with time terms
    double t, dt;
    int32_t steps;  // attempted steps (max_steps is capped at 2^31 - 1)
    int32_t hits;   // accepted steps whose span held a save point
    double next_save;  // save_t[si], +inf once every save is emitted
    double after_save;  // save_t[si + 1] (+inf past the end), loaded a save ahead
    int32_t si;
    int32_t status;
    bool done;
    bool oor;  // a float step took an argument outside the branch-free forms' range
    // streaming statistics over the feature window
    double shift[S], s1[S], s2[S], s3[S];  // s3: sum of window position x (value - shift)
    T mn[S], mx[S];
    int32_t cnt;
    T last[S];
};

template <class Sys, class T>
__device__ __forceinline__ void write_row(const LaneArgs &a, int64_t ic, int32_t si, const T (&v)[Sys::S]) {
    if (a.traj) {
        T *o = reinterpret_cast<T *>(a.traj) + ((int64_t)si * a.n + ic) * Sys::S;
#pragma unroll
        for (int i = 0; i < Sys::S; i++) o[i] = v[i];
    }
}

// One sample of the steady window: running statistics and the peak test.
template <class Sys, class T, class LaneT>
__device__ __forceinline__ void window_sample(const LaneArgs &a, LaneT &L, int64_t ic,
                                              const T (&v)[Sys::S]) {
    constexpr int S = Sys::S;
    if (L.cnt == 0) {
#pragma unroll
        for (int i = 0; i < S; i++) {
            L.shift[i] = (double)v[i];
            L.mn[i] = v[i];
            L.mx[i] = v[i];
        }
    }
#pragma unroll
    for (int i = 0; i < S; i++) {
        const double d = (double)v[i] - L.shift[i];
        L.s1[i] += d;
        L.s2[i] = fma(d, d, L.s2[i]);
        L.s3[i] = fma((double)L.cnt, d, L.s3[i]);
        L.mn[i] = fmin(L.mn[i], v[i]);
        L.mx[i] = fmax(L.mx[i], v[i]);
    }
    if constexpr (LaneT::kP) {
        // window position L.cnt; the value before it is a maximum when it
        // exceeds both neighbours (NaN compares false: no maximum)
        if (L.cnt >= 2) {
#pragma unroll
            for (int i = 0; i < S; i++) {
                if (L.pk.p1[i] > L.pk.p2[i] && L.pk.p1[i] > v[i]) {
                    if (L.pk.count[i] < a.peak_cap)
                        reinterpret_cast<T *>(a.peak_vals)[((int64_t)L.pk.count[i] * a.n + ic) * S + i] =
                            L.pk.p1[i];
                    L.pk.count[i]++;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < S; i++) {
            L.pk.p2[i] = L.pk.p1[i];
            L.pk.p1[i] = v[i];
        }
    }
    L.cnt++;
}

template <class Sys, class T, class LaneT>
__device__ __forceinline__ void emit(const LaneArgs &a, LaneT &L, int64_t ic, int32_t si,
                                     const T (&v)[Sys::S]) {
    write_row<Sys, T>(a, ic, si, v);
    if (si >= a.steady_idx) window_sample<Sys, T>(a, L, ic, v);
#pragma unroll
    for (int i = 0; i < Sys::S; i++) L.last[i] = v[i];
}

template <class Sys, class T, class LaneT>
__device__ __forceinline__ void lane_init(const LaneArgs &a, LaneT &L, int64_t ic) {
    constexpr int S = Sys::S;
    const T *y0 = reinterpret_cast<const T *>(a.y0) + ic * S;
    const double *pp = a.params + ic * a.p_stride;
#pragma unroll
    for (int i = 0; i < Sys::NP; i++) L.p[i] = (T)pp[i];
#pragma unroll
    for (int i = 0; i < S; i++) {
        L.y[i] = y0[i];
        L.s1[i] = 0.0;
        L.s2[i] = 0.0;
        L.s3[i] = 0.0;
        L.shift[i] = 0.0;
    }
    if constexpr (LaneT::kP) {
#pragma unroll
        for (int i = 0; i < S; i++) L.pk.count[i] = 0;
    }
    L.t = a.t0;
    if constexpr (Sys::kTime) dsincos<T>((T)a.t0, L.s0, L.c0);
    L.steps = 0;
    L.hits = 0;
    L.si = 0;
    L.cnt = 0;
    L.status = BH_LANE_OK;
    L.done = false;
    L.oor = false;
    // f0 and the initial step (torchdiffeq _select_initial_step, order 4)
    Sys::f(L.p, (T)a.t0, L.y, L.k1);
    const T ratol = (T)a.atol, rrtol = (T)a.rtol;
    T scale[S], d0 = 0, d1 = 0;
#pragma unroll
    for (int i = 0; i < S; i++) {
        scale[i] = ratol + fabs(L.y[i]) * rrtol;
        const T r0 = L.y[i] / scale[i], r1 = L.k1[i] / scale[i];
        d0 += r0 * r0;
        d1 += r1 * r1;
    }
    d0 = sqrt(d0 / (T)S);
    d1 = sqrt(d1 / (T)S);
    T h0 = (d0 < (T)1e-5 || d1 < (T)1e-5) ? (T)1e-6 : (T)0.01 * d0 / d1;
    h0 = fabs(h0);
    T yt[S], ft[S];
#pragma unroll
    for (int i = 0; i < S; i++) yt[i] = L.y[i] + h0 * L.k1[i];
    Sys::f(L.p, (T)(a.t0 + (double)h0), yt, ft);
    T d2 = 0;
#pragma unroll
    for (int i = 0; i < S; i++) {
        const T r = (ft[i] - L.k1[i]) / scale[i];
        d2 += r * r;
    }
    d2 = fabs(sqrt(d2 / (T)S) / h0);
    T h1;
    if (d1 <= (T)1e-15 && d2 <= (T)1e-15)
        h1 = fmax((T)1e-6, h0 * (T)1e-3);
    else
        h1 = tpow<T>((T)0.01 / fmax(d1, d2), (T)0.2);
    L.dt = (double)fmin((T)100 * h0, fabs(h1));
    // save times at or before t0 return y0 exactly (torchdiffeq solution[0] = y0)
    L.next_save = a.save_t[0];
    L.after_save = a.n_save > 1 ? a.save_t[1] : INFINITY;
    while (L.next_save <= a.t0) {
        emit<Sys, T>(a, L, ic, L.si, L.y);
        L.si++;
        L.next_save = L.after_save;
        L.after_save = L.si + 1 < a.n_save ? a.save_t[L.si + 1] : INFINITY;
    }
    if (L.si >= a.n_save) L.done = true;
}

template <class Sys, class T, class LaneT>
__device__ __forceinline__ void lane_fill_rest(const LaneArgs &a, LaneT &L, int64_t ic, bool inf) {
    constexpr int S = Sys::S;
    T v[S];
#pragma unroll
    for (int i = 0; i < S; i++) v[i] = inf ? (T)INFINITY : L.y[i];
    if (L.si >= a.n_save) return;
    if (a.traj)
        for (int32_t si = L.si; si < a.n_save; si++) write_row<Sys, T>(a, ic, si, v);
#pragma unroll
    for (int i = 0; i < S; i++) L.last[i] = v[i];
    const int32_t w0 = max(L.si, a.steady_idx);
    L.si = a.n_save;
    if (w0 >= a.n_save) return;
    // The first window sample goes through the full update (it may open the
    // window, and it decides whether the sample before it is a peak).  The
    // k others equal it: no peak among them, extremes unchanged, and their
    // sums are added at once, so a lane stopped early costs O(1) here
    // instead of one update per remaining save.
    window_sample<Sys, T>(a, L, ic, v);
    const int32_t k = a.n_save - w0 - 1;
    if (k > 0) {
        // window positions cnt .. cnt + k - 1
        const double pos_sum = (double)k * (double)L.cnt + 0.5 * (double)k * (double)(k - 1);
#pragma unroll
        for (int i = 0; i < S; i++) {
            const double d = (double)v[i] - L.shift[i];
            const double kd = (double)k * d;
            L.s1[i] += kd;
            L.s2[i] = fma(kd, d, L.s2[i]);
            L.s3[i] = fma(pos_sum, d, L.s3[i]);
        }
        if constexpr (LaneT::kP) {
#pragma unroll
            for (int i = 0; i < S; i++) L.pk.p2[i] = v[i];
        }
        L.cnt += k;
    }
}

// The stages of one attempted step from (L.y, L.k1) with step h and the
// step's time terms; returns the squared RMS error ratio (accept iff <= 1;
// the controller takes its -1/10 power, which is the -1/5 power of the
// ratio).  kFast takes the systems' branch-free float forms, which add
// their argument magnitudes to amax.
template <class Sys, class T, bool kFast>
__device__ __forceinline__ T lane_stages(const LaneArgs &a, const T (&p)[Sys::NP], const T (&y)[Sys::S],
                                         const T (&k1)[Sys::S], T h, const T (&aux)[5],
                                         T (&k3)[Sys::S], T (&k4)[Sys::S], T (&k5)[Sys::S],
                                         T (&k6)[Sys::S], T (&k7)[Sys::S], T (&y1)[Sys::S], T &amax) {
    constexpr int S = Sys::S;
    T k2[S], ys[S];
    // Stage inputs y + h * sum_j a_ij k_j: the tableau constants are
    // immediate operands and h is applied once per component.
#pragma unroll
    for (int i = 0; i < S; i++) ys[i] = fma(h, k1[i] * (T)A21, y[i]);
    rhs_aux<Sys, T, kFast>(p, aux[0], ys, k2, amax);
#pragma unroll
    for (int i = 0; i < S; i++) ys[i] = fma(h, k1[i] * (T)A31 + k2[i] * (T)A32, y[i]);
    rhs_aux<Sys, T, kFast>(p, aux[1], ys, k3, amax);
#pragma unroll
    for (int i = 0; i < S; i++)
        ys[i] = fma(h, k1[i] * (T)A41 + k2[i] * (T)A42 + k3[i] * (T)A43, y[i]);
    rhs_aux<Sys, T, kFast>(p, aux[2], ys, k4, amax);
#pragma unroll
    for (int i = 0; i < S; i++)
        ys[i] = fma(h, k1[i] * (T)A51 + k2[i] * (T)A52 + k3[i] * (T)A53 + k4[i] * (T)A54, y[i]);
    rhs_aux<Sys, T, kFast>(p, aux[3], ys, k5, amax);
#pragma unroll
    for (int i = 0; i < S; i++)
        ys[i] = fma(h, k1[i] * (T)A61 + k2[i] * (T)A62 + k3[i] * (T)A63 + k4[i] * (T)A64 +
                           k5[i] * (T)A65, y[i]);
    rhs_aux<Sys, T, kFast>(p, aux[4], ys, k6, amax);
#pragma unroll
    for (int i = 0; i < S; i++)
        y1[i] = fma(h, k1[i] * (T)B1 + k3[i] * (T)B3 + k4[i] * (T)B4 + k5[i] * (T)B5 +
                           k6[i] * (T)B6, y[i]);
    rhs_aux<Sys, T, kFast>(p, aux[4], y1, k7, amax);
    T acc = 0;
    const T ratol = (T)a.atol, rrtol = (T)a.rtol;
#pragma unroll
    for (int i = 0; i < S; i++) {
        const T err = h * (k1[i] * (T)E1 + k3[i] * (T)E3 + k4[i] * (T)E4 + k5[i] * (T)E5 +
                           k6[i] * (T)E6 + k7[i] * (T)E7);
        const T tol = ratol + rrtol * fmax(fabs(y[i]), fabs(y1[i]));
        const T r = err * fast_rcp<T>(tol);
        acc += r * r;
    }
    return acc * (T)(1.0 / S);
}

// One attempted step.  Sets L.done when the lane has finished.
//
// In float the stages use the systems' branch-free forms (the sine of the
// pendulum, the time terms of Duffing): on a wave alone on its SIMD, which
// is how the long trajectories of a launch finish, a divergent branch costs
// about as much as fifteen arithmetic instructions, and the general forms
// have one per evaluation.  A step whose arguments leave the branch-free
// range marks the lane (L.oor); the lane's trajectory is then integrated
// again from its initial state with the general forms (kExact) before its
// results are written, so every result comes from in-range evaluations.
template <class Sys, class T, bool kExact = false, class LaneT>
__device__ __forceinline__ void lane_step(const LaneArgs &a, LaneT &L, int64_t ic) {
    constexpr int S = Sys::S;
    constexpr bool kFast = std::is_same<T, float>::value && !kExact;
    const double tA = L.t, tB = L.t + L.dt;
    const T tf = (T)tA, h = (T)L.dt, tBf = (T)tB, t1p = prev_value<T>(tBf);
    const T tst[4] = {tf + (T)C2 * h, tf + (T)C3 * h, tf + (T)C4 * h, tf + (T)C5 * h};
    T k3[S], k4[S], k5[S], k6[S], k7[S], y1[S], aux[5];
    // sin / cos of the end time: the c = 1 stages here, the start of the next
    // step after an accept
    T sB, cB, amax = 0;
    bool in_range = true;
    if constexpr (kFast)
        in_range = Sys::time_terms_fast(tf, tst, L.s0, L.c0, tBf, t1p, sB, cB, aux);
    else
        Sys::time_terms(L.steps, tf, tst, L.s0, L.c0, tBf, t1p, sB, cB, aux);
    const T ratio = lane_stages<Sys, T, kFast>(a, L.p, L.y, L.k1, h, aux, k3, k4, k5, k6, k7, y1, amax);
    if constexpr (kFast && (Sys::kTime || Sys::kFastRhs)) {
        if (Sys::kFastRhs) in_range = in_range && amax <= kFastTrigMax;
        L.oor = L.oor || !in_range;
    }
    // Failure and acceptance are decided after the stages, so the common path
    // has one divergent branch (the dense output) and one exit test; the
    // state update is a select.
    bool finite = true;
#pragma unroll
    for (int i = 0; i < S; i++) finite = finite && isfinite(L.y[i]);
    const bool bad = !(tB > tA) || !finite || isnan(ratio);
    const bool accept = !bad && ratio <= (T)1;
    L.steps += bad ? 0 : 1;
    if (accept && L.next_save <= tB) {
        L.hits++;
        // dense output on [tA, tB] (torchdiffeq _interp_fit / _interp_evaluate)
        T ca[S], cb[S], cc[S], cd[S];
        const T m1 = h * (T)M1, m3 = h * (T)M3, m4 = h * (T)M4, m5 = h * (T)M5,
                m6 = h * (T)M6, m7 = h * (T)M7;
#pragma unroll
        for (int i = 0; i < S; i++) {
            const T ymid = L.y[i] + (L.k1[i] * m1 + k3[i] * m3 + k4[i] * m4 + k5[i] * m5 +
                                     k6[i] * m6 + k7[i] * m7);
            const T f0 = L.k1[i], f1 = k7[i], a0 = L.y[i], a1 = y1[i];
            ca[i] = (T)2 * h * (f1 - f0) - (T)8 * (a1 + a0) + (T)16 * ymid;
            cb[i] = h * ((T)5 * f0 - (T)3 * f1) + (T)18 * a0 + (T)14 * a1 - (T)32 * ymid;
            cc[i] = h * (f1 - (T)4 * f0) - (T)11 * a0 - (T)5 * a1 + (T)16 * ymid;
            cd[i] = h * f0;
        }
        // x = (s - tA) / (tB - tA) with one reciprocal per step (a save at tB
        // itself takes x = 1 exactly)
        const double inv_span = 1.0 / (tB - tA);
        do {
            const T x = L.next_save == tB ? T(1) : (T)((L.next_save - tA) * inv_span);
            const T x2 = x * x, x3 = x2 * x, x4 = x3 * x;
            T v[S];
#pragma unroll
            for (int i = 0; i < S; i++) {
                T sv = L.y[i] + x * cd[i];
                sv = sv + x2 * cc[i];
                sv = sv + x3 * cb[i];
                v[i] = sv + x4 * ca[i];
            }
            emit<Sys, T>(a, L, ic, L.si, v);
            L.si++;
            // the next save time was loaded one save ahead, so the loop test
            // does not wait on memory
            L.next_save = L.after_save;
            L.after_save = L.si + 1 < a.n_save ? a.save_t[L.si + 1] : INFINITY;
        } while (L.next_save <= tB);
    }
    T ymax = 0;
#pragma unroll
    for (int i = 0; i < S; i++) {
        const T yn = y1[i], yo = L.y[i], kn = k7[i], ko = L.k1[i];
        L.y[i] = accept ? yn : yo;
        L.k1[i] = accept ? kn : ko;
        ymax = fmax(ymax, fabs(L.y[i]));
    }
    L.t = accept ? tB : L.t;
    if constexpr (Sys::kTime) {
        L.s0 = accept ? sB : L.s0;
        L.c0 = accept ? cB : L.c0;
    }
    L.dt = next_dt_sq(L.dt, (float)ratio, ratio < (T)1);
    const bool finished = accept && L.si >= a.n_save;
    const bool stop = accept && !finished && (a.flags & BH_FLAG_STOP) && ymax >= (T)a.stop_thr;
    const bool maxed = L.steps >= (int32_t)a.max_steps;
    // a lane marked out of range ends here and is integrated again by
    // lane_redo_exact before its results are written
    if (bad || finished || stop || maxed || L.oor) {
        if (!finished && !L.oor) {
            bool inf = false;
            if (bad) {
                L.status = finite ? BH_LANE_DT_UNDERFLOW : BH_LANE_NONFINITE;
            } else if (stop) {
                L.status = BH_LANE_DIVERGED;
                inf = (a.flags & BH_FLAG_INF_FILL) != 0;
            } else {
                L.status = BH_LANE_MAX_STEPS;
            }
            lane_fill_rest<Sys, T>(a, L, ic, inf);
        }
        L.done = true;
    }
}

// A lane marked by an out-of-range float step integrates its trajectory
// again with the general forms (see lane_step); its first pass's writes to
// the trajectory and peak buffers are all written again.
template <class Sys, class T, class LaneT>
__device__ __forceinline__ void lane_redo_exact(const LaneArgs &a, LaneT &L, int64_t ic) {
    if constexpr (std::is_same<T, float>::value && (Sys::kTime || Sys::kFastRhs)) {
        if (L.oor) {
            lane_init<Sys, T>(a, L, ic);
            while (!L.done) lane_step<Sys, T, true>(a, L, ic);
        }
    }
}

template <class Sys, class T, class LaneT>
__device__ __forceinline__ void lane_finish(const LaneArgs &a, LaneT &L, int64_t ic) {
    constexpr int S = Sys::S;
    lane_redo_exact<Sys, T>(a, L, ic);
    if (a.feats) {
        const bool trend = (a.flags & BH_FLAG_TREND) != 0;
        const int nf = trend ? BH_NFEAT_TREND : BH_NFEAT;
        double *o = a.feats + ic * S * nf;
#pragma unroll
        for (int i = 0; i < S; i++) {
            if (L.cnt > 0) {
                const double inv = 1.0 / (double)L.cnt;
                const double m1 = L.s1[i] * inv;
                const double var = fmax(L.s2[i] * inv - m1 * m1, 0.0);
                const bool fin = isfinite((double)L.mn[i]) && isfinite((double)L.mx[i]);
                o[i * nf + 0] = fin ? L.shift[i] + m1 : (double)L.mx[i];
                o[i * nf + 1] = fin ? sqrt(var) : NAN;
                o[i * nf + 2] = (double)L.mn[i];
                o[i * nf + 3] = (double)L.mx[i];
                if (trend) {
                    // least-squares slope against the window position 0 .. cnt-1:
                    // sum (t - tbar)(x - xbar) / (sum (t - tbar)^2 + 1e-10), the
                    // shift cancelling in the numerator
                    const double c = (double)L.cnt;
                    const double tbar = 0.5 * (c - 1.0);
                    const double ss_tt = c * (c * c - 1.0) / 12.0;
                    const double ss_tx = L.s3[i] - tbar * L.s1[i];
                    o[i * nf + 4] = fin ? ss_tx / (ss_tt + 1e-10) : NAN;
                }
            } else {
                for (int j = 0; j < nf; j++) o[i * nf + j] = NAN;
            }
        }
    }
    if (a.fin) {
        T *o = reinterpret_cast<T *>(a.fin) + ic * S;
#pragma unroll
        for (int i = 0; i < S; i++) o[i] = L.last[i];
    }
    if (a.status) a.status[ic] = L.status;
    if (a.nsteps)
        a.nsteps[ic] = (a.flags & BH_FLAG_COUNT_SAVE_STEPS)
                           ? (int64_t)(((uint64_t)(uint32_t)L.hits << 32) | (uint32_t)L.steps)
                           : (int64_t)L.steps;
    if constexpr (LaneT::kP) {
#pragma unroll
        for (int i = 0; i < S; i++) a.peak_counts[ic * S + i] = L.pk.count[i];
    }
}

// kRefill = false: every trajectory has its own lane (the launch covers n),
// so a lane steps until its trajectory ends and writes its results; the loop
// carries no refill bookkeeping.  kRefill = true: a resident set smaller
// than n, refilled from a.counter.
template <class Sys, class T, bool kPeaks, bool kRefill>
__global__ void __launch_bounds__(256) dopri5_lane_kernel(LaneArgs a) {
    int64_t ic = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= a.n) return;
    Lane<Sys, T, kPeaks> L;
    lane_init<Sys, T>(a, L, ic);
    if constexpr (!kRefill) {
        while (!L.done) lane_step<Sys, T>(a, L, ic);  // not done unless init emitted every save
        lane_finish<Sys, T>(a, L, ic);
    } else {
        bool live = true;
        int quorum = (int)((a.flags >> BH_REFILL_QUORUM_SHIFT) & 0x7fu);
        quorum = quorum == 0 ? kDefaultRefillQuorum : (quorum > 64 ? 64 : quorum);
        // Every pass runs one step for each lane still integrating; in the
        // same pass finished lanes write their results and take the next ICs
        // from the counter once the refill quorum is met.  The loop ends by
        // a wave-wide vote: had it ended per lane, the finished lanes would
        // leave the loop and wait there until the last lane of the wave
        // finished, and refills could only happen once per wave.
        do {
            if (!L.done) lane_step<Sys, T>(a, L, ic);
            bool go = L.done && live;
            if (quorum > 1) {
                // wave-uniform counts: finished lanes wait for the quorum, or
                // for the last lane still integrating
                const int waiting = __popcll(__ballot(go));
                const int holding = __popcll(__ballot(live));
                go = go && waiting >= (quorum < holding ? quorum : holding);
            }
            if (go) {
                lane_finish<Sys, T>(a, L, ic);
                ic = (int64_t)atomicAdd(a.counter, 1ull);
                live = ic < a.n;
                if (live) lane_init<Sys, T>(a, L, ic);
            }
        } while (__any(live));
    }
}

// Two-phase launch.  Phase 1 gives every trajectory its own lane for at most
// phase_iters attempted steps; a lane still integrating then appends its
// whole state (the Lane record) to a compact buffer at a slot taken from the
// counter, and leaves.  Phase 2 resumes slot r on lane r, so the long
// trajectories continue in dense waves instead of a few live lanes per
// wave.  A trajectory's arithmetic does not depend on the lane, the slot or
// the launch that runs it, so the results equal a single-phase launch.
template <class Sys, class T, int kPhase>
__global__ void __launch_bounds__(256) dopri5_phase_kernel(LaneArgs a) {
    using LaneT = Lane<Sys, T, false>;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    LaneT *state = reinterpret_cast<LaneT *>(a.state);
    LaneT L;
    int64_t ic;
    if constexpr (kPhase == 1) {
        ic = r;
        if (ic >= a.n) return;
        lane_init<Sys, T>(a, L, ic);
        for (int32_t it = 0; !L.done && it < a.phase_iters; it++) lane_step<Sys, T>(a, L, ic);
        if (!L.done) {
            const unsigned long long slot = atomicAdd(a.counter, 1ull);
            state[slot] = L;
            a.state_ic[slot] = ic;
            return;
        }
    } else {
        // the counter holds the number of suspended states (kernel boundary
        // after phase 1)
        if (r >= (int64_t)*a.counter) return;
        L = state[r];
        ic = a.state_ic[r];
        while (!L.done) lane_step<Sys, T>(a, L, ic);
    }
    lane_finish<Sys, T>(a, L, ic);
}

template <class Sys, class T>
constexpr int64_t lane_state_bytes() {
    return (int64_t)sizeof(Lane<Sys, T, false>);
}

__global__ void set_counter_kernel(unsigned long long *c, unsigned long long v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = v;
}

}  // namespace bh
