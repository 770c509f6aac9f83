// Device right-hand sides for the per-lane Dormand-Prince kernel.
//
// Each functor has a compile-time state size S and parameter count NP and
// evaluates dy = f(t, y; p) entirely in registers.  The arithmetic follows
// the reference's torch RHS term by term (same operand order; the compiler
// may contract a multiply-add into one FMA).
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

template <class T> __device__ __forceinline__ T dsin(T x);
template <class T> __device__ __forceinline__ T dcos(T x);
template <> __device__ __forceinline__ float dsin<float>(float x) { return sinf(x); }
template <> __device__ __forceinline__ double dsin<double>(double x) { return sin(x); }
template <> __device__ __forceinline__ float dcos<float>(float x) { return cosf(x); }
template <> __device__ __forceinline__ double dcos<double>(double x) { return cos(x); }

// tests/unit/solvers/test_solver.py:16-18 -- dy/dt = p0 * y
template <int S_> struct Decay {
    static constexpr int S = S_, NP = 1;
    static constexpr int FLOPS = S_;
    template <class T>
    __device__ __forceinline__ static void f(const T (&p)[NP], T, const T (&y)[S], T (&dy)[S]) {
#pragma unroll
        for (int i = 0; i < S; i++) dy[i] = p[0] * y[i];
    }
};

// y' = M y with M = [[p0, p1], [p2, p3]] (harmonic / damped oscillators in tests)
struct Linear2 {
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
struct Pendulum {
    static constexpr int S = 2, NP = 3;
    static constexpr int FLOPS = 5;
    template <class T>
    __device__ __forceinline__ static void f(const T (&p)[NP], T, const T (&y)[S], T (&dy)[S]) {
        dy[0] = y[1];
        dy[1] = -p[0] * y[1] + p[1] - p[2] * dsin<T>(y[0]);
    }
};

// case_studies/duffing_oscillator/duffing_ode.py:81-88
//   x' = v;  v' = -delta*v - k3*x^3 + A*cos(t);  p = [delta, k3, A]
struct Duffing {
    static constexpr int S = 2, NP = 3;
    static constexpr int FLOPS = 8;
    template <class T>
    __device__ __forceinline__ static void f(const T (&p)[NP], T t, const T (&y)[S], T (&dy)[S]) {
        dy[0] = y[1];
        dy[1] = -p[0] * y[1] - p[1] * (y[0] * y[0] * y[0]) + p[2] * dcos<T>(t);
    }
};

// case_studies/lorenz/lorenz_ode.py:128-152, including the freeze mask:
// the derivative is multiplied by (max|y| < 200).
struct Lorenz {
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
