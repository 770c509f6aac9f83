// window.hip -- order statistics of the saved feature window (median,
// quantiles), the non-streaming features of the reference's torch feature
// extractor (median: x.median(dim=0), the lower middle value;
// quantile: torch.quantile, from two order statistics;
// src/pybasin/ts_torch/calculators/torch_features_statistical.py:30-38,154-157).
//
// One lane per (trajectory, state) pair reads its window [start, n_save) of
// the (n_save, n, S) trajectory buffer (consecutive lanes read consecutive
// addresses) and finds the value of each requested rank by radix selection
// on the order-preserving integer image of the floats, 4 bits per pass
// (8 passes for float, 16 for double), 16 counters in registers.  A window
// holding a NaN yields NaN (torch.median and torch.quantile propagate NaN);
// so does a rank outside [0, window length).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bh {
namespace win {

template <class T> struct Key;
template <> struct Key<float> {
    typedef uint32_t U;
    static constexpr int BITS = 32;
    __device__ static U of(float v) {
        const U u = __float_as_uint(v);
        return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    }
    __device__ static float value(U k) {
        return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
    }
};
template <> struct Key<double> {
    typedef uint64_t U;
    static constexpr int BITS = 64;
    __device__ static U of(double v) {
        const U u = (U)__double_as_longlong(v);
        return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    }
    __device__ static double value(U k) {
        return __longlong_as_double((long long)((k & 0x8000000000000000ull)
                                                    ? (k & 0x7fffffffffffffffull) : ~k));
    }
};

template <class T>
__global__ void __launch_bounds__(256) order_stat_kernel(const T *traj, int64_t n_save, int64_t n,
                                                         int S, int64_t start, const int64_t *ranks,
                                                         int n_ranks, T *out) {
    typedef typename Key<T>::U U;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // ic * S + s
    const int64_t np = n * S;
    if (p >= np) return;
    const int64_t stride = n * S;
    const T *col = traj + p;
    bool nan = false;
    for (int64_t t = start; t < n_save; t++) nan |= isnan(col[t * stride]);
    for (int r = 0; r < n_ranks; r++) {
        if (nan) {
            out[(int64_t)r * np + p] = (T)NAN;
            continue;
        }
        int64_t k = ranks[r];  // rank within the window, 0-based
        if (k < 0 || k >= n_save - start) {  // no such order statistic
            out[(int64_t)r * np + p] = (T)NAN;
            continue;
        }
        U prefix = 0, mask = 0;
        for (int shift = Key<T>::BITS - 4; shift >= 0; shift -= 4) {
            int64_t cnt[16];
#pragma unroll
            for (int d = 0; d < 16; d++) cnt[d] = 0;
            for (int64_t t = start; t < n_save; t++) {
                const U key = Key<T>::of(col[t * stride]);
                if ((key & mask) == prefix) {
                    const int d = (int)((key >> shift) & 15u);
#pragma unroll
                    for (int e = 0; e < 16; e++) cnt[e] += (d == e);
                }
            }
            int d = 0;
            while (d < 15 && k >= cnt[d]) {
                k -= cnt[d];
                d++;
            }
            prefix |= (U)d << shift;
            mask |= (U)15u << shift;
        }
        out[(int64_t)r * np + p] = Key<T>::value(prefix);
    }
}

}  // namespace win
}  // namespace bh
