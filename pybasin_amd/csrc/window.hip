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

#include "../../include/basinhip.h"

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

// ---- series kernels: autocorrelation periodicity, spectral frequency ratio ----
//
// A workgroup of 4 waves takes G consecutive (trajectory, state) series --
// consecutive addresses within each saved row -- and copies their window
// [start, n_save) into LDS with coalesced loads (G values per row); each
// wave then processes G / 4 series, one at a time, with its 64 lanes.
constexpr int kSeriesBlock = 256;

// wave-wide sum of a double (xor butterfly over the 64 lanes)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <class T>
__device__ __forceinline__ void load_series(const T *traj, int64_t n, int S, int64_t start,
                                            int64_t w, int64_t p0, int G, int64_t np, T *lds) {
    // lds[g * w + i] = traj[(start + i) * n * S + p0 + g]
    const int64_t stride = n * S;
    for (int64_t e = threadIdx.x; e < (int64_t)G * w; e += blockDim.x) {
        const int64_t i = e / G;
        const int g = (int)(e - i * G);
        const int64_t p = p0 + g;
        lds[(int64_t)g * w + i] = p < np ? traj[(start + i) * stride + p] : T(0);
    }
}

// mean and unbiased / population variance pieces of one series in LDS
template <class T>
__device__ __forceinline__ void series_moments(const T *x, int64_t w, double &mean, double &ss) {
    const int lane = threadIdx.x & 63;
    double s1 = 0.0;
    for (int64_t i = lane; i < w; i += 64) s1 += (double)x[i];
    mean = wave_sum(s1) / (double)w;
    double s2 = 0.0;
    for (int64_t i = lane; i < w; i += 64) {
        const double d = (double)x[i] - mean;
        s2 = fma(d, d, s2);
    }
    ss = wave_sum(s2);  // sum of squared deviations
}

// Autocorrelation periodicity of each series (the reference's
// autocorrelation_periodicity, torch_features_autocorrelation.py:268-331):
// z = (x - mean) / (std + 1e-10) with the unbiased std; r_k = sum_i z_i z_{i+k}
// (the zero-padded FFT product of the reference is this linear
// autocorrelation); ac_k = r_k / (r_0 + 1e-10); the first local maximum of ac
// in scipy's sense (a run of equal values with a smaller value on each side,
// marked at its midpoint rounded down; torch_feature_utilities.py:146-211) at
// a lag >= min_lag with ac > peak_threshold gives strength = ac there and
// period = its lag; none gives 0 and 0.  Lags are computed 64 at a time (one
// per lane, double sums over z in LDS) and scanned in order, so the work
// stops at the first block holding a valid peak.
template <class T>
__global__ void __launch_bounds__(kSeriesBlock)
    autocorr_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                    int min_lag, double thr, T *out_strength, T *out_period) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    double *acbuf = reinterpret_cast<double *>(smem + ((int64_t)G * w * sizeof(T) + 15) / 16 * 16);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *ac = acbuf + wave * 64;
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        T *z = lds + (int64_t)g * w;
        double mean, ss;
        series_moments<T>(z, w, mean, ss);
        const double sd = w > 1 ? sqrt(ss / (double)(w - 1)) : NAN;
        const double inv = 1.0 / (sd + 1e-10);
        for (int64_t i = lane; i < w; i += 64) z[i] = (T)(((double)z[i] - mean) * inv);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        double r0 = 0.0;
        // scipy-style streaming peak search state (lane 0 only)
        double prev = 0.0, cand_val = 0.0;
        int64_t cand = -1;
        bool found = false;
        double strength = 0.0, period = 0.0;
        for (int64_t base = 0; base < w && !found; base += 64) {
            const int64_t k = base + lane;
            double r = 0.0;
            if (k < w)
                for (int64_t i = 0; i + k < w; i++) r = fma((double)z[i], (double)z[i + k], r);
            if (base == 0) r0 = __shfl(r, 0, 64);
            ac[lane] = r / (r0 + 1e-10);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (lane == 0) {
                const int64_t kend = base + 64 < w ? base + 64 : w;
                for (int64_t j = base; j < kend && !found; j++) {
                    const double v = ac[j - base];
                    if (j >= 1) {
                        if (cand >= 0) {
                            if (v == cand_val) {
                                // the plateau goes on (to the last value at most)
                            } else if (v < cand_val) {
                                const int64_t mid = (cand + j - 1) / 2;
                                if (mid >= min_lag && cand_val > thr) {
                                    found = true;
                                    strength = cand_val;
                                    period = (double)mid;
                                }
                                cand = -1;
                            } else {
                                cand = j <= w - 2 ? j : -1;  // a new rise
                                cand_val = v;
                            }
                        } else if (prev < v && j <= w - 2) {
                            cand = j;
                            cand_val = v;
                        }
                    }
                    prev = v;
                }
            }
            found = __shfl((int)found, 0, 64) != 0;
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) {
            out_strength[p] = (T)strength;
            out_period[p] = (T)period;
        }
    }
}

// Spectral frequency ratio of each series (the reference's
// spectral_frequency_ratio, torch_features_frequency.py:390-441): power
// |X_b|^2 of the centred window at the rfft bins b = 1 .. floor(w / 2)
// (bin 0 set to zero), each by Goertzel's recurrence in double
// (s_i = c_i + 2 cos(2 pi b / w) s_{i-1} - s_{i-2}; |X_b|^2 = s1^2 + s2^2 -
// 2 cos(2 pi b / w) s1 s2), four bins per lane at a time; the dominant bin
// f1, the strongest bin f2 at least two bins away from it, ratio f2 / f1;
// 0 when the largest power is below 1e-10.  Equal powers resolve to the
// lower bin.
template <class T>
__global__ void __launch_bounds__(kSeriesBlock)
    spectral_ratio_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                          T *out_ratio) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t nb = w / 2 + 1;  // rfft bins
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        const T *x = lds + (int64_t)g * w;
        double mean, ss;
        series_moments<T>(x, w, mean, ss);
        // best and runner-up over this lane's bins, then across the wave
        double best = -1.0;
        int64_t best_b = 0;
        // powers are kept for the second search: recomputed (cheap relative
        // to memory) rather than stored
        for (int pass = 0; pass < 2; pass++) {
            int64_t first = 0;
            if (pass == 1) {
                // wave argmax of the first pass: largest power, lowest bin
                double bv = best;
                int64_t bb = best_b;
#pragma unroll
                for (int m = 32; m >= 1; m >>= 1) {
                    const double ov = __shfl_xor(bv, m, 64);
                    const int64_t ob = __shfl_xor(bb, m, 64);
                    if (ov > bv || (ov == bv && ob < bb)) {
                        bv = ov;
                        bb = ob;
                    }
                }
                if (!(bv >= 1e-10)) {  // no spectral content (or NaN)
                    if (lane == 0) out_ratio[p] = (T)0;
                    break;
                }
                first = bb;
                best = -1.0;
                best_b = 0;
            }
            for (int64_t b0 = 1 + lane * 4; b0 < nb; b0 += 64 * 4) {
                double cw[4], s1[4], s2[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    cw[q] = 2.0 * cospi(2.0 * (double)(b0 + q) / (double)w);
                    s1[q] = 0.0;
                    s2[q] = 0.0;
                }
                for (int64_t i = 0; i < w; i++) {
                    const double c = (double)x[i] - mean;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const double s0 = fma(cw[q], s1[q], c - s2[q]);
                        s2[q] = s1[q];
                        s1[q] = s0;
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int64_t b = b0 + q;
                    if (b >= nb) continue;
                    if (pass == 1 && (b - first <= 1 && first - b <= 1)) continue;
                    const double pw = fma(s1[q], s1[q], s2[q] * s2[q]) - cw[q] * s1[q] * s2[q];
                    if (pw > best) {
                        best = pw;
                        best_b = b;
                    }
                }
            }
            if (pass == 1) {
                double bv = best;
                int64_t bb = best_b;
#pragma unroll
                for (int m = 32; m >= 1; m >>= 1) {
                    const double ov = __shfl_xor(bv, m, 64);
                    const int64_t ob = __shfl_xor(bb, m, 64);
                    if (ov > bv || (ov == bv && ob < bb)) {
                        bv = ov;
                        bb = ob;
                    }
                }
                // bin 0 (power set to zero) is a candidate too: when no bin
                // two or more away has positive power, the runner-up is a
                // zero-power bin and the ratio is taken as 0
                const int64_t second = bv > 0.0 ? bb : 0;
                if (lane == 0) out_ratio[p] = (T)((double)second / (double)first);
            }
        }
    }
}

// Bins b0 .. b0 + Q - 1 of one staged series by Goertzel, in double; bins at
// or past `top` are skipped.  Writes (re, im) of bins below n_coeff and, with
// pw, keeps each power in the wave's LDS row and adds it into m0 and b * P_b
// into m1.
template <class T, int Q>
__device__ __forceinline__ void goertzel_bins(const T *x, int64_t w, int64_t b0, int64_t top,
                                              int n_coeff, int64_t np, int64_t p,
                                              double *out_coeff, double *pw, double &m0,
                                              double &m1) {
    double cw[Q], s1[Q], s2[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        cw[q] = 2.0 * cospi(2.0 * (double)(b0 + q) / (double)w);
        s1[q] = 0.0;
        s2[q] = 0.0;
    }
#pragma unroll 4
    for (int64_t i = 0; i < w; i++) {
        const double c = (double)x[i];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const double s0 = fma(cw[q], s1[q], c - s2[q]);
            s2[q] = s1[q];
            s1[q] = s0;
        }
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int64_t b = b0 + q;
        if (b >= top) continue;
        const double re = fma(0.5 * cw[q], s1[q], -s2[q]);
        // the DC and Nyquist bins are real: +0, as the FFT gives (their angle is 0 or pi)
        const double im = (b == 0 || 2 * b == w) ? 0.0 : sinpi(2.0 * (double)b / (double)w) * s1[q];
        if (b < n_coeff) {
            out_coeff[(b * 2) * np + p] = re;
            out_coeff[(b * 2 + 1) * np + p] = im;
        }
        if (pw) {
            const double pwr = fma(re, re, im * im);
            pw[b] = pwr;
            m0 += pwr;
            m1 = fma((double)b, pwr, m1);
        }
    }
}

template <class T, int Q>
__device__ __forceinline__ void series_bins(const T *x, int64_t w, int64_t top, int n_coeff,
                                            int64_t np, int64_t p, double *out_coeff,
                                            double *pw, double &m0, double &m1, int lane) {
    for (int64_t b0 = (int64_t)lane * Q; b0 < top; b0 += 64 * Q)
        goertzel_bins<T, Q>(x, w, b0, top, n_coeff, np, p, out_coeff, pw, m0, m1);
}

// Complex rfft bins and the aggregated spectrum of each series of the
// window (the reference's fft_coefficient, spkt_welch_density and
// fft_aggregated, all on torch.fft.rfft of the raw window;
// torch_features_frequency.py:11-70): X_b = sum_i x_i e^{-2 pi i b i / w} by
// Goertzel's recurrence in double, s_i = x_i + 2 cos(om) s_{i-1} - s_{i-2},
// X_b = e^{i om} s_{w-1} - s_{w-2} (om = 2 pi b / w), Q = 1..4 bins per lane
// at a time (the fewest that cover the bins asked for in one round of 64
// lanes).  Bins 0 .. n_coeff - 1 are written as (re, im) (bins past the last
// rfft bin as 0); with out_agg, the powers P_b = |X_b|^2 of all w / 2 + 1
// bins are kept in the wave's LDS row and the centroid, variance, skew and
// kurtosis of the normalised power spectrum p_b = P_b / (sum P + 1e-10)
// follow from centred sums, and its entropy -sum c log c over
// c = max(p_b, 1e-10) (fourier_entropy, torch_features_entropy_complexity.py:97-107),
// the reference's 1e-10 guards included.
template <class T>
__global__ void __launch_bounds__(kSeriesBlock)
    spectrum_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                    int n_coeff, double *out_coeff, double *out_agg) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    const int64_t nb = w / 2 + 1;  // rfft bins
    T *lds = reinterpret_cast<T *>(smem);
    double *pbase = reinterpret_cast<double *>(smem + ((int64_t)G * w * sizeof(T) + 15) / 16 * 16);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *pw = out_agg ? pbase + (int64_t)wave * nb : nullptr;
    const int64_t top = out_agg ? nb : (n_coeff < nb ? (int64_t)n_coeff : nb);
    const int64_t rounds = (top + 63) / 64;
    const int Q = rounds >= 4 ? 4 : (int)rounds;
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        const T *x = lds + (int64_t)g * w;
        double m0 = 0.0, m1 = 0.0;
        switch (Q) {
            case 1: series_bins<T, 1>(x, w, top, n_coeff, np, p, out_coeff, pw, m0, m1, lane); break;
            case 2: series_bins<T, 2>(x, w, top, n_coeff, np, p, out_coeff, pw, m0, m1, lane); break;
            case 3: series_bins<T, 3>(x, w, top, n_coeff, np, p, out_coeff, pw, m0, m1, lane); break;
            default: series_bins<T, 4>(x, w, top, n_coeff, np, p, out_coeff, pw, m0, m1, lane); break;
        }
        for (int64_t b = nb + lane; b < n_coeff; b += 64) {  // past the last bin
            out_coeff[(b * 2) * np + p] = 0.0;
            out_coeff[(b * 2 + 1) * np + p] = 0.0;
        }
        if (out_agg) {
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            const double tot = wave_sum(m0) + 1e-10;
            const double centroid = wave_sum(m1) / tot;
            double c2 = 0.0, c3 = 0.0, c4 = 0.0, ent = 0.0;
            for (int64_t b = lane; b < nb; b += 64) {
                const double d = (double)b - centroid, pn = pw[b] / tot;
                const double d2 = d * d;
                c2 = fma(d2, pn, c2);
                c3 = fma(d2 * d, pn, c3);
                c4 = fma(d2 * d2, pn, c4);
                const double pc = pn < 1e-10 ? 1e-10 : pn;  // NaN stays NaN, as torch's clamp
                ent = fma(-pc, log(pc), ent);
            }
            const double var = wave_sum(c2);
            c3 = wave_sum(c3);
            c4 = wave_sum(c4);
            ent = wave_sum(ent);
            if (lane == 0) {
                out_agg[0 * np + p] = centroid;
                out_agg[1 * np + p] = var;
                out_agg[2 * np + p] = c3 / (pow(var, 1.5) + 1e-10);
                out_agg[3 * np + p] = c4 / (var * var + 1e-10);
                out_agg[4 * np + p] = ent;
            }
            __builtin_amdgcn_wave_barrier();  // pw reused by the wave's next series
        }
    }
}

// ---- comprehensive window features: raw accumulators per series ----------
//
// One lane per (trajectory, state) series of the window rows [start,
// n_save), consecutive lanes on consecutive addresses of each saved row.
// Two passes in double: the first collects sums, extremes and their first
// and last positions, changes, counts, crossings, chunk energies, the lag
// products of c3 and the time-reversal statistic, the seven largest |x| and
// the peaks of support 1, 3 and 5; the second, given the mean and the
// standard deviation of the first (from sums shifted by the first value),
// the central moments, the counts and longest runs about the mean, the
// counts beyond r standard deviations, the autocovariances at lags 1..9,
// the duplicate-extreme counts, the index-mass quantiles and the ten bins of
// binned_entropy.  The slots are
// the BH_WF_* indices of include/basinhip.h; a window holding a non-finite
// value gives NaN in every slot.

constexpr int kWfTop = 7;     // largest |x| kept (mean_n_absolute_max up to 7)
constexpr int kWfLags = 9;    // autocovariance lags 1..9
constexpr int kWfChunks = 10; // energy_ratio_by_chunks segments
constexpr int kWfBins = 10;   // binned_entropy bins (max_bins = 10)
__device__ constexpr double kWfSigma[10] = {0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 5.0, 6.0, 7.0, 10.0};
__device__ constexpr double kWfMass[8] = {0.1, 0.2, 0.3, 0.4, 0.6, 0.7, 0.8, 0.9};

// peak of support k at the centre of the last 2k + 1 values of the window
// w (w[0] newest): strictly greater than each of its k neighbours on both
// sides
template <class T, int K>
__device__ __forceinline__ bool wf_peak(const T (&w)[11]) {
    const T c = w[K];
    bool pk = true;
#pragma unroll
    for (int j = 0; j < K; j++) pk = pk && c > w[j] && c > w[2 * K - j];
    return pk;
}

template <class T>
__global__ void __launch_bounds__(256) window_features_kernel(const T *traj, int64_t n_save, int64_t n,
                                                              int S, int64_t start, double *out) {
    const int64_t np = n * S;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // ic * S + s
    if (p >= np) return;
    const T *col = traj + start * np + p;
    const int64_t N = n_save - start;
    double *o = out + p;
    auto put = [&](int slot, double v) { o[(int64_t)slot * np] = v; };

    // ---- pass 1 ----
    const double x0 = (double)col[0];
    bool fin = true;
    double s1 = 0, s2 = 0, sh1 = 0, sh2 = 0, sabs = 0, mn = x0, mx = x0;
    int64_t first_max = 0, last_max = 0, first_min = 0, last_min = 0;
    double sad = 0, ssd = 0;
    int64_t c_pos = 0, c_neg = 0, c_zero = 0, c_one = 0, c_mone = 0, r_unit = 0, r_neg = 0, r_pos = 0;
    int64_t cr0 = 0, crm = 0, crp = 0;
    // chunk energies: a running sum stored when its chunk ends (the chunk
    // boundaries are the same for every lane)
    const int64_t seg = N / kWfChunks;
    double chunk_sum = 0;
    int chunk_at = 0;
    int64_t chunk_end = seg;
    double c3s[3] = {0, 0, 0}, trs[3] = {0, 0, 0};
    double hist[6];  // x_{i-1} .. x_{i-6}
#pragma unroll
    for (int j = 0; j < 6; j++) hist[j] = 0;
    double top[kWfTop];
#pragma unroll
    for (int j = 0; j < kWfTop; j++) top[j] = -1.0;
    T w[11];  // x_i .. x_{i-10} for the peak tests
#pragma unroll
    for (int j = 0; j < 11; j++) w[j] = T(0);
    int64_t pk1 = 0, pk3 = 0, pk5 = 0;
    double prev = x0;
    // four samples in flight: the load of sample i + 4 is issued while
    // sample i is reduced (indices clamped to the window, so every load is
    // in bounds)
    const int64_t last = N - 1;
    T q0 = col[0], q1 = col[min((int64_t)1, last) * np], q2 = col[min((int64_t)2, last) * np],
      q3 = col[min((int64_t)3, last) * np];
    for (int64_t i = 0; i < N; i++) {
        const T xt = q0;
        q0 = q1;
        q1 = q2;
        q2 = q3;
        q3 = col[min(i + 4, last) * np];
        const double x = (double)xt;
        fin = fin && isfinite(x);
        s1 += x;
        s2 = fma(x, x, s2);
        const double d = x - x0;
        sh1 += d;
        sh2 = fma(d, d, sh2);
        sabs += fabs(x);
        if (x > mx) { mx = x; first_max = i; }
        if (x >= mx) last_max = i;
        if (x < mn) { mn = x; first_min = i; }
        if (x <= mn) last_min = i;
        if (i > 0) {
            const double dx = x - prev;
            sad += fabs(dx);
            ssd = fma(dx, dx, ssd);
            cr0 += (prev > 0.0) != (x > 0.0);
            crm += (prev > -1.0) != (x > -1.0);
            crp += (prev > 1.0) != (x > 1.0);
        }
        c_pos += x > 0.0;
        c_neg += x < 0.0;
        c_zero += x == 0.0;
        c_one += x == 1.0;
        c_mone += x == -1.0;
        r_unit += (x >= -1.0) && (x <= 1.0);
        r_neg += (x >= -1e12) && (x <= 0.0);
        r_pos += (x >= 0.0) && (x <= 1e12);
        if (chunk_at < kWfChunks) {
            chunk_sum = fma(x, x, chunk_sum);
            if (i + 1 == chunk_end) {
                put(BH_WF_CHUNK + chunk_at, chunk_sum);
                chunk_sum = 0;
                chunk_at++;
                chunk_end += seg;
            }
        }
        // lag products ending at i: x_{i-2l} x_{i-l} x_i and x_i^2 x_{i-l} - x_{i-l} x_{i-2l}^2
#pragma unroll
        for (int l = 1; l <= 3; l++) {
            if (i >= 2 * l) {
                const double a = hist[2 * l - 1], b = hist[l - 1];
                c3s[l - 1] += a * b * x;
                trs[l - 1] += x * x * b - b * a * a;
            }
        }
#pragma unroll
        for (int j = 5; j > 0; j--) hist[j] = hist[j - 1];
        hist[0] = x;
        // the seven largest |x| in descending order (most samples fall below
        // the seventh and skip the insertion)
        double v = fabs(x);
        if (v > top[kWfTop - 1]) {
#pragma unroll
            for (int j = 0; j < kWfTop; j++) {
                const double hi = fmax(top[j], v), lo = fmin(top[j], v);
                top[j] = hi;
                v = lo;
            }
        }
#pragma unroll
        for (int j = 10; j > 0; j--) w[j] = w[j - 1];
        w[0] = xt;
        if (i >= 2) pk1 += wf_peak<T, 1>(w);
        if (i >= 6) pk3 += wf_peak<T, 3>(w);
        if (i >= 10) pk5 += wf_peak<T, 5>(w);
        prev = x;
    }
    if (!fin) {
        for (int k = 0; k < BH_WF_SLOTS; k++) put(k, NAN);
        return;
    }
    const double Nd = (double)N;
    put(BH_WF_N, Nd);
    put(BH_WF_SUM, s1);
    put(BH_WF_SUMSQ, s2);
    put(BH_WF_MIN, mn);
    put(BH_WF_MAX, mx);
    put(BH_WF_FIRST_MAX, (double)first_max);
    put(BH_WF_LAST_MAX, (double)last_max);
    put(BH_WF_FIRST_MIN, (double)first_min);
    put(BH_WF_LAST_MIN, (double)last_min);
    put(BH_WF_ABS_CHANGE, sad);
    put(BH_WF_SQ_CHANGE, ssd);
    put(BH_WF_X_FIRST, x0);
    put(BH_WF_X_SECOND, N > 1 ? (double)col[np] : NAN);
    put(BH_WF_X_SECOND_LAST, N > 1 ? hist[1] : NAN);
    put(BH_WF_X_LAST, hist[0]);
    put(BH_WF_COUNT_POS, (double)c_pos);
    put(BH_WF_COUNT_NEG, (double)c_neg);
    put(BH_WF_COUNT_ZERO, (double)c_zero);
    put(BH_WF_COUNT_ONE, (double)c_one);
    put(BH_WF_COUNT_MINUS_ONE, (double)c_mone);
    put(BH_WF_RANGE_UNIT, (double)r_unit);
    put(BH_WF_RANGE_NONPOS, (double)r_neg);
    put(BH_WF_RANGE_NONNEG, (double)r_pos);
    put(BH_WF_CROSS_0, (double)cr0);
    put(BH_WF_CROSS_M1, (double)crm);
    put(BH_WF_CROSS_P1, (double)crp);
    // windows shorter than ten rows have empty chunks
    for (int j = chunk_at; j < kWfChunks; j++) put(BH_WF_CHUNK + j, 0.0);
#pragma unroll
    for (int l = 0; l < 3; l++) {
        put(BH_WF_C3 + l, c3s[l]);
        put(BH_WF_TREV + l, trs[l]);
    }
    put(BH_WF_SUM_ABS, sabs);
#pragma unroll
    for (int j = 0; j < kWfTop; j++) put(BH_WF_TOP_ABS + j, top[j]);
    put(BH_WF_PEAKS_1, (double)pk1);
    put(BH_WF_PEAKS_3, (double)pk3);
    put(BH_WF_PEAKS_5, (double)pk5);

    // ---- pass 2: about the mean ----
    const double mu = s1 / Nd;
    const double m1 = sh1 / Nd;
    const double sd = sqrt(fmax(sh2 / Nd - m1 * m1, 0.0));
    const double tbar = 0.5 * (Nd - 1.0);
    double c2 = 0, c3 = 0, c4 = 0, stx = 0;
    int64_t above = 0, below = 0, run_a = 0, run_b = 0, best_a = 0, best_b = 0, eq_max = 0, eq_min = 0;
    int64_t beyond[10];
#pragma unroll
    for (int j = 0; j < 10; j++) beyond[j] = 0;
    double acv[kWfLags], dh[kWfLags];
#pragma unroll
    for (int j = 0; j < kWfLags; j++) acv[j] = dh[j] = 0;
    int64_t mass_idx[8];
#pragma unroll
    for (int j = 0; j < 8; j++) mass_idx[j] = -1;
    double cum = 0;
    // binned_entropy's bins, formed in the series dtype as the reference
    // does: (x - min) / (range + 1e-10) * (10 - 1e-6) truncated, range 0 -> 1
    const T tmn = (T)mn;
    T trng = (T)mx - tmn;
    trng = trng == T(0) ? T(1) : trng;
    const T tden = trng + T(1e-10), tscale = T(10.0 - 1e-6);
    int64_t bins[kWfBins];
#pragma unroll
    for (int j = 0; j < kWfBins; j++) bins[j] = 0;
    q0 = col[0];
    q1 = col[min((int64_t)1, last) * np];
    q2 = col[min((int64_t)2, last) * np];
    q3 = col[min((int64_t)3, last) * np];
    for (int64_t i = 0; i < N; i++) {
        const T xt = q0;
        const double x = (double)xt;
        q0 = q1;
        q1 = q2;
        q2 = q3;
        q3 = col[min(i + 4, last) * np];
        const double d = x - mu;
        const double d2 = d * d;
        c2 += d2;
        c3 = fma(d2, d, c3);
        c4 = fma(d2, d2, c4);
        stx = fma((double)i - tbar, d, stx);
        above += x > mu;
        below += x < mu;
        run_a = x > mu ? run_a + 1 : 0;
        run_b = x < mu ? run_b + 1 : 0;
        best_a = max(best_a, run_a);
        best_b = max(best_b, run_b);
        eq_max += x == mx;
        eq_min += x == mn;
        const double ad = fabs(d);
#pragma unroll
        for (int j = 0; j < 10; j++) beyond[j] += ad > kWfSigma[j] * sd;
#pragma unroll
        for (int l = 0; l < kWfLags; l++)
            if (i > l) acv[l] = fma(dh[l], d, acv[l]);
#pragma unroll
        for (int j = kWfLags - 1; j > 0; j--) dh[j] = dh[j - 1];
        dh[0] = d;
        cum += fabs(x);
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (mass_idx[j] < 0 && cum >= kWfMass[j] * sabs) mass_idx[j] = i;
        int bin = (int)((xt - tmn) / tden * tscale);
        bin = bin < 0 ? 0 : (bin > kWfBins - 1 ? kWfBins - 1 : bin);
#pragma unroll
        for (int j = 0; j < kWfBins; j++) bins[j] += bin == j;
    }
#pragma unroll
    for (int j = 0; j < kWfBins; j++) put(BH_WF_BINS + j, (double)bins[j]);
    put(BH_WF_CM2, c2);
    put(BH_WF_CM3, c3);
    put(BH_WF_CM4, c4);
    put(BH_WF_SS_TX, stx);
    put(BH_WF_COUNT_ABOVE_MEAN, (double)above);
    put(BH_WF_COUNT_BELOW_MEAN, (double)below);
    put(BH_WF_STRIKE_ABOVE, (double)best_a);
    put(BH_WF_STRIKE_BELOW, (double)best_b);
    put(BH_WF_COUNT_MAX, (double)eq_max);
    put(BH_WF_COUNT_MIN, (double)eq_min);
#pragma unroll
    for (int j = 0; j < 10; j++) put(BH_WF_BEYOND + j, (double)beyond[j]);
#pragma unroll
    for (int l = 0; l < kWfLags; l++) put(BH_WF_AUTOCOV + l, acv[l]);
#pragma unroll
    for (int j = 0; j < 8; j++) put(BH_WF_MASS + j, (double)(mass_idx[j] < 0 ? N - 1 : mass_idx[j]));
}

// ---- autocovariances of many lags ---------------------------------------------

// C_k = sum_{i < w - k} (x_i - mean)(x_{i+k} - mean) for k = 0 .. max_lag of
// each series of the window (0 for k >= w), in double: the linear
// autocorrelation of the centred series that the reference forms as the
// inverse FFT of the zero-padded power spectrum (agg_autocorrelation and
// ar_coefficient, torch_features_autocorrelation.py:72-96,
// torch_features_trend.py:86-131).  One wave per series staged in LDS, one
// lag per lane (64 lags a round); x_i is read by all lanes at once (an LDS
// broadcast), x_{i+k} by consecutive lanes.  out (max_lag + 1, n, dim).
template <class T>
__global__ void __launch_bounds__(kSeriesBlock)
    autocov_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                   int max_lag, double *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        const T *x = lds + (int64_t)g * w;
        double mean, ss;
        series_moments<T>(x, w, mean, ss);
        for (int base = 0; base <= max_lag; base += 64) {
            const int64_t k = base + lane;
            double c0 = 0.0, c1 = 0.0;
            if (k < w) {
                const int64_t m = w - k;
                int64_t i = 0;
                for (; i + 1 < m; i += 2) {  // two independent sums
                    c0 = fma((double)x[i] - mean, (double)x[i + k] - mean, c0);
                    c1 = fma((double)x[i + 1] - mean, (double)x[i + 1 + k] - mean, c1);
                }
                if (i < m) c0 = fma((double)x[i] - mean, (double)x[i + k] - mean, c0);
            }
            if (k <= max_lag) out[k * np + p] = c0 + c1;
        }
    }
}

// ---- chunk trends, polynomial fits, leading digits, wavelets ----------------

// a chunk's maximum, minimum, mean and square deviation sum, merged from two
// disjoint chunks by Chan's update
struct ChunkAgg {
    double mx, mn, mean, m2, cnt;
};

__device__ __forceinline__ void chunk_merge(ChunkAgg &a, const ChunkAgg &b) {
    if (a.cnt == 0.0) {
        a = b;
        return;
    }
    const double tot = a.cnt + b.cnt, d = b.mean - a.mean;
    a.mean = fma(d, b.cnt / tot, a.mean);
    a.m2 += b.m2 + d * d * (a.cnt * b.cnt / tot);
    a.cnt = tot;
    a.mx = fmax(a.mx, b.mx);
    a.mn = fmin(a.mn, b.mn);
}

// shifted sums of one aggregate sequence a_c over chunks c = 0, 1, ..:
// b_c = a_c - a_0, sum b, sum b^2, sum c b
struct TrendSums {
    double a0[4], sb[4], sbb[4], scb[4];
    double nc;
};

__device__ __forceinline__ void trend_add(TrendSums &t, const ChunkAgg &g) {
    const double v[4] = {g.mx, g.mn, g.mean, g.m2 / g.cnt};
#pragma unroll
    for (int f = 0; f < 4; f++) {
        if (t.nc == 0.0) t.a0[f] = v[f];
        const double b = v[f] - t.a0[f];
        t.sb[f] += b;
        t.sbb[f] = fma(b, b, t.sbb[f]);
        t.scb[f] = fma(t.nc, b, t.scb[f]);
    }
    t.nc += 1.0;
}

// rvalue, intercept, slope, stderr of the least-squares line through the
// aggregates (linear_trend_batched, torch_features_trend.py:174-212, over
// the chunk index); zeros when there are fewer than two chunks
__device__ __forceinline__ void trend_write(const TrendSums &t, double *out, int64_t np,
                                            int64_t p, int base) {
    const double nc = t.nc;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        double r = 0.0, icpt = 0.0, slope = 0.0, se = 0.0;
        if (nc >= 2.0) {
            const double tbar = 0.5 * (nc - 1.0);
            const double ss_tt = nc * (nc * nc - 1.0) / 12.0;
            const double ss_xx = fmax(t.sbb[f] - t.sb[f] * t.sb[f] / nc, 0.0);
            const double ss_tx = t.scb[f] - tbar * t.sb[f];
            slope = ss_tx / (ss_tt + 1e-10);
            icpt = (t.a0[f] + t.sb[f] / nc) - slope * tbar;
            r = ss_tx / (sqrt(ss_tt * ss_xx) + 1e-10);
            const double res = fmax(ss_xx - 2.0 * slope * ss_tx + slope * slope * ss_tt, 0.0);
            se = sqrt(res / (nc - 2.0) / (ss_tt + 1e-10));
        }
        out[(base + f * 4 + 0) * np + p] = r;
        out[(base + f * 4 + 1) * np + p] = icpt;
        out[(base + f * 4 + 2) * np + p] = slope;
        out[(base + f * 4 + 3) * np + p] = se;
    }
}

// The reference's leading digit of a > 0 (torch_features_advanced.py:39-52):
// floor(a / 10^floor(log10 a)), counted only when it is 1..9 (0 otherwise;
// values within rounding of a power of ten can give 0 or 10 there), in
// double.  The fast path takes the exponent from a float logarithm and
// scales by a table of 10^-e (e in [-31, 32]); an exponent estimate within
// 1e-4 of an integer, a scaled value within 1e-9 of a digit boundary or
// outside [1, 10), or an a outside [1e-30, 1e30] takes the exact formula.
__device__ __forceinline__ int leading_digit_exact(double a) {
    const double e = floor(log10(a));
    const int d = (int)floor(a / pow(10.0, e));
    return d >= 1 && d <= 9 ? d : 0;
}

__device__ __forceinline__ int leading_digit_fast(double a, const double *p10inv) {
    if (!(a >= 1e-30 && a <= 1e30)) return leading_digit_exact(a);
    const float lf = __log10f((float)a);
    const float ef = floorf(lf);
    if (lf - ef < 1e-4f || lf - ef > 1.0f - 1e-4f) return leading_digit_exact(a);
    const double sc = a * p10inv[(int)ef + 31];
    const double fr = sc - floor(sc);
    if (!(sc >= 1.0 && sc < 10.0) || fr < 1e-9 || fr > 1.0 - 1e-9) return leading_digit_exact(a);
    return (int)sc;
}

constexpr int kCwtWidths[4] = {2, 5, 10, 20};
constexpr int kCwtCoeffs = 15;

// One lane per series, two passes over the window (in double):
// * agg_linear_trend (torch_features_trend.py:60-83): chunks of 5 samples
//   give their max, min, mean and population variance; pairs of them merge
//   into chunks of 10 and fives of those into chunks of 50 (Chan's update);
//   each aggregate sequence is fitted against its chunk index (out rows
//   BH_TR_ALT + (size * 4 + agg) * 4 + attr: sizes 5, 10, 50; aggs max, min,
//   mean, var; attrs rvalue, intercept, slope, stderr; a sample tail shorter
//   than a chunk is left out, fewer than two chunks give zeros);
// * friedrich_coefficients and max_langevin_fixed_point (m = 3;
//   torch_features_dynamical.py:87-177): the normal equations of the cubic
//   through (x_i, x_{i+1} - x_i) plus 1e-6 I, solved by Gaussian elimination
//   with partial pivoting (rows BH_TR_FRIED + 0..3, coefficients of x^3 ..
//   x^0); pass 2 evaluates the cubic at every x_i and keeps the largest x_i
//   whose value changes sign to the next one (row BH_TR_LANGEVIN, 0 if none);
// * benford_correlation (torch_features_advanced.py:21-78): counts of the
//   leading digits floor(|x| / 10^floor(log10 |x|)) in 1..9 of the nonzero
//   |x_i| (in double) and their correlation with Benford's law as the
//   reference forms it (unbiased deviations, divided by 9 s_f s_b + 1e-10; 0
//   for fewer than 10 counted samples; row BH_TR_BENFORD);
// * with want_cwt, cwt_coefficients (torch_features_frequency.py:71-122):
//   the first 15 values of the correlation of the padded window with the
//   Ricker wavelet of width 2, 5, 10, 20 (taps t = -4w .. 4w, reflect padding,
//   replicate when the window has at most 4w rows; rows BH_TR_CWT + wi * 15 +
//   coeff; coefficients past the window are 0).
// A window holding a non-finite value gives NaN throughout.
template <class T>
__global__ void __launch_bounds__(256)
    trends_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int want_cwt,
                  double *out) {
    __shared__ double psi[17 + 41 + 81 + 161];
    __shared__ double p10inv[64];  // 10^-e for e = -31 .. 32
    if (threadIdx.x < 64) p10inv[threadIdx.x] = pow(10.0, (double)(31 - (int)threadIdx.x));
    for (int i = threadIdx.x; i < 300; i += blockDim.x) {
        const int wi = i < 17 ? 0 : (i < 58 ? 1 : (i < 139 ? 2 : 3));
        const int off = wi == 0 ? 0 : (wi == 1 ? 17 : (wi == 2 ? 58 : 139));
        const double sg = (double)kCwtWidths[wi], t = (double)(i - off - 4 * kCwtWidths[wi]);
        const double u = t / sg;
        psi[i] = 2.0 / (sqrt(3.0 * sg) * 1.3313353638003897) * (1.0 - u * u) * exp(-0.5 * u * u);
    }
    __syncthreads();
    const int64_t np = n * S;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const T *col = traj + start * np + p;
    const int64_t w = n_save - start;

    TrendSums tr[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        tr[k].nc = 0.0;
#pragma unroll
        for (int f = 0; f < 4; f++) tr[k].a0[f] = tr[k].sb[f] = tr[k].sbb[f] = tr[k].scb[f] = 0.0;
    }
    ChunkAgg c10 = {0, 0, 0, 0, 0}, c50 = {0, 0, 0, 0, 0};
    int k10 = 0, k50 = 0, c5 = 0;
    double x0 = 0.0, s1 = 0.0, s2 = 0.0, mx = 0.0, mn = 0.0;
    double sp[7] = {0, 0, 0, 0, 0, 0, 0}, svp[4] = {0, 0, 0, 0};
    int dig[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool bad = false;
    double prev = 0.0;
    for (int64_t i = 0; i < w; i++) {
        const double xv = (double)col[i * np];
        bad |= !isfinite(xv);
        // chunks of 5, merged into 10 and 50
        if (c5 == 0) {
            x0 = xv;
            s1 = s2 = 0.0;
            mx = mn = xv;
        }
        const double d = xv - x0;
        s1 += d;
        s2 = fma(d, d, s2);
        mx = fmax(mx, xv);
        mn = fmin(mn, xv);
        if (++c5 == 5) {
            c5 = 0;
            const ChunkAgg g5 = {mx, mn, x0 + s1 / 5.0, fmax(s2 - s1 * s1 / 5.0, 0.0), 5.0};
            trend_add(tr[0], g5);
            chunk_merge(c10, g5);
            if (++k10 == 2) {
                trend_add(tr[1], c10);
                chunk_merge(c50, c10);
                c10.cnt = 0.0;
                k10 = 0;
                if (++k50 == 5) {
                    trend_add(tr[2], c50);
                    c50.cnt = 0.0;
                    k50 = 0;
                }
            }
        }
        // the cubic's normal equations over the pairs (x_{i-1}, x_i - x_{i-1})
        if (i > 0) {
            const double pp = prev, v = xv - prev;
            double pk = 1.0;
#pragma unroll
            for (int k = 0; k < 7; k++) {
                sp[k] += pk;
                if (k < 4) svp[k] = fma(v, pk, svp[k]);
                pk *= pp;
            }
        }
        prev = xv;
        // leading digits
        const double a = fabs(xv);
        if (a > 0.0 && a < INFINITY) {
            const int dg = leading_digit_fast(a, p10inv);  // 0: not counted
#pragma unroll
            for (int k = 0; k < 9; k++) dig[k] += dg == k + 1 ? 1 : 0;
        }
    }
    if (bad) {
        const int rows = want_cwt ? BH_TR_SLOTS : BH_TR_CWT;
        for (int k = 0; k < rows; k++) out[k * np + p] = NAN;
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) trend_write(tr[k], out, np, p, BH_TR_ALT + k * 16);

    // (V^T V + 1e-6 I) c = V^T v, V = [x^3 x^2 x 1]
    double A[4][5];
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int c = 0; c < 4; c++) A[r][c] = sp[6 - r - c] + (r == c ? 1e-6 : 0.0);
        A[r][4] = svp[3 - r];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int piv = k;
        double big = fabs(A[k][k]);
#pragma unroll
        for (int r = k + 1; r < 4; r++)
            if (fabs(A[r][k]) > big) {
                big = fabs(A[r][k]);
                piv = r;
            }
        if (piv != k) {
#pragma unroll
            for (int c = 0; c < 5; c++) {
                // swap rows k and piv without dynamic register indices
                double hold = A[k][c];
#pragma unroll
                for (int r = k + 1; r < 4; r++)
                    if (r == piv) {
                        A[k][c] = A[r][c];
                        A[r][c] = hold;
                    }
            }
        }
#pragma unroll
        for (int r = k + 1; r < 4; r++) {
            const double f = A[r][k] / A[k][k];
#pragma unroll
            for (int c = k; c < 5; c++) A[r][c] = fma(-f, A[k][c], A[r][c]);
        }
    }
    double cf[4];
#pragma unroll
    for (int r = 3; r >= 0; r--) {
        double v = A[r][4];
#pragma unroll
        for (int c = r + 1; c < 4; c++) v = fma(-A[r][c], cf[c], v);
        cf[r] = v / A[r][r];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) out[(BH_TR_FRIED + k) * np + p] = cf[k];

    // the largest x_i at which the fitted cubic changes sign to x_{i+1}
    double best = -INFINITY;
    if (w >= 3) {
        double pa = (double)col[0];
        double fa = fma(fma(fma(cf[0], pa, cf[1]), pa, cf[2]), pa, cf[3]);
        for (int64_t i = 1; i + 1 < w; i++) {
            const double pb = (double)col[i * np];
            const double fb = fma(fma(fma(cf[0], pb, cf[1]), pb, cf[2]), pb, cf[3]);
            if (fa * fb < 0.0) best = fmax(best, pa);
            pa = pb;
            fa = fb;
        }
    }
    out[BH_TR_LANGEVIN * np + p] = best == -INFINITY ? 0.0 : best;

    // Benford: Pearson-like correlation as the reference forms it
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) tot += (double)dig[k];
    double corr = 0.0;
    if (tot >= 10.0) {
        double bf[9], fr[9], bm = 0.0, fm = 0.0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            bf[k] = log10(1.0 + 1.0 / (double)(k + 1));
            fr[k] = (double)dig[k] / (tot + 1e-10);
            bm += bf[k];
            fm += fr[k];
        }
        bm /= 9.0;
        fm /= 9.0;
        double sbb = 0.0, sff = 0.0, sfb = 0.0;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            sbb += (bf[k] - bm) * (bf[k] - bm);
            sff += (fr[k] - fm) * (fr[k] - fm);
            sfb += (fr[k] - fm) * (bf[k] - bm);
        }
        corr = sfb / (sqrt(sff / 8.0) * sqrt(sbb / 8.0) * 9.0 + 1e-10);
    }
    out[BH_TR_BENFORD * np + p] = corr;

    if (!want_cwt) return;
    int off = 0;
#pragma unroll
    for (int wi = 0; wi < 4; wi++) {
        const int wd = kCwtWidths[wi], pad = 4 * wd;
        const bool reflect = w > pad;
        double acc[kCwtCoeffs];
#pragma unroll
        for (int c = 0; c < kCwtCoeffs; c++) acc[c] = 0.0;
        // samples past x_{w-1+pad} feed only coefficients c >= w (written as 0)
        const int uend = (int)(w < kCwtCoeffs ? w : kCwtCoeffs) + pad;
        for (int u = -pad; u < uend; u++) {
            int64_t m = u;
            if (reflect) {
                if (m < 0) m = -m;
                if (m > w - 1) m = 2 * (w - 1) - m;
            } else {
                m = m < 0 ? 0 : (m > w - 1 ? w - 1 : m);
            }
            const double xu = (double)col[m * np];
#pragma unroll
            for (int c = 0; c < kCwtCoeffs; c++) {
                const int j = u - c + pad;
                if (j >= 0 && j <= 2 * pad) acc[c] = fma(psi[off + j], xu, acc[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < kCwtCoeffs; c++)
            out[(BH_TR_CWT + wi * kCwtCoeffs + c) * np + p] = c < w ? acc[c] : 0.0;
        off += 2 * pad + 1;
    }
}

// ---- peaks of wide support, ordinal patterns -------------------------------

constexpr int kPeakSupports = 4;
constexpr int kPeakWidths = 8;

// The number of consecutive samples x[i + dir], x[i + 2 dir], .. (at most
// cap, inside [0, w)) that c exceeds (Strict) or is not below: eight
// neighbours are loaded at a time, so a long run costs one LDS latency per
// eight samples rather than one per sample.
template <bool Strict, class T>
__device__ __forceinline__ int run_length(const T *x, int64_t i, int dir, int cap, int64_t w, T c) {
    int run = 0;
    while (run < cap) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int64_t idx = i + (int64_t)dir * (run + u + 1);
            v[u] = (idx >= 0 && idx < w) ? x[idx] : (T)INFINITY;
        }
        int f = 8;
#pragma unroll
        for (int u = 7; u >= 0; u--)
            if (!(Strict ? c > v[u] : c >= v[u])) f = u;
        run += f < cap - run ? f : cap - run;
        if (f < 8) break;
    }
    return run;
}

// One wave per series staged in LDS, one position per lane at a time: the
// strict peak radius of x_i (the largest k, up to the largest support asked
// for, with x_i > x_{i +- j} for all j <= k inside the window) and its
// non-strict radius (x_i >= x_{i +- j}, at most 8); the right-hand run is
// walked first and bounds the left-hand one, so a position on a slope costs
// one comparison.  out (12, n, dim): rows 0..3 count the positions whose strict
// radius reaches supports[r] (number_peaks, the reference's find_peak_mask:
// torch_features_pattern.py:64-105; a support of 0 counts nothing), rows
// 4..11 the positions whose non-strict radius reaches width 1..8 (the
// per-width counts of number_cwt_peaks, :154-190: x_i equal to the maximum of
// x_{i-w} .. x_{i+w}, the w first and last positions left out).  A window
// holding a non-finite value gives NaN.
template <class T>
__global__ void __launch_bounds__(kSeriesBlock)
    peak_counts_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                       int4 supports, double *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    unsigned short *cbase =
        reinterpret_cast<unsigned short *>(smem + ((int64_t)G * w * sizeof(T) + 15) / 16 * 16);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned short *cand = cbase + (int64_t)wave * w;  // the wave's candidate positions
    const int sup[kPeakSupports] = {supports.x, supports.y, supports.z, supports.w};
    int cap = 1;  // the strict run need not go past the largest support
#pragma unroll
    for (int k = 0; k < kPeakSupports; k++) cap = sup[k] > cap ? sup[k] : cap;
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        const T *x = lds + (int64_t)g * w;
        bool bad = false;
        for (int64_t i = lane; i < w; i += 64) bad |= !isfinite((double)x[i]);
        if (__any(bad)) {
            if (lane < kPeakSupports + kPeakWidths) out[lane * np + p] = NAN;
            continue;
        }
        // pass 1: the positions not below either neighbour (both inside the
        // window) are the only ones with a radius of 1 or more; they are
        // listed (ballot order) so that pass 2 spreads them over the lanes
        int nc = 0;
        for (int64_t i0 = 0; i0 < w; i0 += 64) {
            const int64_t i = i0 + lane;
            bool keep = false;
            if (i >= 1 && i + 1 < w) {
                const T c = x[i];
                keep = c >= x[i - 1] && c >= x[i + 1];
            }
            const unsigned long long m = __ballot(keep);
            if (keep) cand[nc + __popcll(m & ((1ull << lane) - 1ull))] = (unsigned short)i;
            nc += __popcll(m);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        double cs[kPeakSupports] = {0, 0, 0, 0};
        double cw[kPeakWidths] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k0 = lane; k0 < nc; k0 += 64) {
            const int64_t i = cand[k0];
            const T c = x[i];
            // only min(left, right) matters: the right run caps the left one
            const int rs = run_length<true, T>(x, i, 1, cap, w, c);
            const int r = rs ? run_length<true, T>(x, i, -1, rs, w, c) : 0;
            const int rn = run_length<false, T>(x, i, 1, kPeakWidths, w, c);
            const int q = rn ? run_length<false, T>(x, i, -1, rn, w, c) : 0;
#pragma unroll
            for (int k = 0; k < kPeakSupports; k++) cs[k] += (sup[k] > 0 && r >= sup[k]) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < kPeakWidths; k++) cw[k] += q >= k + 1 ? 1.0 : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kPeakSupports; k++) cs[k] = wave_sum(cs[k]);
#pragma unroll
        for (int k = 0; k < kPeakWidths; k++) cw[k] = wave_sum(cw[k]);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < kPeakSupports; k++) out[k * np + p] = cs[k];
#pragma unroll
            for (int k = 0; k < kPeakWidths; k++) out[(kPeakSupports + k) * np + p] = cw[k];
        }
        __builtin_amdgcn_wave_barrier();  // the list is rebuilt for the next series
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

constexpr int kPermDims = 5;  // embedding dimensions 3 .. 7
// first histogram bin of dimension 3 + k: sums of d! for the smaller d
__constant__ int kPermBase[kPermDims + 1] = {0, 6, 30, 150, 870, 5910};
__constant__ int kFact[8] = {1, 1, 2, 6, 24, 120, 720, 5040};

// permutation_entropy (torch_features_entropy_complexity.py:11-53) with
// tau = 1 and dimensions 3 .. 7, one wave per series staged in LDS: each
// pattern x_i .. x_{i+d-1} gets the reference's id sum_k argsort_k (d-1-k)!
// (argsort ascending, equal values in index order), counted in the wave's
// LDS histogram (16-bit counts, two per word); the entropy of the counts over
// the n - d + 1 patterns is divided by ln d!.  out (5, n, dim); a window
// shorter than d gives 0, a non-finite window NaN.
template <class T>
__global__ void __launch_bounds__(kSeriesBlock)
    perm_entropy_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                        double *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    unsigned *hbase = reinterpret_cast<unsigned *>(smem + ((int64_t)G * w * sizeof(T) + 15) / 16 * 16);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int kWords = 5910 / 2;
    unsigned *hist = hbase + wave * kWords;
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        const T *x = lds + (int64_t)g * w;
        bool bad = false;
        for (int64_t i = lane; i < w; i += 64) bad |= !isfinite((double)x[i]);
        if (__any(bad)) {
            if (lane < kPermDims) out[lane * np + p] = NAN;
            continue;
        }
        for (int k = lane; k < kWords; k += 64) hist[k] = 0u;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int64_t i = lane; i + 2 < w; i += 64) {
            T v[7];
#pragma unroll
            for (int j = 0; j < 7; j++) v[j] = i + j < w ? x[i + j] : (T)0;
            // rank of element j among the first d (ties by index), grown one
            // element at a time: element d - 1 joins with its comparisons to
            // the d - 1 before it; the pattern id is sum_j j * (d - 1 - rank_j)!
            int rk[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int j = 1; j < 3; j++)
#pragma unroll
                for (int m = 0; m < j; m++) {
                    const bool below = !(v[j] < v[m]);  // v[m] ranks below v[j] (m < j on ties)
                    rk[j] += below ? 1 : 0;
                    rk[m] += below ? 0 : 1;
                }
#pragma unroll
            for (int k = 0; k < kPermDims; k++) {
                const int d = 3 + k;
                if (i + d > w) break;
                if (d > 3) {
                    const int jn = d - 1;
#pragma unroll
                    for (int m = 0; m < 6; m++) {
                        if (m >= jn) break;
                        const bool below = !(v[jn] < v[m]);
                        rk[jn] += below ? 1 : 0;
                        rk[m] += below ? 0 : 1;
                    }
                }
                int id = 0;
#pragma unroll
                for (int j = 1; j < 7; j++) {
                    if (j >= d) break;
                    id += j * kFact[d - 1 - rk[j]];
                }
                const int bin = kPermBase[k] + id;
                atomicAdd(&hist[bin >> 1], 1u << (16 * (bin & 1)));
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int k = 0; k < kPermDims; k++) {
            const int d = 3 + k;
            const double npat = (double)(w - d + 1);
            double e = 0.0;
            for (int b = kPermBase[k] + lane; b < kPermBase[k + 1]; b += 64) {
                const unsigned c = (hist[b >> 1] >> (16 * (b & 1))) & 0xffffu;
                if (c) {
                    const double q = (double)c / npat;
                    e = fma(-q, log(q), e);
                }
            }
            e = wave_sum(e);
            if (lane == 0) out[k * np + p] = npat > 0.0 ? e / log((double)kFact[d]) : 0.0;
        }
        __builtin_amdgcn_wave_barrier();  // the histogram is cleared for the next series
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// ---- delay-embedding estimates: correlation dimension, largest Lyapunov ----

constexpr int kCdRadii = 50;

// correlation_dimension (torch_corr_dim.py:44-124; emb_dim = 4, lag = 1,
// n_rvals = 50): radii r_k = exp(L0 + k / 49 (L1 - L0)), L0 = ln(0.1 sd +
// 1e-10), L1 = ln(0.5 sd + 1e-10) with the unbiased sd of the window; C_k = the
// share of ordered pairs i != j of the m = w - 3 embedded vectors at
// Euclidean distance <= r_k; the slope of ln(C_k + 1e-30) on ln r_k over the
// nonzero C_k (NaN for fewer than two, or a degenerate fit).  One wave per
// series staged in LDS, rows i one per lane, the four coordinates of row j
// shared (an LDS broadcast); each pair i < j's squared distance is placed by a
// float logarithm and settled against the exact r_k^2 (double), then counted
// in the wave's LDS histogram (the counts of radius k add up the bins <= k).
template <class T>
__global__ void __launch_bounds__(kSeriesBlock)
    corr_dim_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                    double *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    unsigned char *tail = smem + ((int64_t)G * w * sizeof(T) + 15) / 16 * 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *r2 = reinterpret_cast<double *>(tail) + wave * kCdRadii;
    unsigned *hist = reinterpret_cast<unsigned *>(tail + 4 * kCdRadii * sizeof(double)) + wave * 64;
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        const T *x = lds + (int64_t)g * w;
        bool bad = false;
        for (int64_t i = lane; i < w; i += 64) bad |= !isfinite((double)x[i]);
        const int64_t m = w - 3;
        if (__any(bad) || m < 2) {
            if (lane == 0) out[p] = NAN;
            continue;
        }
        double mean, ss;
        series_moments<T>(x, w, mean, ss);
        const double sd = sqrt(ss / (double)(w - 1));
        const double L0 = log(0.1 * sd + 1e-10), L1 = log(0.5 * sd + 1e-10);
        const double dl = (L1 - L0) / (double)(kCdRadii - 1);
        if (lane < kCdRadii) {
            const double r = exp(fma((double)lane, dl, L0));
            r2[lane] = r * r;
        }
        hist[lane] = 0u;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const float fl0 = (float)L0, finv = dl > 0.0 ? (float)(1.0 / dl) : 0.0f;
        const double rmin2 = r2[0], rmax2 = r2[kCdRadii - 1];
        const bool fast = dl > 1e-3;  // the float guess is within 1e-4 of exact
        for (int64_t i0 = 0; i0 < m; i0 += 64) {
            const int64_t i = i0 + lane;
            const bool row = i < m;
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            if (row) {
                a0 = (double)x[i];
                a1 = (double)x[i + 1];
                a2 = (double)x[i + 2];
                a3 = (double)x[i + 3];
            }
            // pairs (i, j > i): j runs from the chunk's first row + 1, rows
            // below the lane's own are skipped; eight rows j at a time, their
            // 11 samples loaded together (LDS broadcasts)
            for (int64_t j0 = i0 + 1; j0 < m; j0 += 8) {
                double v[11];
#pragma unroll
                for (int u = 0; u < 11; u++) v[u] = j0 + u < w ? (double)x[j0 + u] : 0.0;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int64_t j = j0 + u;
                    const double e0 = a0 - v[u], e1 = a1 - v[u + 1];
                    const double e2 = a2 - v[u + 2], e3 = a3 - v[u + 3];
                    const double d2 = fma(e0, e0, fma(e1, e1, fma(e2, e2, e3 * e3)));
                    if (!row || j <= i || j >= m || !(d2 <= rmax2)) continue;
                    // smallest k with r_k^2 >= d2: from a float logarithm,
                    // settled against the exact r_k^2 only within 1e-3 of a
                    // boundary (the float error is ~3e-5 there)
                    int k = 0;
                    if (d2 > rmin2) {
                        const float t = (0.5f * __logf((float)d2) - fl0) * finv;
                        const float fr = t - floorf(t);
                        k = (int)ceilf(t);
                        k = k < 1 ? 1 : (k > kCdRadii - 1 ? kCdRadii - 1 : k);
                        if (!fast || fr < 1e-3f || fr > 0.999f) {
                            while (k > 1 && r2[k - 1] >= d2) k--;
                            while (k < kCdRadii - 1 && r2[k] < d2) k++;
                        }
                    }
                    atomicAdd(&hist[k], 1u);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // counts per radius (twice: ordered pairs) and the log-log fit
        const double pairs = (double)m * (double)(m - 1);
        double cum = 0.0, nv = 0.0, sr = 0.0, sc = 0.0, srr = 0.0, src = 0.0;
        for (int k = 0; k < kCdRadii; k++) {  // wave-uniform
            cum += 2.0 * (double)hist[k];
            const double c = cum / pairs;
            if (c > 0.0) {
                const double lr = fma((double)k, dl, L0), lc = log(c + 1e-30);
                nv += 1.0;
                sr += lr;
                sc += lc;
                srr = fma(lr, lr, srr);
                src = fma(lr, lc, src);
            }
        }
        const double den = nv * srr - sr * sr;
        if (lane == 0) out[p] = (nv >= 2.0 && fabs(den) > 1e-10) ? (nv * src - sr * sc) / den : NAN;
        __builtin_amdgcn_wave_barrier();  // the histogram is cleared for the next series
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

constexpr int kLrEmb = 10;
constexpr int kLrTraj = 20;

// lyapunov_r (torch_lyapunov_r.py:20-150; emb_dim = 10, lag = 1,
// trajectory_len = 20; the result is the slope, the caller divides by tau).
// One wave per series staged in LDS:
// 1. the minimum temporal separation from the mean frequency of the
//    zero-padded (2w - 1)-point spectrum of the raw window (Goertzel, bins 1 ..
//    w - 1 spread over the lanes): ceil(1 / (mf + 1e-10)) clamped to
//    [1, floor(w / 4)];
// 2. for each of the ntraj = m - 19 first embedded vectors (m = w - 9) its
//    nearest other vector among them at a temporal separation above the
//    minimum (Euclidean, the first on ties; rows one per lane, the ten
//    coordinates of row j shared in registers as a sliding window);
// 3. the mean over the vectors of ln |v_{i+k} - v_{nb_i+k}| for k = 0..19
//    (zero distances left out) and the least-squares slope over the finite
//    means (NaN for fewer than two).
// out (n, dim): the slope.  A non-finite window gives NaN.
template <class T>
__global__ void __launch_bounds__(kSeriesBlock)
    lyap_r_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                  double *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        const T *x = lds + (int64_t)g * w;
        bool bad = false;
        for (int64_t i = lane; i < w; i += 64) bad |= !isfinite((double)x[i]);
        const int64_t m = w - (kLrEmb - 1);
        if (__any(bad) || m < 1) {
            if (lane == 0) out[p] = NAN;
            continue;
        }
        // 1. mean frequency of the (2w - 1)-point spectrum, bins 1 .. w - 1
        const double M = (double)(2 * w - 1);
        double sp = 0.0, sfp = 0.0;
        for (int64_t k = 1 + lane; k < w; k += 64) {
            const double cw = 2.0 * cospi(2.0 * (double)k / M);
            double s1 = 0.0, s2 = 0.0;
            for (int64_t i = 0; i < w; i++) {
                const double s0 = fma(cw, s1, (double)x[i] - s2);
                s2 = s1;
                s1 = s0;
            }
            const double pw = fma(s1, s1, s2 * s2) - cw * s1 * s2;
            sp += pw;
            sfp = fma((double)k / M, pw, sfp);
        }
        sp = wave_sum(sp);
        sfp = wave_sum(sfp);
        const double mf = sfp / (sp + 1e-10);
        int64_t tsep = (int64_t)ceil(1.0 / (mf + 1e-10));
        const int64_t tmax = (int64_t)(0.25 * (double)w);
        tsep = tsep < 1 ? 1 : tsep;
        tsep = tsep > tmax ? tmax : tsep;
        // 2. nearest neighbours among the first ntraj vectors
        const int64_t ntraj = m - kLrTraj + 1 > 1 ? m - kLrTraj + 1 : 1;
        double lsum[kLrTraj], lcnt[kLrTraj];
#pragma unroll
        for (int k = 0; k < kLrTraj; k++) lsum[k] = lcnt[k] = 0.0;
        for (int64_t i0 = 0; i0 < ntraj; i0 += 64) {
            const int64_t i = i0 + lane;
            const bool row = i < ntraj;
            double a[kLrEmb];
#pragma unroll
            for (int e = 0; e < kLrEmb; e++) a[e] = row ? (double)x[i + e] : 0.0;
            double best = INFINITY;
            int nb = 0;
            // eight rows j at a time: their 17 samples loaded together (LDS
            // broadcasts), the distances from registers; |i - j| > tsep as
            // one unsigned comparison (window rows fit 32 bits)
            const int ii = (int)i, nt = (int)ntraj, ts = (int)tsep;
            for (int j0 = 0; j0 < nt; j0 += 8) {
                double v[kLrEmb + 7];
#pragma unroll
                for (int u = 0; u < kLrEmb + 7; u++) v[u] = j0 + u < w ? (double)x[j0 + u] : 0.0;
                double dd[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {  // independent chains, two partial sums each
                    double s0 = 0.0, s1 = 0.0;
#pragma unroll
                    for (int e = 0; e < kLrEmb; e += 2) {
                        const double d0 = a[e] - v[u + e], d1 = a[e + 1] - v[u + e + 1];
                        s0 = fma(d0, d0, s0);
                        s1 = fma(d1, d1, s1);
                    }
                    dd[u] = s0 + s1;
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int j = j0 + u;
                    const double d2 = dd[u];
                    const bool far = (unsigned)(j - ii + ts) > (unsigned)(2 * ts);
                    if (j < nt && far && d2 < best) {
                        best = d2;
                        nb = j;
                    }
                }
            }
            // 3. the divergence of the pair over the next 20 steps
            if (row) {
#pragma unroll
                for (int k = 0; k < kLrTraj; k++) {
                    const int64_t ia = i + k, ib = nb + k;
                    if (ia >= m || ib >= m) continue;  // past the last vector
                    const int64_t sep = ia > ib ? ia - ib : ib - ia;
                    double d2 = 0.0;
                    for (int e = 0; e < kLrEmb; e++) {
                        const double df = (double)x[ia + e] - (double)x[ib + e];
                        d2 = fma(df, df, d2);
                    }
                    const double d = sep <= tsep ? INFINITY : sqrt(d2);
                    if (d > 0.0) {
                        lsum[k] += log(d);
                        lcnt[k] += 1.0;
                    }
                }
            }
        }
        double n_f = 0.0, sk = 0.0, sd = 0.0, skk = 0.0, skd = 0.0;
#pragma unroll
        for (int k = 0; k < kLrTraj; k++) {
            const double c = wave_sum(lcnt[k]);
            const double sm = wave_sum(lsum[k]);
            const double dk = c > 0.0 ? sm / c : NAN;
            if (isfinite(dk)) {
                n_f += 1.0;
                sk += (double)k;
                sd += dk;
                skk += (double)(k * k);
                skd = fma((double)k, dk, skd);
            }
        }
        const double den = n_f * skk - sk * sk;
        if (lane == 0) out[p] = (n_f >= 2.0 && fabs(den) > 1e-10) ? (n_f * skd - sk * sd) / den : NAN;
    }
}

constexpr int kLeEmb = 10, kLeDim = 4, kLeStep = 3, kLeNb = 8;
constexpr int kLeRec = 14;  // the normal equations of one point: 10 of X^T X, 4 of X^T b
constexpr uint64_t kLeKeyMask = 0x1FFF;  // 13 bits of point index: windows below 8,204 rows

// The search key of orbit point j at Chebyshev distance d (0 <= d < 2^1023):
// d's bits with the exponent raised by one (no key is subnormal, so min and
// max order keys whatever the denormal mode) and the low 13 bits replaced by
// j.  Keys of non-negative doubles order as their bit patterns, so keys order
// as (d truncated to 39 fraction bits, j): a key below another has a distance
// no larger, and a distance part strictly below means a smaller distance.
__device__ __forceinline__ double le_key(double d, int j) {
    const uint64_t b = (uint64_t)__double_as_longlong(d) + (1ull << 52);
    return __longlong_as_double((long long)((b & ~kLeKeyMask) | (uint64_t)j));
}

// Solves the 4 x 4 system A a = b (A symmetric, given by its upper triangle
// u[0..9] row by row, plus reg on the diagonal) by Gaussian elimination with
// partial pivoting; a is written to a[0..3].
__device__ __forceinline__ void solve4(const double *u, const double *b, double reg, double *a) {
    double M[4][5];
    int t = 0;
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = r; c < 4; c++, t++) {
            M[r][c] = u[t] + (r == c ? reg : 0.0);
            M[c][r] = M[r][c];
        }
#pragma unroll
    for (int r = 0; r < 4; r++) M[r][4] = b[r];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int piv = k;
        double big = fabs(M[k][k]);
#pragma unroll
        for (int r = k + 1; r < 4; r++)
            if (fabs(M[r][k]) > big) {
                big = fabs(M[r][k]);
                piv = r;
            }
#pragma unroll
        for (int r = k + 1; r < 4; r++)
            if (r == piv) {
#pragma unroll
                for (int c = 0; c < 5; c++) {
                    const double h = M[k][c];
                    M[k][c] = M[r][c];
                    M[r][c] = h;
                }
            }
#pragma unroll
        for (int r = k + 1; r < 4; r++) {
            const double f = M[r][k] / M[k][k];
#pragma unroll
            for (int c = k; c < 5; c++) M[r][c] = fma(-f, M[k][c], M[r][c]);
        }
    }
#pragma unroll
    for (int r = 3; r >= 0; r--) {
        double v = M[r][4];
#pragma unroll
        for (int c = r + 1; c < 4; c++) v = fma(-M[r][c], a[c], v);
        a[r] = v / M[r][r];
    }
}

// lyapunov_e (torch_lyapunov_e.py:15-150; emb_dim 10, matrix_dim 4, min_nb 8,
// min_tsep 0; the caller divides by tau): Eckmann's method.  The orbit is the
// 10-sample delay embedding of x[0 .. w - 4] (m = 3, n_orb = w - 12 points);
// for every point i its neighbours are the points within the 8th smallest
// Chebyshev distance to it (ties included, itself left out); the least-squares
// map of the neighbours' offsets x[j + 3k] - x[i + 3k] (k = 0..3) to
// x[j + 12] - x[i + 12] (normal equations + 1e-10 I) gives the last row of a
// companion matrix T_i; the running Q of (T_i Q) = Q R (diag R made positive)
// accumulates ln diag R where positive; the exponents are the means over the
// points, over 3 (inf where nothing was accumulated).  One wave per series in
// LDS: the points' neighbour searches and normal equations run 64 at a time
// (one point per lane, the nine smallest distances kept in registers as
// keys that carry their points, le_key), their
// 14 sums go through the wave's LDS rows, and the sequential QR chain runs
// over them in every lane alike.  out (4, n, dim) in descending accumulation
// order as the reference leaves it; a non-finite window gives inf.
template <class T>
__global__ void __launch_bounds__(kSeriesBlock)
    lyap_e_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                  double *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    double *recs = reinterpret_cast<double *>(smem + ((int64_t)G * w * sizeof(T) + 15) / 16 * 16);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *rec = recs + wave * 64 * (kLeRec + 1);
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        const T *x = lds + (int64_t)g * w;
        bool bad = false, big = false;
        for (int64_t i = lane; i < w; i += 64) {
            bad |= !isfinite((double)x[i]);
            big |= fabs((double)x[i]) >= 0x1p1021;
        }
        if (__any(bad)) {
            if (lane < kLeDim) out[lane * np + p] = INFINITY;
            continue;
        }
        const bool huge = __any(big);  // distances may reach 2^1022: no finite keys
        const int no = (int)(w - kLeStep) - (kLeEmb - 1);  // orbit points
        double Q[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
        double lexp[4] = {0, 0, 0, 0}, lcnt[4] = {0, 0, 0, 0};
        for (int i0 = 0; i0 < no; i0 += 64) {
            const int i = i0 + lane;
            double sums[kLeRec];
#pragma unroll
            for (int k = 0; k < kLeRec; k++) sums[k] = 0.0;
            bool ok = false;
            if (i < no) {
                double a[kLeEmb];
#pragma unroll
                for (int e = 0; e < kLeEmb; e++) a[e] = (double)x[i + e];
                // the nine smallest Chebyshev distances as keys (le_key): one
                // branch-free min/max network keeps distances and points
                // together; a window with |x| >= 2^1021 makes no finite key
                // and takes the exact search below
                double top[kLeNb + 1];
#pragma unroll
                for (int k = 0; k <= kLeNb; k++) top[k] = INFINITY;
                if (!huge) {
                  // eight points j at a time: their 17 samples loaded together
                  // (LDS broadcasts), the distances from registers
                  for (int j0 = 0; j0 < no; j0 += 8) {
                    double xv[kLeEmb + 7];
#pragma unroll
                    for (int u = 0; u < kLeEmb + 7; u++) xv[u] = j0 + u < w ? (double)x[j0 + u] : 0.0;
                    double dd[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {  // independent chains, a balanced max tree
                      double t[kLeEmb];
#pragma unroll
                      for (int e = 0; e < kLeEmb; e++) t[e] = fabs(a[e] - xv[u + e]);
                      dd[u] = fmax(fmax(fmax(fmax(t[0], t[1]), fmax(t[2], t[3])),
                                        fmax(fmax(t[4], t[5]), fmax(t[6], t[7]))), fmax(t[8], t[9]));
                    }
                    // one branch per block of eight: a key at or above the
                    // ninth passes through the network unchanged
                    double kv[8];
                    double kmin = INFINITY;
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                      const int j = j0 + u;
                      kv[u] = (j < no && j != i) ? le_key(dd[u], j) : INFINITY;
                      kmin = fmin(kmin, kv[u]);
                    }
                    if (kmin < top[kLeNb]) {
#pragma unroll
                      for (int u = 0; u < 8; u++) {
                        double v = kv[u];
#pragma unroll
                        for (int k = 0; k <= kLeNb; k++) {
                          const double lo = fmin(top[k], v);
                          v = fmax(top[k], v);
                          top[k] = lo;
                        }
                      }
                    }
                  }
                }
                // The keys settle the neighbours when the ninth key's distance
                // part lies above the eighth's: the ninth distance then lies
                // above the eighth, and the neighbours are the eight keyed
                // points.  Otherwise (a tie at the eighth distance, or fewer
                // than eight points) the exact search decides.
                const uint64_t k8 = (uint64_t)__double_as_longlong(top[kLeNb - 1]) & ~kLeKeyMask;
                const uint64_t k9 = (uint64_t)__double_as_longlong(top[kLeNb]) & ~kLeKeyMask;
                double nd[kLeNb], r = INFINITY;
                int nj[kLeNb];
                bool tie = false;
                if (!huge && k9 > k8) {
#pragma unroll
                    for (int k = 0; k < kLeNb; k++) {
                        nj[k] = (int)((uint64_t)__double_as_longlong(top[k]) & kLeKeyMask);
                        double d = 0.0;
#pragma unroll
                        for (int e = 0; e < kLeEmb; e++) d = fmax(d, fabs(a[e] - (double)x[nj[k] + e]));
                        nd[k] = d;
                    }
                    r = nd[kLeNb - 1];
                } else {
                    // exact search: the nine smallest distances (ascending,
                    // ties by point) and their points; the neighbours are the
                    // first eight when the ninth lies above the eighth, else
                    // every point within the eighth distance
                    double td[kLeNb + 1];
                    int tj[kLeNb + 1];
#pragma unroll
                    for (int k = 0; k <= kLeNb; k++) {
                        td[k] = INFINITY;
                        tj[k] = -1;
                    }
                    for (int j = 0; j < no; j++) {
                        double d = 0.0;
#pragma unroll
                        for (int e = 0; e < kLeEmb; e++) d = fmax(d, fabs(a[e] - (double)x[j + e]));
                        if (j != i && d < td[kLeNb]) {
                            double v = d;
                            int vj = j;
#pragma unroll
                            for (int k = 0; k <= kLeNb; k++) {
                                const bool sw = v < td[k];
                                const double hv = sw ? td[k] : v;
                                const int hj = sw ? tj[k] : vj;
                                td[k] = sw ? v : td[k];
                                tj[k] = sw ? vj : tj[k];
                                v = hv;
                                vj = hj;
                            }
                        }
                    }
                    r = td[kLeNb - 1];
                    tie = !(td[kLeNb] > r);
#pragma unroll
                    for (int k = 0; k < kLeNb; k++) {
                        nd[k] = td[k];
                        nj[k] = tj[k];
                    }
                }
                // the eight neighbours in the order of (distance, point), on
                // either path (keys order by truncated distances, and the
                // exact insertion can carry an element past an equal one)
                if (!tie) {
#pragma unroll
                    for (int k = 1; k < kLeNb; k++)
#pragma unroll
                        for (int m = k; m > 0; m--) {
                            const bool sw = nd[m] < nd[m - 1] || (nd[m] == nd[m - 1] && nj[m] < nj[m - 1]);
                            const double hd = nd[m];
                            const int hj = nj[m];
                            nd[m] = sw ? nd[m - 1] : nd[m];
                            nj[m] = sw ? nj[m - 1] : nj[m];
                            nd[m - 1] = sw ? hd : nd[m - 1];
                            nj[m - 1] = sw ? hj : nj[m - 1];
                        }
                }
                if (r < INFINITY) {
                    ok = true;
                    double xi[kLeDim];
#pragma unroll
                    for (int k = 0; k < kLeDim; k++) xi[k] = (double)x[i + kLeStep * k];
                    const double bi = (double)x[i + kLeStep * kLeDim];
                    auto add = [&](int j) {
                        double xd[kLeDim];
#pragma unroll
                        for (int k = 0; k < kLeDim; k++) xd[k] = (double)x[j + kLeStep * k] - xi[k];
                        const double bd = (double)x[j + kLeStep * kLeDim] - bi;
                        int t = 0;
#pragma unroll
                        for (int rr = 0; rr < kLeDim; rr++)
#pragma unroll
                            for (int c = rr; c < kLeDim; c++, t++) sums[t] = fma(xd[rr], xd[c], sums[t]);
#pragma unroll
                        for (int k = 0; k < kLeDim; k++) sums[10 + k] = fma(xd[k], bd, sums[10 + k]);
                    };
                    if (!tie) {
#pragma unroll
                        for (int k = 0; k < kLeNb; k++) add(nj[k]);
                    } else {
                        for (int j = 0; j < no; j++) {
                            if (j == i) continue;
                            double d = 0.0;
#pragma unroll
                            for (int e = 0; e < kLeEmb; e++) d = fmax(d, fabs(a[e] - (double)x[j + e]));
                            if (d <= r) add(j);
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < kLeRec; k++) rec[lane * (kLeRec + 1) + k] = sums[k];
            rec[lane * (kLeRec + 1) + kLeRec] = ok ? 1.0 : 0.0;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            // the QR chain over this chunk's points, in order (every lane alike)
            const int cnt = no - i0 < 64 ? no - i0 : 64;
            for (int q = 0; q < cnt; q++) {
                const double *rq = rec + q * (kLeRec + 1);
                if (rq[kLeRec] == 0.0) continue;
                double co[4];
                solve4(rq, rq + 10, 1e-10, co);
                // M = T Q: rows 0..2 of T shift (row r picks Q's row r + 1), row 3 is co
                double Mt[4][4];
#pragma unroll
                for (int c = 0; c < 4; c++) {
#pragma unroll
                    for (int r = 0; r < 3; r++) Mt[r][c] = Q[r + 1][c];
                    Mt[3][c] = fma(co[0], Q[0][c], fma(co[1], Q[1][c], fma(co[2], Q[2][c], co[3] * Q[3][c])));
                }
                // Householder QR of Mt: R in Mt's upper triangle, Q accumulated
                double Qn[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    double nrm = 0.0;
#pragma unroll
                    for (int r = k; r < 4; r++) nrm = fma(Mt[r][k], Mt[r][k], nrm);
                    nrm = sqrt(nrm);
                    if (nrm == 0.0) continue;
                    const double alpha = Mt[k][k] > 0.0 ? -nrm : nrm;
                    double v[4] = {0, 0, 0, 0};
#pragma unroll
                    for (int r = k; r < 4; r++) v[r] = Mt[r][k];
                    v[k] -= alpha;
                    double vv = 0.0;
#pragma unroll
                    for (int r = k; r < 4; r++) vv = fma(v[r], v[r], vv);
                    if (vv == 0.0) continue;
#pragma unroll
                    for (int c = 0; c < 4; c++) {  // Mt = (I - 2 v v^T / vv) Mt
                        double dot = 0.0;
#pragma unroll
                        for (int r = k; r < 4; r++) dot = fma(v[r], Mt[r][c], dot);
                        const double f = 2.0 * dot / vv;
#pragma unroll
                        for (int r = k; r < 4; r++) Mt[r][c] = fma(-f, v[r], Mt[r][c]);
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {  // Qn = Qn (I - 2 v v^T / vv)
                        double dot = 0.0;
#pragma unroll
                        for (int c = k; c < 4; c++) dot = fma(Qn[r][c], v[c], dot);
                        const double f = 2.0 * dot / vv;
#pragma unroll
                        for (int c = k; c < 4; c++) Qn[r][c] = fma(-f, v[c], Qn[r][c]);
                    }
                }
                // positive diagonal: flip the signs of R's rows and Q's columns
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const double sg = Mt[k][k] < 0.0 ? -1.0 : 1.0;
                    const double dg = sg * Mt[k][k];
#pragma unroll
                    for (int r = 0; r < 4; r++) Q[r][k] = sg * Qn[r][k];
                    if (dg > 0.0) {
                        lexp[k] += log(dg);
                        lcnt[k] += 1.0;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();  // the records are rewritten by the next chunk
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
        if (lane < kLeDim) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < kLeDim; k++)
                if (k == lane) v = lcnt[k] > 0.0 ? lexp[k] / lcnt[k] / (double)kLeStep : INFINITY;
            out[lane * np + p] = v;
        }
    }
}

// ---- Lempel-Ziv phrase counts ----------------------------------------------

constexpr int kLzMaxBins = 8;  // bins settings per launch

// lempel_ziv_complexity (torch_features_entropy_complexity.py:140-180): the
// window is cut into symbols b_i = clamp(floor((x_i - min) / (max - min +
// 1e-10) * bins), 0, bins - 1) (in the series dtype), then parsed left to
// right into phrases, each the shortest string at its position not seen
// before; a tail whose every prefix was seen is not counted.  The phrases
// form a trie (a new phrase extends a seen one by a symbol), kept as an
// open-addressing hash table (parent * bins + symbol + 1 in the upper 20
// bits, the node in the lower 12; 2^cap_log >= 2 w entries, so at most half
// full and every probe ends) in LDS.  One wave per block, G series staged
// together, lane k parsing each for bins setting k (nb settings at most 8).
// out (nb, n, dim): the phrase count over w / (ln w + 1e-10) (over 1 for
// w = 1); NaN for a non-finite window.
template <class T>
__global__ void __launch_bounds__(64)
    lempel_ziv_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                      int cap_log, int4 bins_a, int4 bins_b, int nb, double *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    unsigned *tabs = reinterpret_cast<unsigned *>(smem + ((int64_t)G * w * sizeof(T) + 15) / 16 * 16);
    const int lane = threadIdx.x;
    const unsigned cap = 1u << cap_log;
    unsigned char *syms = reinterpret_cast<unsigned char *>(tabs + (int64_t)nb * cap);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int binsv[kLzMaxBins] = {bins_a.x, bins_a.y, bins_a.z, bins_a.w,
                                   bins_b.x, bins_b.y, bins_b.z, bins_b.w};
    int bins = 2;
#pragma unroll
    for (int k = 0; k < kLzMaxBins; k++)
        if (k == lane) bins = binsv[k];
    unsigned *tab = tabs + (lane < nb ? lane : 0) * cap;
    const unsigned char *sy = syms + (int64_t)(lane < nb ? lane : 0) * w;
    const double bn = w > 1 ? (double)w / (log((double)w) + 1e-10) : 1.0;
    for (int g = 0; g < G; g++) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // block-uniform
        const T *x = lds + (int64_t)g * w;
        // minimum, maximum and finiteness by the whole wave
        double mnd = INFINITY, mxd = -INFINITY;
        bool bad = false;
        for (int64_t i = lane; i < w; i += 64) {
            const double v = (double)x[i];
            bad |= !isfinite(v);
            mnd = fmin(mnd, v);
            mxd = fmax(mxd, v);
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            mnd = fmin(mnd, __shfl_xor(mnd, m, 64));
            mxd = fmax(mxd, __shfl_xor(mxd, m, 64));
        }
        if (__any(bad)) {
            if (lane < nb) out[lane * np + p] = NAN;
            continue;
        }
        // the symbols of every setting (one byte each), by the whole wave,
        // and the hash tables cleared
        const T mn = (T)mnd, rng = (T)mxd - mn + (T)1e-10;
        for (int64_t i = lane; i < w; i += 64) {
            const T v = x[i] - mn;
#pragma unroll
            for (int k = 0; k < kLzMaxBins; k++) {
                if (k >= nb) break;
                const int bk = binsv[k];
                int sb = (int)(v / rng * (T)bk);
                sb = sb < 0 ? 0 : (sb > bk - 1 ? bk - 1 : sb);
                syms[(int64_t)k * w + i] = (unsigned char)sb;
            }
        }
        for (int64_t t = lane; t < (int64_t)nb * cap; t += 64) tabs[t] = 0u;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane < nb) {
            int count = 0, nodes = 1;  // node 0 is the root
            int64_t i = 0;
            while (i < w) {
                int cur = 0;
                int64_t len = 1;
                bool added = false;
                unsigned sym = sy[i];
                while (i + len <= w) {
                    // the next symbol is read while this one's probe runs
                    const unsigned nsym = i + len < w ? sy[i + len] : 0u;
                    const unsigned key = (unsigned)cur * (unsigned)bins + sym + 1u;
                    unsigned h = (key * 2654435761u) >> (32 - cap_log);
                    int child = -1;
                    for (;;) {
                        const unsigned e = tab[h];
                        if (e == 0u) break;
                        if ((e >> 12) == key) {
                            child = (int)(e & 0xfffu);
                            break;
                        }
                        h = (h + 1u) & (cap - 1u);
                    }
                    if (child < 0) {  // a new phrase: add it as cur's child
                        tab[h] = (key << 12) | (unsigned)nodes;
                        nodes++;
                        count++;
                        added = true;
                        break;
                    }
                    cur = child;
                    len++;
                    sym = nsym;
                }
                i += len;
                if (!added) break;
            }
            out[lane * np + p] = (double)count / bn;
        }
        __builtin_amdgcn_wave_barrier();  // symbols and tables are rebuilt for the next series
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// ---- sorted-window statistics: duplicates and change quantiles -------------

// the six quantile levels whose corridors change_quantiles takes
// (ql in 0 .. 0.8, qh in 0.2 .. 1, ql < qh; settings.py:315-322)
__constant__ double kCqLevel[6] = {0.0, 0.2, 0.4, 0.6, 0.8, 1.0};
constexpr int kCqLevels = 6;
constexpr int kCqCorr = 15;

// torch.quantile's linear interpolation on a sorted series s of w values, in
// the series dtype: rank = q * (w - 1) with q in that dtype, below = floor,
// above = ceil, weight = rank - below, then lerp as a fused multiply-add
// (weight < 0.5: below + weight (above - below); else above - (above -
// below)(1 - weight)).
template <class T>
__device__ __forceinline__ T torch_quantile(const T *s, int64_t w, double q) {
    const T rank = (T)q * (T)(w - 1);
    const int64_t lo = (int64_t)rank;
    const int64_t hi = (int64_t)ceil((double)rank);
    const T wgt = rank - (T)lo;
    const T a = s[lo], b = s[hi < w ? hi : w - 1];
    const T d = b - a;
    return wgt < (T)0.5 ? fma(wgt, d, a) : fma(-d, (T)1 - wgt, b);
}

// Ascending bitonic sort of the P = 64 R values held R per lane, element e =
// lane * R + r in v[r], by one wave: a compare-exchange stage (k, j) pairs e
// with e ^ j and keeps the smaller value at the lower index where e & k is
// clear (the larger where it is set).  For j < R both elements sit in the
// lane's registers; for j >= R the partner is the same register of lane
// lane ^ (j / R), exchanged by a cross-lane shuffle.  All loop bounds are
// compile-time, so every register index is.
template <class T, int R>
__device__ __forceinline__ void wave_bitonic_sort(T (&v)[R], int lane) {
    constexpr int P = 64 * R;
#pragma unroll
    for (int k = 2; k <= P; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= R) {
                const int m = j / R;
                const bool lower = (lane & m) == 0;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const bool asc = (((lane * R + r) & k) == 0);
                    const T o = __shfl_xor(v[r], m, 64);
                    v[r] = (lower == asc) ? (o < v[r] ? o : v[r]) : (o > v[r] ? o : v[r]);
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if (r & j) continue;
                    const bool asc = (((lane * R + r) & k) == 0);
                    const T a = v[r], b = v[r + j];
                    const T lo = a < b ? a : b, hi = a < b ? b : a;
                    v[r] = asc ? lo : hi;
                    v[r + j] = asc ? hi : lo;
                }
            }
        }
    }
}

// One wave per series: the series (staged in LDS with the block's others) is
// taken into registers, R values per lane, padded with +inf to P = 64 R (a
// power of two), sorted (wave_bitonic_sort) and written to the wave's sort
// row.  On the sorted values: out_dup (5, n, dim) =
// the adjacent equal pairs D, the values in groups of equal values R, the
// groups G, the sum of the values in groups and the sum of one value per group
// (has_duplicate, ratio_value_number_to_time_series_length and the
// reoccurrence counts, torch_features_boolean.py:11-19,
// torch_features_reocurrance.py:11-161), sums in double; out_levels (6, n,
// dim) in the series dtype, or null: torch.quantile at the levels of
// kCqLevel (the corridor bounds of change_quantiles); out_median (n, dim) in
// the series dtype, or null: the lower middle value (torch.median).  A
// window holding a non-finite value gives NaN throughout.
template <class T, int R>
__global__ void __launch_bounds__(kSeriesBlock)
    sorted_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int G,
                  double *out_dup, T *out_levels, T *out_median) {
    constexpr int P = 64 * R;
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    T *lds = reinterpret_cast<T *>(smem);
    T *sortbase = reinterpret_cast<T *>(smem + ((int64_t)G * w * sizeof(T) + 15) / 16 * 16);
    load_series<T>(traj, n, S, start, w, p0, G, np, lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    T *buf = sortbase + (int64_t)wave * P;
    for (int g = wave; g < G; g += kSeriesBlock / 64) {
        const int64_t p = p0 + g;
        if (p >= np) break;  // wave-uniform
        const T *x = lds + (int64_t)g * w;
        bool bad = false;
        for (int64_t i = lane; i < w; i += 64) bad |= !isfinite((double)x[i]);
        if (__any(bad)) {
            if (lane < 5) out_dup[lane * np + p] = NAN;
            if (out_levels && lane < kCqLevels) out_levels[lane * np + p] = (T)NAN;
            if (out_median && lane == 0) out_median[p] = (T)NAN;
            continue;
        }
        T v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int64_t e = (int64_t)lane * R + r;
            v[r] = e < w ? x[e] : (T)INFINITY;
        }
        wave_bitonic_sort<T, R>(v, lane);
#pragma unroll
        for (int r = 0; r < R; r++) buf[lane * R + r] = v[r];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        double nd = 0.0, nr = 0.0, ng = 0.0, sp = 0.0, sv = 0.0;
        for (int64_t i = lane; i < w; i += 64) {
            const T v = buf[i];
            const bool prev = i > 0 && buf[i - 1] == v;
            const bool next = i + 1 < w && buf[i + 1] == v;
            nd += prev ? 1.0 : 0.0;
            if (prev || next) {
                nr += 1.0;
                sp += (double)v;
            }
            if (next && !prev) {
                ng += 1.0;
                sv += (double)v;
            }
        }
        nd = wave_sum(nd);
        nr = wave_sum(nr);
        ng = wave_sum(ng);
        sp = wave_sum(sp);
        sv = wave_sum(sv);
        if (lane == 0) {
            out_dup[0 * np + p] = nd;
            out_dup[1 * np + p] = nr;
            out_dup[2 * np + p] = ng;
            out_dup[3 * np + p] = sp;
            out_dup[4 * np + p] = sv;
        }
        if (out_levels && lane < kCqLevels) out_levels[lane * np + p] = torch_quantile<T>(buf, w, kCqLevel[lane]);
        if (out_median && lane == 0) out_median[p] = buf[(w - 1) / 2];
        __builtin_amdgcn_wave_barrier();  // buf reused by the wave's next series
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// The same statistics for windows too long for one wave's registers: one
// series per block of 256 threads, padded with +inf to P (a power of two,
// P * sizeof(T) <= 32 KB) and sorted in LDS by a block-wide bitonic network
// (each stage's pairs loaded by every thread before any is stored; a block
// barrier between stages); the sums by a block reduction through LDS.
template <class T>
__global__ void __launch_bounds__(256)
    sorted_block_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start, int P,
                        double *out_dup, T *out_levels, T *out_median) {
    extern __shared__ __align__(16) unsigned char smem[];
    T *buf = reinterpret_cast<T *>(smem);
    double *red = reinterpret_cast<double *>(smem + (int64_t)P * sizeof(T));  // 4 waves x 5
    const int64_t w = n_save - start;
    const int64_t np = n * S;
    const int64_t p = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    bool bad = false;
    for (int64_t i = tid; i < P; i += 256) {
        T v = (T)INFINITY;
        if (i < w) {
            v = traj[(start + i) * np + p];
            bad |= !isfinite((double)v);
        }
        buf[i] = v;
    }
    const int anybad = __syncthreads_or(bad ? 1 : 0);
    if (anybad) {
        if (tid < 5) out_dup[tid * np + p] = NAN;
        if (out_levels && tid < kCqLevels) out_levels[tid * np + p] = (T)NAN;
        if (out_median && tid == 0) out_median[p] = (T)NAN;
        return;
    }
    constexpr int kMaxPairs = 16;  // P / 2 / 256 for P <= 8192
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            T a[kMaxPairs], b[kMaxPairs];
            int idx[kMaxPairs];
#pragma unroll
            for (int r = 0; r < kMaxPairs; r++) {
                const int t = tid + 256 * r;
                idx[r] = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                if (t < (P >> 1)) {
                    a[r] = buf[idx[r]];
                    b[r] = buf[idx[r] + j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kMaxPairs; r++) {
                const int t = tid + 256 * r;
                if (t < (P >> 1)) {
                    const bool asc = (idx[r] & k) == 0;
                    const T lo = a[r] < b[r] ? a[r] : b[r], hi = a[r] < b[r] ? b[r] : a[r];
                    buf[idx[r]] = asc ? lo : hi;
                    buf[idx[r] + j] = asc ? hi : lo;
                }
            }
            __syncthreads();
        }
    }
    double v5[5] = {0, 0, 0, 0, 0};  // D, R, G, sum in groups, sum of one per group
    for (int64_t i = tid; i < w; i += 256) {
        const T v = buf[i];
        const bool prev = i > 0 && buf[i - 1] == v;
        const bool next = i + 1 < w && buf[i + 1] == v;
        v5[0] += prev ? 1.0 : 0.0;
        if (prev || next) {
            v5[1] += 1.0;
            v5[3] += (double)v;
        }
        if (next && !prev) {
            v5[2] += 1.0;
            v5[4] += (double)v;
        }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) {
        v5[k] = wave_sum(v5[k]);
        if (lane == 0) red[wave * 5 + k] = v5[k];
    }
    __syncthreads();
    if (tid < 5) out_dup[tid * np + p] = red[tid] + red[5 + tid] + red[10 + tid] + red[15 + tid];
    if (out_levels && tid < kCqLevels) out_levels[tid * np + p] = torch_quantile<T>(buf, w, kCqLevel[tid]);
    if (out_median && tid == 0) out_median[p] = buf[(w - 1) / 2];
}

// change_quantiles (torch_features_change.py:20-139) for the 15 corridors
// [v_a, v_b], a < b, of the six levels of kCqLevel, one lane per series: a
// change d_i = x_{i+1} - x_i counts in a corridor when both samples lie in it
// (compared in the series dtype with the levels of bh_window_sorted).  Pass 1
// counts and sums d and |d| per corridor; pass 2 sums the squared deviations
// from the corridor means.  out (15, 4, n, dim): mean, population variance of
// d, then of |d|, 0 for a corridor no change falls in; NaN where a level is
// NaN (a non-finite window).  Sums in double.
template <class T>
__global__ void __launch_bounds__(256)
    change_quantiles_kernel(const T *traj, int64_t n_save, int64_t n, int S, int64_t start,
                            const T *levels, double *out) {
    const int64_t np = n * S;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    T lv[kCqLevels];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < kCqLevels; j++) {
        lv[j] = levels[j * np + p];
        ok &= !isnan((double)lv[j]);
    }
    if (!ok) {
        for (int k = 0; k < kCqCorr * 4; k++) out[k * np + p] = NAN;
        return;
    }
    const T *col = traj + start * np + p;
    const int64_t w = n_save - start;
    double cnt[kCqCorr], s1[kCqCorr], sa[kCqCorr];
#pragma unroll
    for (int c = 0; c < kCqCorr; c++) cnt[c] = s1[c] = sa[c] = 0.0;
    T u = col[0];
    for (int64_t i = 1; i < w; i++) {
        const T v = col[i * np];
        const double d = (double)v - (double)u, ad = fabs(d);
        const T lo = u < v ? u : v, hi = u < v ? v : u;
        u = v;
        int c = 0;
#pragma unroll
        for (int a = 0; a < kCqLevels - 1; a++) {
            const bool ge = lo >= lv[a];
#pragma unroll
            for (int b = a + 1; b < kCqLevels; b++, c++) {
                const bool in = ge && hi <= lv[b];
                cnt[c] += in ? 1.0 : 0.0;
                s1[c] += in ? d : 0.0;
                sa[c] += in ? ad : 0.0;
            }
        }
    }
    double vd[kCqCorr], va[kCqCorr];
#pragma unroll
    for (int c = 0; c < kCqCorr; c++) {
        s1[c] /= cnt[c] + 1e-10;  // the corridor means
        sa[c] /= cnt[c] + 1e-10;
        vd[c] = va[c] = 0.0;
    }
    u = col[0];
    for (int64_t i = 1; i < w; i++) {
        const T v = col[i * np];
        const double d = (double)v - (double)u, ad = fabs(d);
        const T lo = u < v ? u : v, hi = u < v ? v : u;
        u = v;
        int c = 0;
#pragma unroll
        for (int a = 0; a < kCqLevels - 1; a++) {
            const bool ge = lo >= lv[a];
#pragma unroll
            for (int b = a + 1; b < kCqLevels; b++, c++) {
                const bool in = ge && hi <= lv[b];
                const double e = d - s1[c], ea = ad - sa[c];
                vd[c] += in ? e * e : 0.0;
                va[c] += in ? ea * ea : 0.0;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < kCqCorr; c++) {
        const bool any = cnt[c] > 0.0;
        out[(c * 4 + 0) * np + p] = any ? s1[c] : 0.0;
        out[(c * 4 + 1) * np + p] = any ? vd[c] / (cnt[c] + 1e-10) : 0.0;
        out[(c * 4 + 2) * np + p] = any ? sa[c] : 0.0;
        out[(c * 4 + 3) * np + p] = any ? va[c] / (cnt[c] + 1e-10) : 0.0;
    }
}

}  // namespace win
}  // namespace bh
