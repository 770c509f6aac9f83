// rossler_tile.hip -- Dormand-Prince 5(4) for a network of 100 diffusively
// coupled Rossler oscillators (state dim 300), 16 trajectories per workgroup.
//
// Right-hand side (case_studies/rossler_network/rossler_network_jax_ode.py:63-79):
//   x_i' = -y_i - z_i - K (L x)_i,  y_i' = x_i + a y_i,  z_i' = b + z_i (x_i - c)
// with L the 100 x 100 graph Laplacian (rossler_network_topology.py:64-77).
// The coupling for the 16 trajectories of a tile is one (16 x 100) . (100 x 100)
// product per stage on the matrix cores (v_mfma_f32_16x16x4f32, FP32 in and
// out): seven waves each own a 16-column block of L for the first 22 of its
// 25 k-blocks, the eighth wave the last 3 k-blocks of all seven blocks, so
// every SIMD runs the same 43-44 products per stage; the B operands stay in
// registers for the whole launch, the A operand comes from the stage inputs
// staged in LDS, and the two partial products meet in the consumer.
//
// Step control is per trajectory with the torchdiffeq rules used by the
// per-lane kernel (error RMS over the 300 components with the quotient by
// the hardware reciprocal, as there; I-controller, Hairer initial step,
// dense output at the last save time without clamping the step).
// Per-trajectory scalars are replicated in the 32 threads that own the
// trajectory; the reductions they need (initial-step norms, error norm,
// max |y| for the divergence stop) go through LDS in a fixed order, so every
// copy takes the same decision.
//
// A slot whose trajectory finishes (end time reached, |y|inf >= stop
// threshold, failure) takes the next trajectory index from a device counter
// and spends one step-iteration on its initial derivative and initial step,
// so bounded and quickly-diverging trajectories share tiles without idling.
//
// Thread layout: tid = m + 16 q, m = slot (0..15), q = node group (0..31);
// the thread owns node positions q, q + 32, q + 64 and q + 96 (real below
// 100), i.e. x, y, z of up to four nodes.
//
// The edge-list form of the coupling (bh_solve_rossler_edges) runs in the
// wavefront kernel of rossler_wave.hip; the graph table and the argument
// block are shared.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/basinhip.h"

namespace bh {

namespace ross {
constexpr int NN = 100;       // nodes
constexpr int SD = 3 * NN;    // state dim
constexpr int TILE = 16;      // trajectories per workgroup
constexpr int THREADS = 512;  // 8 waves
constexpr int NPT = 4;        // node slots per thread (last one real for q < 4)
constexpr int XS = 101;       // LDS row stride (odd: spreads the A-operand reads over banks)
constexpr int KB = NN / 4;    // k-blocks of the 16x16x4 product
constexpr int NB = 7;         // 16-column blocks covering 100 nodes
constexpr int KS = 22;        // k-blocks of each block done by its own wave; the
                              // eighth wave does the last KB - KS of all seven

constexpr int DMAX = 12;      // neighbour slots per node position (edge-list kernel)
constexpr int NPOS = 128;     // node positions: 32 node groups x NPT

enum Mode : int { IDLE = 0, INIT = 1, STEP = 2 };

// Edge-list graph of the coupling (wavefront kernel), built on the host by
// bh_solve_rossler_edges and passed by value with the launch arguments.
struct Graph {
    uint8_t node[NPOS];       // position -> node (positions >= NN unused)
    uint8_t deg[NPOS];        // edge entries of the position's node
    uint8_t nbr[NPOS][DMAX];  // neighbour nodes; unused slots hold NN (the zero cell)
};

struct Args {
    const float *y0;
    const double *params;
    int64_t p_stride;
    int64_t n;
    double t0, t_end;
    double rtol, atol;
    int64_t max_steps;
    uint32_t flags;
    double stop_thr;
    const float *lap;  // (100, 100) row-major (dense kernel)
    Graph g;           // edge-list (wavefront) kernel
    float *fin;        // (n, 300)
    int32_t *status;
    int64_t *nsteps;
    unsigned long long *counter;
};

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Dormand-Prince coefficients (see dopri5_lane.hip)
__device__ constexpr float A21 = 1.0 / 5.0;
__device__ constexpr float A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
__device__ constexpr float A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
__device__ constexpr float A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0,
                           A54 = -212.0 / 729.0;
__device__ constexpr float A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0,
                           A64 = 49.0 / 176.0, A65 = -5103.0 / 18656.0;
__device__ constexpr float B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0,
                           B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
__device__ constexpr float E1 = 35.0 / 384.0 - 1951.0 / 21600.0;
__device__ constexpr float E3 = 500.0 / 1113.0 - 22642.0 / 50085.0;
__device__ constexpr float E4 = 125.0 / 192.0 - 451.0 / 720.0;
__device__ constexpr float E5 = -2187.0 / 6784.0 - -12231.0 / 42400.0;
__device__ constexpr float E6 = 11.0 / 84.0 - 649.0 / 6300.0;
__device__ constexpr float E7 = -1.0 / 60.0;
__device__ constexpr float M1 = 6025192743.0 / 30085553152.0 / 2.0;
__device__ constexpr float M3 = 51252292925.0 / 65400821598.0 / 2.0;
__device__ constexpr float M4 = -2691868925.0 / 45128329728.0 / 2.0;
__device__ constexpr float M5 = 187940372067.0 / 1594534317056.0 / 2.0;
__device__ constexpr float M6 = -1776094331.0 / 19743644256.0 / 2.0;
__device__ constexpr float M7 = 11237099.0 / 235043384.0 / 2.0;

__device__ __forceinline__ double next_dt(double dt, double ratio) {
    if (ratio == 0.0) return dt * 10.0;
    float f = 0.9f * exp2f(-0.2f * log2f((float)ratio));
    const float dfac = ratio < 1.0 ? 1.0f : 0.2f;
    f = fminf(10.0f, fmaxf(f, dfac));
    return dt * (double)f;
}

// Sum (or max) of one value over the 32 threads of each slot: lanes
// m, m+16, m+32, m+48 of a wave by shuffles, the 8 waves through LDS.
template <bool MAX>
__device__ __forceinline__ float slot_reduce(float v, float (*buf)[TILE], int wave, int lane, int m) {
    float o = __shfl_xor(v, 16);
    v = MAX ? fmaxf(v, o) : v + o;
    o = __shfl_xor(v, 32);
    v = MAX ? fmaxf(v, o) : v + o;
    if (lane < TILE) buf[wave][lane] = v;
    __syncthreads();
    float r = buf[0][m];
#pragma unroll
    for (int w = 1; w < THREADS / 64; w++) r = MAX ? fmaxf(r, buf[w][m]) : r + buf[w][m];
    return r;
}

// The sum of one value and the maximum of another over the threads of each
// slot, with one barrier.
__device__ __forceinline__ void slot_sum_max(float &sum, float &mx, float (*bs)[TILE],
                                             float (*bm)[TILE], int wave, int lane, int m) {
    sum += __shfl_xor(sum, 16);
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    sum += __shfl_xor(sum, 32);
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    if (lane < TILE) {
        bs[wave][lane] = sum;
        bm[wave][lane] = mx;
    }
    __syncthreads();
    float rs = bs[0][m], rm = bm[0][m];
#pragma unroll
    for (int w = 1; w < THREADS / 64; w++) {
        rs += bs[w][m];
        rm = fmaxf(rm, bm[w][m]);
    }
    sum = rs;
    mx = rm;
}

__global__ void __launch_bounds__(THREADS) tile_kernel(Args a) {
    __shared__ float Xs[TILE * XS];
    __shared__ float Cs[TILE * XS];
    __shared__ float Cs2[TILE * XS];  // the eighth wave's partial products
    __shared__ float red0[THREADS / 64][TILE];
    __shared__ float red1[THREADS / 64][TILE];
    __shared__ long long slot_ic[TILE];
    __shared__ int slot_new[TILE];  // the slot took a new trajectory at the last refill

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m = tid & 15, q = tid >> 4;
    // node of each position slot (position q + 32 r is node q + 32 r, real
    // below 100)
    bool real[NPT];
    int node[NPT];
#pragma unroll
    for (int r = 0; r < NPT; r++) {
        node[r] = q + 32 * r;
        real[r] = node[r] < NN;
    }

    // B operands: L[4 kb .. 4 kb + 3][16-column block] for this wave's share
    // (columns past 99 are zero)
    float Bop[KS];
    if (wave < NB) {
#pragma unroll
        for (int kb = 0; kb < KS; kb++) {
            const int k = 4 * kb + (lane >> 4), j = 16 * wave + (lane & 15);
            Bop[kb] = (j < NN) ? a.lap[k * NN + j] : 0.0f;
        }
    } else {
        // blocks b = 0..6, k-blocks KS..KB-1: Bop[(KB - KS) * b + t]
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int t = 0; t < KB - KS; t++) {
                const int k = 4 * (KS + t) + (lane >> 4), j = 16 * b + (lane & 15);
                Bop[(KB - KS) * b + t] = (j < NN) ? a.lap[k * NN + j] : 0.0f;
            }
    }
    // per-thread state for its node slots: x, y, z of nodes q + 32 r
    float y[3][NPT], k1[3][NPT], k2[3][NPT], k3[3][NPT], k4[3][NPT], k5[3][NPT], k6[3][NPT],
        k7[3][NPT], ys[3][NPT];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < NPT; r++)
            y[c][r] = k1[c][r] = k2[c][r] = k3[c][r] = k4[c][r] = k5[c][r] = k6[c][r] = k7[c][r] =
                ys[c][r] = 0.0f;
    // per-slot scalars (replicated over the slot's 32 threads)
    long long ic = (long long)blockIdx.x * TILE + m;
    int mode = ic < a.n ? INIT : IDLE;
    double t = a.t0, dt = 0.0;
    float h0 = 0.0f, pa = 0.f, pb = 0.f, pc = 0.f, pK = 0.f;
    int64_t steps = 0;
    const float ratol = (float)a.atol, rrtol = (float)a.rtol;

    auto load_slot = [&](long long idx) {
        const double *pp = a.params + idx * a.p_stride;
        pa = (float)pp[0];
        pb = (float)pp[1];
        pc = (float)pp[2];
        pK = (float)pp[3];
        const float *src = a.y0 + idx * SD;
#pragma unroll
        for (int r = 0; r < NPT; r++) {
#pragma unroll
            for (int c = 0; c < 3; c++) y[c][r] = real[r] ? src[c * NN + node[r]] : 0.0f;
        }
        // the FSAL slot is multiplied by a zero step in the initial evaluations
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < NPT; r++) k1[c][r] = 0.0f;
        t = a.t0;
        dt = 0.0;
        steps = 0;
    };
    if (mode == INIT) load_slot(ic);

    auto write_final = [&](long long idx, const float (&v)[3][NPT]) {
        float *dst = a.fin + idx * SD;
#pragma unroll
        for (int r = 0; r < NPT; r++) {
            if (real[r]) {
#pragma unroll
                for (int c = 0; c < 3; c++) dst[c * NN + node[r]] = v[c][r];
            }
        }
    };

    // one right-hand-side evaluation of ys for all slots -> out
    auto rhs = [&](float (&out)[3][NPT]) {
#pragma unroll
        for (int r = 0; r < NPT; r++)
            if (real[r]) Xs[m * XS + q + 32 * r] = ys[0][r];
        __syncthreads();
        // local terms (no coupling) next to the coupling product
        float part[NPT];
#pragma unroll
        for (int r = 0; r < NPT; r++) {
            const float xx = ys[0][r], yy = ys[1][r], zz = ys[2][r];
            part[r] = -yy - zz;
            out[1][r] = xx + pa * yy;
            out[2][r] = pb + zz * (xx - pc);
        }
        const float *xrow = Xs + (lane & 15) * XS + (lane >> 4);
        if (wave < NB) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < KS; kb++)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xrow[4 * kb], Bop[kb], acc, 0, 0, 0);
            const int col = 16 * wave + (lane & 15);
            if (col < NN) {
#pragma unroll
                for (int it = 0; it < 4; it++) Cs[(4 * (lane >> 4) + it) * XS + col] = acc[it];
            }
        } else {
#pragma unroll
            for (int b = 0; b < NB; b++) {
                floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < KB - KS; t++)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xrow[4 * (KS + t)],
                                                               Bop[(KB - KS) * b + t], acc, 0, 0, 0);
                const int col = 16 * b + (lane & 15);
                if (col < NN) {
#pragma unroll
                    for (int it = 0; it < 4; it++) Cs2[(4 * (lane >> 4) + it) * XS + col] = acc[it];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NPT; r++) {
            const int e = m * XS + q + 32 * r;
            const float cpl = real[r] ? Cs[e] + Cs2[e] : 0.0f;
            out[0][r] = part[r] - pK * cpl;
        }
    };

    // The initial-step norms are needed only in iterations where some slot of
    // the tile starts a trajectory: every slot at the first iteration, later
    // the slots refilled at the end of the previous one.  any_init is the
    // same in every thread, so the reductions it guards are reached by every
    // thread or by none.
    bool any_init = true;
    for (;;) {
        if (!__syncthreads_or(mode != IDLE)) break;
        const bool stepping = mode == STEP, init = mode == INIT;
        const double tB = t + dt;
        const float h = stepping ? (float)dt : 0.0f;

        // ---- evaluation 1: k2 = f(y + a21 h k1); INIT: f0 = f(y0) into k2 ----
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < NPT; r++) ys[c][r] = y[c][r] + k1[c][r] * (A21 * h);
        rhs(k2);
        if (any_init) {
            // initial-step norms d0, d1 (used by INIT slots only)
            float s0 = 0.f, s1 = 0.f;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++) {
                    if (real[r]) {
                        const float sc = ratol + fabsf(y[c][r]) * rrtol;
                        const float r0 = y[c][r] / sc, r1 = k2[c][r] / sc;
                        s0 += r0 * r0;
                        s1 += r1 * r1;
                    }
                }
            s0 = slot_reduce<false>(s0, red0, wave, lane, m);
            s1 = slot_reduce<false>(s1, red1, wave, lane, m);
            if (init) {
                const float d0 = sqrtf(s0 / SD), d1 = sqrtf(s1 / SD);
                h0 = (d0 < 1e-5f || d1 < 1e-5f) ? 1e-6f : 0.01f * d0 / d1;
                h0 = fabsf(h0);
                // stash d1 in dt until the second evaluation
                dt = (double)d1;
            }
            __syncthreads();
        }

        // ---- evaluation 2: k3 = f(y + h(a31 k1 + a32 k2)); INIT: f1 = f(y0 + h0 f0) ----
        {
            const float c1 = stepping ? A31 * h : 0.0f;
            const float c2 = stepping ? A32 * h : (init ? h0 : 0.0f);
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++) ys[c][r] = y[c][r] + (k1[c][r] * c1 + k2[c][r] * c2);
        }
        rhs(k3);
        if (any_init) {
            float s2 = 0.f;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++) {
                    if (real[r]) {
                        const float sc = ratol + fabsf(y[c][r]) * rrtol;
                        const float v = (k3[c][r] - k2[c][r]) / sc;
                        s2 += v * v;
                    }
                }
            s2 = slot_reduce<false>(s2, red0, wave, lane, m);
            if (init) {
                const float d1 = (float)dt;
                const float d2 = fabsf(sqrtf(s2 / SD) / h0);
                float h1;
                if (d1 <= 1e-15f && d2 <= 1e-15f)
                    h1 = fmaxf(1e-6f, h0 * 1e-3f);
                else
                    h1 = powf(0.01f / fmaxf(d1, d2), 0.2f);
                dt = (double)fminf(100.0f * h0, fabsf(h1));
            }
            __syncthreads();
        }

        // ---- evaluations 3..6 (stage inputs; zero coefficients for non-stepping slots) ----
        {
            const float c1 = A41 * h, c2 = A42 * h, c3 = A43 * h;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++)
                    ys[c][r] = y[c][r] + (k1[c][r] * c1 + k2[c][r] * c2 + k3[c][r] * c3);
        }
        rhs(k4);
        {
            const float c1 = A51 * h, c2 = A52 * h, c3 = A53 * h, c4 = A54 * h;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++)
                    ys[c][r] = y[c][r] + (k1[c][r] * c1 + k2[c][r] * c2 + k3[c][r] * c3 + k4[c][r] * c4);
        }
        rhs(k5);
        {
            const float c1 = A61 * h, c2 = A62 * h, c3 = A63 * h, c4 = A64 * h, c5 = A65 * h;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++)
                    ys[c][r] = y[c][r] + (k1[c][r] * c1 + k2[c][r] * c2 + k3[c][r] * c3 +
                                          k4[c][r] * c4 + k5[c][r] * c5);
        }
        rhs(k6);
        {
            const float c1 = B1 * h, c3 = B3 * h, c4 = B4 * h, c5 = B5 * h, c6 = B6 * h;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++)
                    ys[c][r] = y[c][r] + (k1[c][r] * c1 + k3[c][r] * c3 + k4[c][r] * c4 +
                                          k5[c][r] * c5 + k6[c][r] * c6);
        }
        rhs(k7);  // ys now holds y1

        // ---- error norm and max |y1| ----
        float se = 0.f, ymax = 0.f;
        {
            const float e1 = h * E1, e3 = h * E3, e4 = h * E4, e5 = h * E5, e6 = h * E6, e7 = h * E7;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++) {
                    if (real[r]) {
                        const float err = k1[c][r] * e1 + k3[c][r] * e3 + k4[c][r] * e4 +
                                          k5[c][r] * e5 + k6[c][r] * e6 + k7[c][r] * e7;
                        const float tol = ratol + rrtol * fmaxf(fabsf(y[c][r]), fabsf(ys[c][r]));
                        const float v = err * __builtin_amdgcn_rcpf(tol);
                        se += v * v;
                        ymax = fmaxf(ymax, fabsf(ys[c][r]));
                    }
                }
        }
        slot_sum_max(se, ymax, red0, red1, wave, lane, m);

        bool finished = false;
        int32_t st = BH_LANE_OK;
        if (init) {
            // f0 (held in k2) becomes the FSAL derivative; the slot starts stepping
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++) k1[c][r] = k2[c][r];
            mode = STEP;
            if (a.t_end <= a.t0) {
                write_final(ic, y);
                finished = true;
            }
        } else if (stepping) {
            const double ratio = (double)sqrtf(se / SD);
            steps++;
            if (!(tB > t) || isnan(ratio)) {
                st = isnan(ratio) ? BH_LANE_NONFINITE : BH_LANE_DT_UNDERFLOW;
                write_final(ic, y);
                finished = true;
            } else {
                if (ratio <= 1.0) {
                    if (a.t_end <= tB) {
                        // dense output at the end time (torchdiffeq _interp_fit / _interp_evaluate)
                        const float x = (float)((a.t_end - t) / (tB - t));
                        const float x2 = x * x, x3 = x2 * x, x4 = x3 * x;
                        float v[3][NPT];
#pragma unroll
                        for (int c = 0; c < 3; c++)
#pragma unroll
                            for (int r = 0; r < NPT; r++) {
                                const float ymid = y[c][r] + (k1[c][r] * (h * M1) + k3[c][r] * (h * M3) +
                                                              k4[c][r] * (h * M4) + k5[c][r] * (h * M5) +
                                                              k6[c][r] * (h * M6) + k7[c][r] * (h * M7));
                                const float f0 = k1[c][r], f1 = k7[c][r], a0 = y[c][r], a1 = ys[c][r];
                                const float ca = 2.f * h * (f1 - f0) - 8.f * (a1 + a0) + 16.f * ymid;
                                const float cb = h * (5.f * f0 - 3.f * f1) + 18.f * a0 + 14.f * a1 - 32.f * ymid;
                                const float cc = h * (f1 - 4.f * f0) - 11.f * a0 - 5.f * a1 + 16.f * ymid;
                                float s = a0 + x * (h * f0);
                                s = s + x2 * cc;
                                s = s + x3 * cb;
                                v[c][r] = s + x4 * ca;
                            }
                        write_final(ic, v);
                        finished = true;
                    } else {
#pragma unroll
                        for (int c = 0; c < 3; c++)
#pragma unroll
                            for (int r = 0; r < NPT; r++) {
                                y[c][r] = ys[c][r];
                                k1[c][r] = k7[c][r];
                            }
                        t = tB;
                        if ((a.flags & BH_FLAG_STOP) && ymax >= (float)a.stop_thr) {
                            st = BH_LANE_DIVERGED;
                            if (a.flags & BH_FLAG_INF_FILL) {
                                float v[3][NPT];
#pragma unroll
                                for (int c = 0; c < 3; c++)
#pragma unroll
                                    for (int r = 0; r < NPT; r++) v[c][r] = INFINITY;
                                write_final(ic, v);
                            } else {
                                write_final(ic, y);
                            }
                            finished = true;
                        }
                    }
                }
                if (!finished) {
                    dt = next_dt(dt, ratio);
                    if (steps >= a.max_steps) {
                        st = BH_LANE_MAX_STEPS;
                        write_final(ic, y);
                        finished = true;
                    }
                }
            }
        }

        // ---- finished slots record their result and take the next trajectory ----
        if (q == 0) {
            int renewed = 0;
            if (finished) {
                if (a.status) a.status[ic] = st;
                if (a.nsteps) a.nsteps[ic] = steps;
                const long long next = (long long)atomicAdd(a.counter, 1ull);
                slot_ic[m] = next;
                renewed = next < a.n;
            }
            slot_new[m] = renewed;
        }
        __syncthreads();
        any_init = false;
#pragma unroll
        for (int j = 0; j < TILE; j++) any_init = any_init || slot_new[j] != 0;
        if (finished) {
            ic = slot_ic[m];
            if (ic < a.n) {
                load_slot(ic);
                mode = INIT;
            } else {
                mode = IDLE;
            }
        }
    }
}

// the refill counter's start value, written by a kernel on the launch's
// stream (no memset, so the pair can be captured in a graph)
__global__ void set_counter_kernel(unsigned long long *c, unsigned long long v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = v;
}

}  // namespace ross

}  // namespace bh
