// rossler_wave.hip -- the Rossler network with the edge-list coupling, one
// wavefront per pair of trajectories and no workgroup barrier.
//
// Same right-hand side, step rules and per-trajectory control as the tile
// kernel of rossler_tile.hip, with the coupling of the reference's own
// edge-list form (rossler_network_jax_ode.py:72-75): coupling_i =
// sum over the entries e with edges_i[e] = i of (x_i - x_edges_j[e]) =
// deg_i x_i - sum x_j.  Each wavefront integrates two trajectories, one per
// half: lane l holds slot h = l / 32 and node group g = l % 32, i.e. the
// node positions g, g + 32, g + 64, g + 96 (real below 100; positions map
// to nodes in order of decreasing degree, so a half's 32 groups hold nodes
// of nearly equal degree), as the tile kernel's threads do for its 16
// slots.  The stage input's x goes to the wavefront's own LDS rows, and the
// neighbours' values come back by gathers from them; LDS serves one
// wavefront's instructions in order, so the gathers see the writes without
// a barrier.  The per-trajectory reductions (initial-step norms, error norm,
// max |y|) run over the half by cross-lane exchanges in a fixed order.  A
// wavefront never waits for another: the workgroup-wide barriers that pace
// the tile kernel's 16 trajectories (one or two per evaluation) are gone,
// and a finished slot takes the next trajectory from the device counter on
// its own.
namespace bh {
namespace ross {

constexpr int WTHREADS = 64;  // one wavefront per workgroup
constexpr int WSLOTS = 2;     // trajectories per wavefront
// LDS row stride of a slot: 16 modulo the 32 banks, so the two halves' gathers
// of the same node ids land in banks 16 apart; the host orders every node's
// neighbour list so that the 32 ids one gather instruction reads per half
// are distinct modulo 32 where it can (each bank then serves two lanes)
constexpr int WXS = 112;

// sum (or max) over the 32 lanes of a half, the same value in every lane
template <bool MAX>
__device__ __forceinline__ float half_reduce(float v) {
#pragma unroll
    for (int s = 16; s >= 1; s >>= 1) {
        const float o = __shfl_xor(v, s, 32);
        v = MAX ? fmaxf(v, o) : v + o;
    }
    return v;
}

// Neighbour slots gathered per position slot r (compile time, even): the
// positions g + 32 r of slot r hold nodes of nearly equal degree (sorted),
// so the host picks the smallest instantiation whose bounds cover every
// slot's largest degree.
template <int G0, int G1, int G2, int G3>
__global__ void __launch_bounds__(WTHREADS) wave_kernel(Args a) {
    constexpr int GATHER[NPT] = {G0, G1, G2, G3};
    // per slot: the x of its 100 nodes and a zero cell for unused neighbour slots
    __shared__ float X[WSLOTS * WXS];
    const int lane = threadIdx.x;
    const int h = lane >> 5, g = lane & 31;

    bool real[NPT];
    int node[NPT];
    uint32_t off[NPT][DMAX];  // LDS byte offsets of the neighbour slots
    float deg[NPT];
#pragma unroll
    for (int r = 0; r < NPT; r++) {
        const int pos = g + 32 * r;
        real[r] = pos < NN;
        node[r] = real[r] ? (int)a.g.node[pos] : NN;
        deg[r] = real[r] ? (float)a.g.deg[pos] : 0.0f;
#pragma unroll
        for (int d = 0; d < DMAX; d++)
            off[r][d] = (uint32_t)(h * WXS + (real[r] ? a.g.nbr[pos][d] : NN)) * 4u;
    }
    if (g == 0) X[h * WXS + NN] = 0.0f;

    float y[3][NPT], k1[3][NPT], k2[3][NPT], k3[3][NPT], k4[3][NPT], k5[3][NPT], k6[3][NPT],
        k7[3][NPT], ys[3][NPT];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < NPT; r++)
            y[c][r] = k1[c][r] = k2[c][r] = k3[c][r] = k4[c][r] = k5[c][r] = k6[c][r] = k7[c][r] =
                ys[c][r] = 0.0f;
    // per-slot scalars (replicated over the half's 32 lanes)
    long long ic = (long long)blockIdx.x * WSLOTS + h;
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

    // one right-hand-side evaluation of ys -> out
    const char *Xb = reinterpret_cast<const char *>(X);
    auto rhs = [&](float (&out)[3][NPT]) {
        float part[NPT];
#pragma unroll
        for (int r = 0; r < NPT; r++) {
            const float xx = ys[0][r], yy = ys[1][r], zz = ys[2][r];
            part[r] = -yy - zz;
            out[1][r] = xx + pa * yy;
            out[2][r] = pb + zz * (xx - pc);
            if (real[r]) X[h * WXS + node[r]] = xx;
        }
        // the gathers below follow the writes in this wavefront's order
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < NPT; r++) {
            float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
            for (int d = 0; d < GATHER[r]; d += 2) {
                s0 += *reinterpret_cast<const float *>(Xb + off[r][d]);
                s1 += *reinterpret_cast<const float *>(Xb + off[r][d + 1]);
            }
            // a position without a node gathers the zero cell and has degree
            // 0: it integrates an uncoupled oscillator from zero, bounded,
            // never written and outside every norm
            const float cpl = fmaf(deg[r], ys[0][r], -(s0 + s1));
            out[0][r] = part[r] - pK * cpl;
        }
        // the next evaluation's writes follow these gathers
        __builtin_amdgcn_wave_barrier();
    };

    // The initial-step norms are needed only in iterations where a slot of
    // the wavefront starts a trajectory; the test is wavefront-uniform.
    bool any_init = true;
    for (;;) {
        if (!__any(mode != IDLE)) break;
        const bool stepping = mode == STEP, init = mode == INIT;
        const double tB = t + dt;
        const float hs = stepping ? (float)dt : 0.0f;

        // ---- evaluation 1: k2 = f(y + a21 h k1); INIT: f0 = f(y0) into k2 ----
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < NPT; r++) ys[c][r] = y[c][r] + k1[c][r] * (A21 * hs);
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
            s0 = half_reduce<false>(s0);
            s1 = half_reduce<false>(s1);
            if (init) {
                const float d0 = sqrtf(s0 / SD), d1 = sqrtf(s1 / SD);
                h0 = (d0 < 1e-5f || d1 < 1e-5f) ? 1e-6f : 0.01f * d0 / d1;
                h0 = fabsf(h0);
                // stash d1 in dt until the second evaluation
                dt = (double)d1;
            }
        }

        // ---- evaluation 2: k3 = f(y + h(a31 k1 + a32 k2)); INIT: f1 = f(y0 + h0 f0) ----
        {
            const float c1 = stepping ? A31 * hs : 0.0f;
            const float c2 = stepping ? A32 * hs : (init ? h0 : 0.0f);
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
            s2 = half_reduce<false>(s2);
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
        }

        // ---- evaluations 3..6 (stage inputs; zero coefficients for non-stepping slots) ----
        {
            const float c1 = A41 * hs, c2 = A42 * hs, c3 = A43 * hs;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++)
                    ys[c][r] = y[c][r] + (k1[c][r] * c1 + k2[c][r] * c2 + k3[c][r] * c3);
        }
        rhs(k4);
        {
            const float c1 = A51 * hs, c2 = A52 * hs, c3 = A53 * hs, c4 = A54 * hs;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++)
                    ys[c][r] = y[c][r] + (k1[c][r] * c1 + k2[c][r] * c2 + k3[c][r] * c3 + k4[c][r] * c4);
        }
        rhs(k5);
        {
            const float c1 = A61 * hs, c2 = A62 * hs, c3 = A63 * hs, c4 = A64 * hs, c5 = A65 * hs;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < NPT; r++)
                    ys[c][r] = y[c][r] + (k1[c][r] * c1 + k2[c][r] * c2 + k3[c][r] * c3 +
                                          k4[c][r] * c4 + k5[c][r] * c5);
        }
        rhs(k6);
        {
            const float c1 = B1 * hs, c3 = B3 * hs, c4 = B4 * hs, c5 = B5 * hs, c6 = B6 * hs;
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
            const float e1 = hs * E1, e3 = hs * E3, e4 = hs * E4, e5 = hs * E5, e6 = hs * E6,
                        e7 = hs * E7;
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
        se = half_reduce<false>(se);
        ymax = half_reduce<true>(ymax);

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
                                const float ymid = y[c][r] + (k1[c][r] * (hs * M1) + k3[c][r] * (hs * M3) +
                                                              k4[c][r] * (hs * M4) + k5[c][r] * (hs * M5) +
                                                              k6[c][r] * (hs * M6) + k7[c][r] * (hs * M7));
                                const float f0 = k1[c][r], f1 = k7[c][r], a0 = y[c][r], a1 = ys[c][r];
                                const float ca = 2.f * hs * (f1 - f0) - 8.f * (a1 + a0) + 16.f * ymid;
                                const float cb = hs * (5.f * f0 - 3.f * f1) + 18.f * a0 + 14.f * a1 - 32.f * ymid;
                                const float cc = hs * (f1 - 4.f * f0) - 11.f * a0 - 5.f * a1 + 16.f * ymid;
                                float sv = a0 + x * (hs * f0);
                                sv = sv + x2 * cc;
                                sv = sv + x3 * cb;
                                v[c][r] = sv + x4 * ca;
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
        long long next = 0;
        if (finished && g == 0) {
            if (a.status) a.status[ic] = st;
            if (a.nsteps) a.nsteps[ic] = steps;
            next = (long long)atomicAdd(a.counter, 1ull);
        }
        // every lane of the half takes its first lane's index
        next = __shfl(next, h * 32);
        any_init = __any(finished && next < a.n);
        if (finished) {
            ic = next;
            if (ic < a.n) {
                load_slot(ic);
                mode = INIT;
            } else {
                mode = IDLE;
            }
        }
    }
}

}  // namespace ross
}  // namespace bh
