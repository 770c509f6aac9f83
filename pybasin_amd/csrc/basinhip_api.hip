// basinhip_api.hip -- C-ABI entry points of libbasinhip.so (see include/basinhip.h);
// the per-lane solver's entry points are in lane_api.hip.
#include "api_common.hpp"
#include <algorithm>

#include "rossler_tile.hip"
#include "rossler_wave.hip"
#include "mlp.hip"
#include "cluster.hip"
#include "window.hip"

#include <system_error>
#include <thread>
#include <vector>

thread_local char bhapi::g_err[512] = "";

namespace {

using bhapi::device_lanes;
using bhapi::fail;
using bhapi::g_err;

// feature slots for F features: 2, 4, 8 or 16
inline int feature_slots(int F) { return F <= 2 ? 2 : F <= 4 ? 4 : F <= 8 ? 8 : F <= 16 ? 16 : 32; }

template <int K, int FC>
int32_t launch_knn_fc(const double *Q, int64_t nq, const double *R, int64_t nr, int32_t F,
                      int32_t k, double *out_d, int64_t *out_i, hipStream_t s) {
    namespace C = bh::clu;
    const int64_t blocks = (nq + C::TILE - 1) / C::TILE;
    hipLaunchKernelGGL((C::knn_kernel<K, FC>), dim3((unsigned)blocks), dim3(C::TILE), 0, s, Q, nq,
                       R, nr, F, k, out_d, out_i);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

template <int K>
int32_t launch_knn(const double *Q, int64_t nq, const double *R, int64_t nr, int32_t F, int32_t k,
                   double *out_d, int64_t *out_i, hipStream_t s) {
    switch (feature_slots(F)) {
    case 2: return launch_knn_fc<K, 2>(Q, nq, R, nr, F, k, out_d, out_i, s);
    case 4: return launch_knn_fc<K, 4>(Q, nq, R, nr, F, k, out_d, out_i, s);
    case 8: return launch_knn_fc<K, 8>(Q, nq, R, nr, F, k, out_d, out_i, s);
    case 16: return launch_knn_fc<K, 16>(Q, nq, R, nr, F, k, out_d, out_i, s);
    default: return launch_knn_fc<K, 32>(Q, nq, R, nr, F, k, out_d, out_i, s);
    }
}

template <int FC>
int32_t tile_bounds(const double *X, int64_t n, int32_t F, const double *core_sq, int64_t n_tiles,
                    double *lo, double *hi, double *mincore, hipStream_t s) {
    hipLaunchKernelGGL(bh::clu::tile_bounds_kernel<FC>, dim3((unsigned)((n_tiles + 255) / 256)),
                       dim3(256), 0, s, X, n, F, core_sq, n_tiles, lo, hi, mincore);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

template <int FC>
int32_t core_dist_launch(const double *X, int64_t n, int32_t F, int32_t k, const double *lo,
                         const double *hi, int64_t n_tiles, double *core, hipStream_t s) {
    namespace C = bh::clu;
    const dim3 grid((unsigned)n_tiles), block(C::QT);
    if (k <= 8)
        hipLaunchKernelGGL((C::core_distance_kernel<8, FC>), grid, block, 0, s, X, n, F, k, lo, hi, n_tiles, core);
    else if (k <= 16)
        hipLaunchKernelGGL((C::core_distance_kernel<16, FC>), grid, block, 0, s, X, n, F, k, lo, hi, n_tiles, core);
    else
        hipLaunchKernelGGL((C::core_distance_kernel<32, FC>), grid, block, 0, s, X, n, F, k, lo, hi, n_tiles, core);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

template <int FC>
int32_t boruvka_round(const double *X, int64_t n, int32_t F, const double *core_sq,
                      const int32_t *comp, const double *lo, const double *hi, const double *mincore,
                      int32_t *tcomp, int64_t n_tiles, double *best_key, int64_t *best_j,
                      hipStream_t s) {
    namespace C = bh::clu;
    hipLaunchKernelGGL(C::tile_comp_kernel, dim3((unsigned)((n_tiles + 255) / 256)), dim3(256), 0, s,
                       comp, n, n_tiles, tcomp);
    BH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(C::boruvka_kernel<FC>, dim3((unsigned)n_tiles), dim3(C::QT), 0, s, X, n, F,
                       core_sq, comp, lo, hi, mincore, tcomp, n_tiles, best_key, best_j);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

template <int FC>
int32_t edge_weights(const double *X, int32_t F, const double *core, const int64_t *ea,
                     const int64_t *eb, int64_t m, double *w, hipStream_t s) {
    hipLaunchKernelGGL(bh::clu::edge_weight_kernel<FC>, dim3((unsigned)((m + 255) / 256)), dim3(256),
                       0, s, X, F, core, ea, eb, m, w);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

template <int FC>
int32_t silhouette_launch(const double *X, const double *w, int64_t n, int32_t F, const int32_t *lab,
                          const int64_t *starts, const double *counts, int32_t C_, double *sums,
                          double *s_out, hipStream_t s) {
    namespace C = bh::clu;
    hipLaunchKernelGGL(C::silhouette_sums_kernel<FC>,
                       dim3((unsigned)((n + C::TILE - 1) / C::TILE), (unsigned)C_), dim3(C::TILE), 0,
                       s, X, w, n, F, starts, sums);
    BH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(C::silhouette_value_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       sums, n, lab, starts, counts, C_, s_out);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

#define BH_BY_SLOTS(F, call)                      \
    switch (feature_slots(F)) {                   \
    case 2: { constexpr int FCV = 2; rc = call; } break; \
    case 4: { constexpr int FCV = 4; rc = call; } break; \
    case 8: { constexpr int FCV = 8; rc = call; } break; \
    case 16: { constexpr int FCV = 16; rc = call; } break; \
    default: { constexpr int FCV = 32; rc = call; } break; \
    }

}  // namespace

extern "C" {

int32_t bh_version(void) { return 10000; }

const char *bh_last_error(void) { return g_err; }

}  // extern "C"

namespace {

// Argument checks and launch shared by the dense (tile kernel) and the
// edge-list (wavefront kernel) entry points; slots = trajectories per
// workgroup.
template <class Kernel>
int32_t rossler_launch(Kernel kernel, int threads, int slots, bh::ross::Args &a, int32_t dtype,
                       const void *y0s, int64_t n, const double *params, int64_t params_stride,
                       double t0, double t_end, double rtol, double atol, int64_t max_steps,
                       uint32_t flags, double stop_threshold, void *out_final, int32_t *out_status,
                       int64_t *out_nsteps, void *workspace, int64_t workspace_bytes,
                       int32_t max_tiles, void *stream) {
    if (dtype != BH_F32) return fail(-5, "the Rossler network kernel runs in float32 only");
    if (n < 0) return fail(-5, "n must be >= 0");
    if (n == 0) return 0;
    if (!y0s || !params || !out_final) return fail(-5, "y0s, params and out_final are required");
    if (params_stride != 0 && params_stride < 4) return fail(-5, "params_stride must be 0 or >= 4");
    if (!(rtol >= 0.0) || !(atol >= 0.0) || rtol + atol <= 0.0)
        return fail(-5, "rtol/atol must be >= 0 with rtol + atol > 0");
    if (!(t_end >= t0)) return fail(-5, "t_end must be >= t0");
    if (max_steps < 1) return fail(-5, "max_steps must be >= 1");
    if (!workspace || workspace_bytes < BH_WORKSPACE_BYTES)
        return fail(-5, "workspace of at least %d bytes is required", BH_WORKSPACE_BYTES);
    if ((flags & BH_FLAG_STOP) && !(stop_threshold > 0.0))
        return fail(-5, "BH_FLAG_STOP needs stop_threshold > 0");
    if (max_tiles < 0) return fail(-5, "max_tiles must be >= 0");
    a.y0 = reinterpret_cast<const float *>(y0s);
    a.params = params;
    a.p_stride = params_stride;
    a.n = n;
    a.t0 = t0;
    a.t_end = t_end;
    a.rtol = rtol;
    a.atol = atol;
    a.max_steps = max_steps;
    a.flags = flags;
    a.stop_thr = stop_threshold;
    a.fin = reinterpret_cast<float *>(out_final);
    a.status = out_status;
    a.nsteps = out_nsteps;
    a.counter = reinterpret_cast<unsigned long long *>(workspace);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int64_t lanes = 0;
    int32_t rc = device_lanes(kernel, threads, &lanes);
    if (rc) return rc;
    // max_tiles caps the resident trajectories at 16 * max_tiles
    int64_t blocks = lanes / threads;
    if (max_tiles > 0 && (int64_t)max_tiles * 16 / slots < blocks)
        blocks = std::max<int64_t>(1, (int64_t)max_tiles * 16 / slots);
    const int64_t need = (n + slots - 1) / slots;
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(bh::ross::set_counter_kernel, dim3(1), dim3(64), 0, s, a.counter,
                       (unsigned long long)(blocks * slots));
    BH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(threads), 0, s, a);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int32_t bh_solve_rossler(int32_t dtype, const void *y0s, int64_t n, const double *params,
                         int64_t params_stride, const float *laplacian, double t0, double t_end,
                         double rtol, double atol, int64_t max_steps, uint32_t flags,
                         double stop_threshold, void *out_final, int32_t *out_status,
                         int64_t *out_nsteps, void *workspace, int64_t workspace_bytes,
                         int32_t max_tiles, void *stream) {
    g_err[0] = 0;
    if (n > 0 && !laplacian) return fail(-5, "laplacian is required");
    bh::ross::Args a{};
    a.lap = laplacian;
    return rossler_launch(bh::ross::tile_kernel, bh::ross::THREADS, bh::ross::TILE, a, dtype,
                          y0s, n, params, params_stride, t0, t_end, rtol, atol, max_steps, flags,
                          stop_threshold, out_final, out_status, out_nsteps, workspace,
                          workspace_bytes, max_tiles, stream);
}

int32_t bh_solve_rossler_edges(int32_t dtype, const void *y0s, int64_t n, const double *params,
                               int64_t params_stride, const int32_t *edges_i,
                               const int32_t *edges_j, int64_t n_edges, double t0, double t_end,
                               double rtol, double atol, int64_t max_steps, uint32_t flags,
                               double stop_threshold, void *out_final, int32_t *out_status,
                               int64_t *out_nsteps, void *workspace, int64_t workspace_bytes,
                               int32_t max_tiles, void *stream) {
    g_err[0] = 0;
    namespace R = bh::ross;
    if (n_edges < 0) return fail(-5, "n_edges must be >= 0");
    if (n_edges > 0 && (!edges_i || !edges_j)) return fail(-5, "edges_i and edges_j are required");
    // host-side graph table: degrees, neighbour lists, positions by degree
    int deg[R::NN] = {0};
    uint8_t nbr[R::NN][R::DMAX];
    for (int64_t e = 0; e < n_edges; e++) {
        const int32_t i = edges_i[e], j = edges_j[e];
        if (i < 0 || i >= R::NN || j < 0 || j >= R::NN)
            return fail(-5, "edge %lld (%d, %d) outside the %d nodes", (long long)e, (int)i, (int)j,
                        R::NN);
        if (deg[i] == R::DMAX)
            return fail(-5, "node %d has more than %d edge entries (use bh_solve_rossler with the "
                        "dense Laplacian)", (int)i, R::DMAX);
        nbr[i][deg[i]++] = (uint8_t)j;
    }
    bh::ross::Args a{};
    // positions in order of decreasing degree, ties by node index
    int order[R::NN];
    for (int i = 0; i < R::NN; i++) order[i] = i;
    std::stable_sort(order, order + R::NN, [&](int u, int v) { return deg[u] > deg[v]; });
    for (int p = 0; p < R::NPOS; p++) {
        const int node = p < R::NN ? order[p] : 0;
        a.g.node[p] = (uint8_t)(p < R::NN ? node : R::NN);
        a.g.deg[p] = (uint8_t)(p < R::NN ? deg[node] : 0);
        for (int d = 0; d < R::DMAX; d++) a.g.nbr[p][d] = (uint8_t)R::NN;
    }
    // gather bounds: the largest degree among each slot's 32 positions,
    // covered by the tighter instantiation when it fits
    int cap[R::NPT];
    for (int r = 0; r < R::NPT; r++) {
        cap[r] = 0;
        for (int p = 32 * r; p < 32 * r + 32 && p < R::NN; p++) cap[r] = std::max(cap[r], deg[order[p]]);
    }
    static constexpr int kTight[R::NPT] = {12, 10, 8, 6};
    bool tight = true;
    for (int r = 0; r < R::NPT; r++) tight = tight && cap[r] <= kTight[r];
    for (int r = 0; r < R::NPT; r++) cap[r] = tight ? kTight[r] : R::DMAX;
    // Neighbour slot of each edge entry: one gather instruction reads slot d
    // of the 32 positions r * 32 .. r * 32 + 31, so for each such group the
    // entries are coloured greedily (positions of high degree first, then
    // neighbours in order) with the lowest slot below the group's bound whose
    // gather has not yet used the neighbour's bank (its id modulo 32); an
    // entry that finds none takes the lowest free slot.  Any order gives the
    // same sums; this one spreads the gathers over the LDS banks.
    for (int r = 0; r < R::NPT; r++) {
        uint32_t used[R::DMAX] = {0};  // banks taken per slot
        for (int p = 32 * r; p < 32 * r + 32 && p < R::NN; p++) {
            const int node = order[p];
            uint32_t taken = 0;  // slots of this position
            for (int e = 0; e < deg[node]; e++) {
                const int j = nbr[node][e], bank = j & 31;
                int slot = -1;
                for (int d = 0; d < cap[r] && slot < 0; d++)
                    if (!(taken >> d & 1u) && !(used[d] >> bank & 1u)) slot = d;
                for (int d = 0; d < cap[r] && slot < 0; d++)
                    if (!(taken >> d & 1u)) slot = d;
                taken |= 1u << slot;
                used[slot] |= 1u << bank;
                a.g.nbr[p][slot] = (uint8_t)j;
            }
        }
    }
    if (tight)
        return rossler_launch(bh::ross::wave_kernel<12, 10, 8, 6>, bh::ross::WTHREADS,
                              bh::ross::WSLOTS, a, dtype, y0s, n, params, params_stride, t0, t_end,
                              rtol, atol, max_steps, flags, stop_threshold, out_final, out_status,
                              out_nsteps, workspace, workspace_bytes, max_tiles, stream);
    return rossler_launch(bh::ross::wave_kernel<12, 12, 12, 12>, bh::ross::WTHREADS, bh::ross::WSLOTS, a, dtype, y0s,
                          n, params, params_stride, t0, t_end, rtol, atol, max_steps, flags,
                          stop_threshold, out_final, out_status, out_nsteps, workspace,
                          workspace_bytes, max_tiles, stream);
}

// ---- surrogate MLP --------------------------------------------------------

#define BH_GRID(count, block) dim3((unsigned)(((count) + (block) - 1) / (block)))

int32_t bh_gemm_bf16(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk,
                     int64_t sbn, float *C, int64_t ldc, int64_t M, int64_t N, int64_t K,
                     int32_t accumulate, void *stream) {
    g_err[0] = 0;
    if (M < 0 || N < 0 || K < 0) return fail(-5, "negative GEMM size");
    if (M == 0 || N == 0) return 0;
    if (!A || !B || !C) return fail(-5, "null GEMM operand");
    if ((M + 63) / 64 > 65535) return fail(-5, "M too large for one launch (%lld)", (long long)M);
    bh::mlp::GemmArgs g{reinterpret_cast<const unsigned short *>(A),
                        reinterpret_cast<const unsigned short *>(B), C, M, N, K, sam, sak, sbk, sbn,
                        ldc, accumulate};
    hipLaunchKernelGGL(bh::mlp::gemm_bf16_kernel, dim3((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64)),
                       dim3(256), 0, reinterpret_cast<hipStream_t>(stream), g);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_mlp_bias_act(const float *Z, int64_t M, int64_t N, const float *bias, int32_t silu,
                        float *zsave, void *out_bf16, int64_t ldo, float *out_f32, void *stream) {
    g_err[0] = 0;
    if (M <= 0 || N <= 0) return M == 0 ? 0 : fail(-5, "bad sizes");
    if (!Z || !bias) return fail(-5, "null argument");
    hipLaunchKernelGGL(bh::mlp::bias_act_kernel, BH_GRID(M * N, 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), Z, M, N, bias, silu, zsave,
                       reinterpret_cast<unsigned short *>(out_bf16), ldo, out_f32);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_mlp_silu_bwd(const float *dA, const float *Z, int64_t count, float *dZ,
                        void *dZ_bf16, void *stream) {
    g_err[0] = 0;
    if (count <= 0) return count == 0 ? 0 : fail(-5, "bad count");
    if (!dA || !Z || !dZ || !dZ_bf16) return fail(-5, "null argument");
    hipLaunchKernelGGL(bh::mlp::silu_bwd_kernel, BH_GRID(count, 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), dA, Z, count, dZ,
                       reinterpret_cast<unsigned short *>(dZ_bf16));
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_mlp_softmax_ce(const float *logits, const int32_t *labels, int64_t M, float *dlogits,
                          void *dlogits_bf16, float *loss, void *stream) {
    g_err[0] = 0;
    if (M <= 0) return M == 0 ? 0 : fail(-5, "bad M");
    if (!logits || !labels || !dlogits || !dlogits_bf16 || !loss) return fail(-5, "null argument");
    hipLaunchKernelGGL(bh::mlp::softmax_ce_kernel, BH_GRID(M, 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), logits, labels, M, dlogits,
                       reinterpret_cast<unsigned short *>(dlogits_bf16), loss);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_mlp_colsum(const float *G, int64_t M, int64_t N, float *out, void *stream) {
    g_err[0] = 0;
    if (M < 0 || N <= 0) return N == 0 ? 0 : fail(-5, "bad sizes");
    if (!G || !out) return fail(-5, "null argument");
    hipLaunchKernelGGL(bh::mlp::colsum_kernel, dim3((unsigned)N), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), G, M, N, out);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_mlp_adam(float *W, const float *G, float *m, float *v, int64_t count, float lr,
                    float beta1, float beta2, float eps, int64_t step, void *W_bf16, void *stream) {
    g_err[0] = 0;
    if (count <= 0) return count == 0 ? 0 : fail(-5, "bad count");
    if (!W || !G || !m || !v || step < 1) return fail(-5, "bad Adam arguments");
    const float bc1 = 1.0f - powf(beta1, (float)step), bc2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(bh::mlp::adam_kernel, BH_GRID(count, 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), W, G, m, v, count, lr, beta1, beta2,
                       eps, bc1, bc2, reinterpret_cast<unsigned short *>(W_bf16));
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_mlp_prep_input(const float *x, int64_t M, float c0, float s0, float c1, float s1,
                          float c2, float s2, void *out_bf16, void *stream) {
    g_err[0] = 0;
    if (M <= 0) return M == 0 ? 0 : fail(-5, "bad M");
    if (!x || !out_bf16) return fail(-5, "null argument");
    hipLaunchKernelGGL(bh::mlp::prep_input_kernel, BH_GRID(M * 16, 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), x, M, c0, s0, c1, s1, c2, s2,
                       reinterpret_cast<unsigned short *>(out_bf16));
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_mlp_classify(const float *Z, const float *bias, int64_t M, int32_t *label, float *conf,
                        void *stream) {
    g_err[0] = 0;
    if (M <= 0) return M == 0 ? 0 : fail(-5, "bad M");
    if (!Z || !bias || !label || !conf) return fail(-5, "null argument");
    hipLaunchKernelGGL(bh::mlp::classify_kernel, BH_GRID(M, 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), Z, bias, M, label, conf);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_mlp_infer(const float *x, int64_t M, const void *w_packed, const float *biases,
                     float c0, float s0, float c1, float s1, float c2, float s2, int32_t *label,
                     float *conf, void *stream) {
    g_err[0] = 0;
    namespace P = bh::mlp;
    if (M <= 0) return M == 0 ? 0 : fail(-5, "bad M");
    if (!x || !w_packed || !biases || !label || !conf) return fail(-5, "null argument");
    const int64_t tiles = (M + P::IT_ROWS - 1) / P::IT_ROWS;
    if (tiles > 2147483647LL) return fail(-5, "M too large");
    const P::bf16x8v *w = reinterpret_cast<const P::bf16x8v *>(w_packed);
    P::InferArgs a;
    a.x = x;
    a.M = M;
    a.w1 = w;                      // 1 kstep x 16 colblocks x 64 lanes (x 8 bf16)
    a.w2 = a.w1 + 1 * 16 * 64;     // 8 x 16 x 64
    a.w3 = a.w2 + 8 * 16 * 64;
    a.w4 = a.w3 + 8 * 16 * 64;     // 8 x 1 x 64
    a.b1 = biases;
    a.b2 = biases + 256;
    a.b3 = biases + 512;
    a.b4 = biases + 768;
    a.c0 = c0; a.s0 = s0; a.c1 = c1; a.s1 = s1; a.c2 = c2; a.s2 = s2;
    a.label = label;
    a.conf = conf;
    hipLaunchKernelGGL(P::mlp_infer_kernel, dim3((unsigned)tiles), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), a);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

// Workspace of bh_mlp_train_step, carved in this order (each piece aligned to
// 256 bytes): input activations (M, 16) bf16; hidden activations 3 x (M, H)
// bf16; saved pre-activations 3 x (M, H) f32; GEMM output (M, H) f32; output
// logits and their gradient (M, 2) f32 x 2 + (M, 2) bf16; per-row loss (M)
// f32; back-propagated gradients 2 x (M, H) f32 + 2 x (M, H) bf16 (ping-pong)
// and dA (M, H) f32; one layer's weight and bias gradients max(16, H) x H f32
// + H f32.
namespace {
constexpr int64_t kMlpIn = 16;  // padded input width (bh_mlp_prep_input)

struct TrainLayout {
    int64_t a0, act[3], z[3], zbuf, logits, dlog, dlogb, loss, g[2], gb16[2], dA, gw, gbias, total;
};

TrainLayout train_layout(int64_t M, int64_t H) {
    TrainLayout L{};
    int64_t off = 0;
    auto take = [&off](int64_t bytes) {
        const int64_t at = off;
        off += (bytes + 255) / 256 * 256;
        return at;
    };
    L.a0 = take(M * kMlpIn * 2);
    for (int i = 0; i < 3; i++) L.act[i] = take(M * H * 2);
    for (int i = 0; i < 3; i++) L.z[i] = take(M * H * 4);
    L.zbuf = take(M * H * 4);
    L.logits = take(M * 2 * 4);
    L.dlog = take(M * 2 * 4);
    L.dlogb = take(M * 2 * 2);
    L.loss = take(M * 4);
    for (int i = 0; i < 2; i++) L.g[i] = take(M * H * 4);
    for (int i = 0; i < 2; i++) L.gb16[i] = take(M * H * 2);
    L.dA = take(M * H * 4);
    L.gw = take((H > kMlpIn ? H : kMlpIn) * H * 4);
    L.gbias = take(H * 4);
    L.total = off;
    return L;
}
}  // namespace

int64_t bh_mlp_train_workspace_bytes(int64_t M, int32_t hidden) {
    if (M < 0 || hidden < 2) return -5;
    return train_layout(M, hidden).total;
}

int32_t bh_mlp_train_step(const float *x, const int32_t *labels, int64_t M, const float *norm,
                          int32_t hidden, float *const *W, void *const *W_bf16, float *const *b,
                          float *const *mW, float *const *vW, float *const *mb, float *const *vb,
                          float lr, float beta1, float beta2, float eps, int64_t step,
                          float *loss_out, void *workspace, int64_t workspace_bytes,
                          void *stream) {
    g_err[0] = 0;
    if (M < 0) return fail(-5, "M must be >= 0, got %lld", (long long)M);
    if (hidden < 2) return fail(-5, "hidden must be >= 2, got %d", hidden);
    if (!W || !W_bf16 || !b || !mW || !vW || !mb || !vb || !norm)
        return fail(-5, "the parameter, moment and norm arrays are required");
    for (int l = 0; l < 4; l++)
        if (!W[l] || !W_bf16[l] || !b[l] || !mW[l] || !vW[l] || !mb[l] || !vb[l])
            return fail(-5, "null pointer in the arrays of layer %d", l);
    if (step < 1) return fail(-5, "step must be >= 1 (Adam's bias correction)");
    if (M == 0) return 0;
    if (!x || !labels) return fail(-5, "x and labels are required");
    const int64_t H = hidden;
    const TrainLayout L = train_layout(M, H);
    if (!workspace || workspace_bytes < L.total)
        return fail(-5, "workspace of %lld bytes is required", (long long)L.total);
    char *ws = reinterpret_cast<char *>(workspace);
    auto f32 = [ws](int64_t off) { return reinterpret_cast<float *>(ws + off); };
    auto b16 = [ws](int64_t off) { return reinterpret_cast<void *>(ws + off); };
    int32_t rc;
#define BH_TRY(call)          \
    do {                      \
        rc = (call);          \
        if (rc) return rc;    \
    } while (0)
    // forward: input features, three SiLU layers, logits (pre-activations kept)
    BH_TRY(bh_mlp_prep_input(x, M, norm[0], norm[1], norm[2], norm[3], norm[4], norm[5],
                             b16(L.a0), stream));
    const int64_t K_in[4] = {kMlpIn, H, H, H}, N_out[4] = {H, H, H, 2};
    const void *act_in[4] = {b16(L.a0), b16(L.act[0]), b16(L.act[1]), b16(L.act[2])};
    for (int l = 0; l < 3; l++) {
        BH_TRY(bh_gemm_bf16(act_in[l], K_in[l], 1, W_bf16[l], N_out[l], 1, f32(L.zbuf), N_out[l], M,
                            N_out[l], K_in[l], 0, stream));
        BH_TRY(bh_mlp_bias_act(f32(L.zbuf), M, N_out[l], b[l], 1, f32(L.z[l]), b16(L.act[l]),
                               N_out[l], nullptr, stream));
    }
    BH_TRY(bh_gemm_bf16(act_in[3], H, 1, W_bf16[3], 2, 1, f32(L.zbuf), 2, M, 2, H, 0, stream));
    BH_TRY(bh_mlp_bias_act(f32(L.zbuf), M, 2, b[3], 0, nullptr, nullptr, 0, f32(L.logits), stream));
    float *loss = loss_out ? loss_out : f32(L.loss);
    BH_TRY(bh_mlp_softmax_ce(f32(L.logits), labels, M, f32(L.dlog), b16(L.dlogb), loss, stream));
    // backward, last layer first; each layer's Adam update runs once its weights
    // have served the gradient of the layer below
    const float *gz = f32(L.dlog);
    const void *gzb = b16(L.dlogb);
    for (int l = 3; l >= 0; l--) {
        BH_TRY(bh_gemm_bf16(act_in[l], 1, K_in[l], gzb, N_out[l], 1, f32(L.gw), N_out[l], K_in[l],
                            N_out[l], M, 0, stream));
        BH_TRY(bh_mlp_colsum(gz, M, N_out[l], f32(L.gbias), stream));
        if (l > 0) {
            const int nxt = l & 1;
            BH_TRY(bh_gemm_bf16(gzb, N_out[l], 1, W_bf16[l], 1, N_out[l], f32(L.dA), K_in[l], M,
                                K_in[l], N_out[l], 0, stream));
            BH_TRY(bh_mlp_silu_bwd(f32(L.dA), f32(L.z[l - 1]), M * K_in[l], f32(L.g[nxt]),
                                   b16(L.gb16[nxt]), stream));
            gz = f32(L.g[nxt]);
            gzb = b16(L.gb16[nxt]);
        }
        BH_TRY(bh_mlp_adam(W[l], f32(L.gw), mW[l], vW[l], K_in[l] * N_out[l], lr, beta1, beta2, eps,
                           step, W_bf16[l], stream));
        BH_TRY(bh_mlp_adam(b[l], f32(L.gbias), mb[l], vb[l], N_out[l], lr, beta1, beta2, eps, step,
                           nullptr, stream));
    }
#undef BH_TRY
    return 0;
}


// ---- label layer ------------------------------------------------------------

int32_t bh_knn(const double *Q, int64_t nq, const double *R, int64_t nr, int32_t F, int32_t k,
               double *out_d, int64_t *out_i, void *stream) {
    g_err[0] = 0;
    namespace C = bh::clu;
    if (nq < 0 || nr < 1) return fail(-5, "bh_knn: nq >= 0 and nr >= 1 required");
    if (F < 1 || F > C::MAXF) return fail(-5, "bh_knn: 1 <= F <= %d, got %d", C::MAXF, F);
    if (k < 1 || k > C::MAXK) return fail(-5, "bh_knn: 1 <= k <= %d, got %d", C::MAXK, k);
    if (nq == 0) return 0;
    if (!Q || !R || !out_d || !out_i) return fail(-5, "bh_knn: null argument");
    if ((nq + C::TILE - 1) / C::TILE > 2147483647LL) return fail(-5, "bh_knn: nq too large");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (k <= 8) return launch_knn<8>(Q, nq, R, nr, F, k, out_d, out_i, s);
    if (k <= 16) return launch_knn<16>(Q, nq, R, nr, F, k, out_d, out_i, s);
    return launch_knn<32>(Q, nq, R, nr, F, k, out_d, out_i, s);
}

int32_t bh_core_distances(const double *X, int64_t n, int32_t F, int32_t k, double *out_core,
                          void *workspace, int64_t workspace_bytes, void *stream) {
    g_err[0] = 0;
    namespace C = bh::clu;
    if (n < 1) return fail(-5, "bh_core_distances: n >= 1 required");
    if (F < 1 || F > C::MAXF) return fail(-5, "bh_core_distances: 1 <= F <= %d, got %d", C::MAXF, F);
    if (k < 1 || k > C::MAXK || k > n) return fail(-5, "bh_core_distances: 1 <= k <= min(%d, n), got %d", C::MAXK, k);
    if (!X || !out_core || !workspace) return fail(-5, "bh_core_distances: null argument");
    if (workspace_bytes < bh_hdbscan_workspace_bytes(n))
        return fail(-5, "bh_core_distances: workspace of %lld bytes required",
                    (long long)bh_hdbscan_workspace_bytes(n));
    if (n > 2147483647LL) return fail(-5, "bh_core_distances: n too large");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t n_tiles = (n + C::QT - 1) / C::QT;
    double *tlo = reinterpret_cast<double *>(workspace);
    double *thi = tlo + n_tiles * C::MAXF;
    int32_t rc = 0;
    BH_BY_SLOTS(F, (tile_bounds<FCV>(X, n, F, nullptr, n_tiles, tlo, thi, nullptr, s)));
    if (rc) return rc;
    BH_BY_SLOTS(F, (core_dist_launch<FCV>(X, n, F, k, tlo, thi, n_tiles, out_core, s)));
    return rc;
}

int64_t bh_hdbscan_workspace_bytes(int64_t n) {
    namespace C = bh::clu;
    const int64_t tiles = (n + C::QT - 1) / C::QT;
    // squared cores, best keys, best partners, components; per tile a box
    // of 2 x MAXF slots, the smallest core, the shared component; alignment
    return n * (8 + 8 + 8 + 4) + tiles * ((2 * C::MAXF + 1) * 8 + 4) + 1024;
}

int32_t bh_hdbscan_mst(const double *X, int64_t n, int32_t F, const double *core, int64_t *out_a,
                       int64_t *out_b, double *out_w, void *workspace, int64_t workspace_bytes,
                       void *stream) {
    g_err[0] = 0;
    namespace C = bh::clu;
    if (n < 2) return fail(-5, "bh_hdbscan_mst: n >= 2 required");
    if (F < 1 || F > C::MAXF) return fail(-5, "bh_hdbscan_mst: 1 <= F <= %d", C::MAXF);
    if (!X || !core || !out_a || !out_b || !out_w || !workspace)
        return fail(-5, "bh_hdbscan_mst: null argument");
    if (workspace_bytes < bh_hdbscan_workspace_bytes(n))
        return fail(-5, "bh_hdbscan_mst: workspace of %lld bytes required",
                    (long long)bh_hdbscan_workspace_bytes(n));
    if (n > 2147483647LL) return fail(-5, "bh_hdbscan_mst: n too large");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char *ws = reinterpret_cast<char *>(workspace);
    const int64_t n_tiles = (n + C::QT - 1) / C::QT;
    double *core_sq = reinterpret_cast<double *>(ws);
    double *best_key = core_sq + n;
    int64_t *best_j = reinterpret_cast<int64_t *>(best_key + n);
    double *tlo = reinterpret_cast<double *>(best_j + n);
    double *thi = tlo + n_tiles * C::MAXF;
    double *tmin = thi + n_tiles * C::MAXF;
    int32_t *comp_d = reinterpret_cast<int32_t *>(tmin + n_tiles);
    int32_t *tcomp = comp_d + n;
    // squared core distances for the comparison key
    std::vector<double> core_h(n);
    BH_HIP_CHECK(hipMemcpyAsync(core_h.data(), core, n * sizeof(double), hipMemcpyDeviceToHost, s));
    BH_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<double> csq(n);
    for (int64_t i = 0; i < n; i++) csq[i] = core_h[i] * core_h[i];
    BH_HIP_CHECK(hipMemcpyAsync(core_sq, csq.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    int32_t rc = 0;
    BH_BY_SLOTS(F, (tile_bounds<FCV>(X, n, F, core_sq, n_tiles, tlo, thi, tmin, s)));
    if (rc) return rc;
    C::UnionFind uf(n);
    std::vector<int32_t> comp(n);
    std::iota(comp.begin(), comp.end(), 0);
    std::vector<double> bk(n);
    std::vector<int64_t> bj(n), ea, eb;
    ea.reserve(n - 1);
    eb.reserve(n - 1);
    int64_t n_comp = n;
    while (n_comp > 1) {
        BH_HIP_CHECK(hipMemcpyAsync(comp_d, comp.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
        BH_BY_SLOTS(F, (boruvka_round<FCV>(X, n, F, core_sq, comp_d, tlo, thi, tmin, tcomp, n_tiles,
                                           best_key, best_j, s)));
        if (rc) return rc;
        BH_HIP_CHECK(hipMemcpyAsync(bk.data(), best_key, n * sizeof(double), hipMemcpyDeviceToHost, s));
        BH_HIP_CHECK(hipMemcpyAsync(bj.data(), best_j, n * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        BH_HIP_CHECK(hipStreamSynchronize(s));
        // cheapest edge per component (ties: lower key, then lower point, then lower partner)
        std::vector<int64_t> pick(n, -1);
        for (int64_t i = 0; i < n; i++) {
            if (bj[i] < 0) continue;
            const int32_t c = comp[i];
            const int64_t p = pick[c];
            if (p < 0 || bk[i] < bk[p] || (bk[i] == bk[p] && (i < p || (i == p && bj[i] < bj[p]))))
                pick[c] = i;
        }
        int64_t merged = 0;
        for (int64_t c = 0; c < n; c++) {
            const int64_t i = pick[c];
            if (i < 0) continue;
            const int64_t a = uf.find(i), b = uf.find(bj[i]);
            if (a == b) continue;  // the partner component chose the same edge
            int64_t big = a, small = b;
            if (uf.size[big] < uf.size[small]) std::swap(big, small);
            uf.parent[small] = big;
            uf.size[big] += uf.size[small];
            ea.push_back(i);
            eb.push_back(bj[i]);
            merged++;
        }
        if (merged == 0) return fail(-3, "bh_hdbscan_mst: no edge between components");
        n_comp -= merged;
        for (int64_t i = 0; i < n; i++) comp[i] = (int32_t)uf.find(i);
    }
    const int64_t m = (int64_t)ea.size();
    BH_HIP_CHECK(hipMemcpyAsync(out_a, ea.data(), m * sizeof(int64_t), hipMemcpyHostToDevice, s));
    BH_HIP_CHECK(hipMemcpyAsync(out_b, eb.data(), m * sizeof(int64_t), hipMemcpyHostToDevice, s));
    BH_BY_SLOTS(F, (edge_weights<FCV>(X, F, core, out_a, out_b, m, out_w, s)));
    if (rc) return rc;
    BH_HIP_CHECK(hipStreamSynchronize(s));
    return 0;
}

int32_t bh_hdbscan_labels(const int64_t *edge_a, const int64_t *edge_b, const double *edge_w,
                          int64_t n, const int32_t *min_cluster_sizes, int32_t n_sizes,
                          int32_t *labels, int32_t *n_clusters) {
    g_err[0] = 0;
    namespace C = bh::clu;
    if (n < 2) return fail(-5, "bh_hdbscan_labels: n >= 2 required");
    if (n_sizes < 1 || !min_cluster_sizes) return fail(-5, "bh_hdbscan_labels: no min_cluster_size");
    for (int32_t k = 0; k < n_sizes; k++)
        if (min_cluster_sizes[k] < 2)
            return fail(-5, "bh_hdbscan_labels: min_cluster_size >= 2 required");
    if (!edge_a || !edge_b || !edge_w || !labels) return fail(-5, "bh_hdbscan_labels: null argument");
    for (int64_t e = 0; e < n - 1; e++)
        if (edge_a[e] < 0 || edge_a[e] >= n || edge_b[e] < 0 || edge_b[e] >= n || !(edge_w[e] >= 0.0))
            return fail(-5, "bh_hdbscan_labels: edge %lld out of range", (long long)e);
    // one hierarchy, condensed once per requested min_cluster_size; the sizes
    // are independent and run on host threads of their own
    const C::Hierarchy h = C::single_linkage(n, edge_a, edge_b, edge_w);
    auto one = [&](int32_t k) {
        const C::Condensed c = C::condense(h, n, min_cluster_sizes[k]);
        const int32_t nc = C::select_and_label(c, n, labels + (int64_t)k * n);
        if (n_clusters) n_clusters[k] = nc;
    };
    if (n_sizes == 1) {
        one(0);
        return 0;
    }
    std::vector<std::thread> pool;
    pool.reserve(n_sizes - 1);
    try {
        for (int32_t k = 1; k < n_sizes; k++) pool.emplace_back(one, k);
    } catch (const std::system_error &) {
        // no thread to spare: the remaining sizes run here
        for (int32_t k = (int32_t)pool.size() + 1; k < n_sizes; k++) one(k);
    }
    one(0);
    for (std::thread &t : pool) t.join();
    return 0;
}

int32_t bh_silhouette(const double *X_sorted, int64_t n, int32_t F, const int32_t *labels_sorted,
                      const int64_t *starts, int32_t n_labels, const double *weights,
                      const double *label_counts, double *sums, double *s_out, void *stream) {
    g_err[0] = 0;
    namespace C = bh::clu;
    if (n < 1 || F < 1 || F > C::MAXF || n_labels < 1)
        return fail(-5, "bh_silhouette: n >= 1, 1 <= F <= %d, n_labels >= 1 required", C::MAXF);
    if (!X_sorted || !labels_sorted || !starts || !sums || !s_out)
        return fail(-5, "bh_silhouette: null argument");
    if ((weights == nullptr) != (label_counts == nullptr))
        return fail(-5, "bh_silhouette: weights and label_counts go together");
    if (n_labels > 65535) return fail(-5, "bh_silhouette: at most 65535 labels");
    if ((n + C::TILE - 1) / C::TILE > 2147483647LL) return fail(-5, "bh_silhouette: n too large");
    int32_t rc = 0;
    BH_BY_SLOTS(F, (silhouette_launch<FCV>(X_sorted, weights, n, F, labels_sorted, starts,
                                            label_counts, n_labels, sums, s_out,
                                            reinterpret_cast<hipStream_t>(stream))));
    return rc;
}

// ---- windowed order statistics ---------------------------------------------

int32_t bh_window_order_stats(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                              int32_t dim, int64_t start, const int64_t *ranks, int32_t n_ranks,
                              void *out, void *stream) {
    g_err[0] = 0;
    if (n < 0 || dim < 1 || n_save < 1) return fail(-5, "bh_window_order_stats: bad sizes");
    if (start < 0 || start >= n_save) return fail(-5, "bh_window_order_stats: empty window");
    if (n_ranks < 1 || !ranks) return fail(-5, "bh_window_order_stats: no ranks");
    if (n == 0) return 0;
    if (!traj || !out) return fail(-5, "bh_window_order_stats: null argument");
    const int64_t np = n * dim;
    if ((np + 255) / 256 > 2147483647LL) return fail(-5, "bh_window_order_stats: too many rows");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)((np + 255) / 256));
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::order_stat_kernel<float>, grid, dim3(256), 0, s,
                           reinterpret_cast<const float *>(traj), n_save, n, dim, start, ranks,
                           n_ranks, reinterpret_cast<float *>(out));
    else if (dtype == BH_F64)
        hipLaunchKernelGGL(bh::win::order_stat_kernel<double>, grid, dim3(256), 0, s,
                           reinterpret_cast<const double *>(traj), n_save, n, dim, start, ranks,
                           n_ranks, reinterpret_cast<double *>(out));
    else
        return fail(-5, "bh_window_order_stats: dtype must be BH_F32 or BH_F64");
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---- series features of a kept window ---------------------------------------

namespace {

// series per workgroup and dynamic LDS bytes for a window of w rows (62 KB
// of series data plus the autocorrelation blocks of the four waves)
int32_t series_launch(int64_t w, size_t elem, int *G, size_t *lds, bool ac) {
    const int64_t budget = 62 * 1024;
    const int64_t per = w * (int64_t)elem;
    if (per > budget)
        return fail(-5, "window of %lld rows exceeds the %lld-byte LDS budget of a series",
                    (long long)w, (long long)budget);
    int64_t g = budget / per;
    if (g > 16) g = 16;
    *G = (int)g;
    *lds = (size_t)((g * per + 15) / 16 * 16 + (ac ? 4 * 64 * sizeof(double) : 0));
    return 0;
}

int32_t check_series_args(const char *fn, int32_t dtype, const void *traj, int64_t n_save,
                          int64_t n, int32_t dim, int64_t start) {
    if (n < 0 || dim < 1 || n_save < 1) return fail(-5, "%s: bad sizes", fn);
    if (start < 0 || start >= n_save) return fail(-5, "%s: empty window", fn);
    if (dtype != BH_F32 && dtype != BH_F64) return fail(-5, "%s: dtype must be BH_F32 or BH_F64", fn);
    if (n > 0 && !traj) return fail(-5, "%s: null trajectory buffer", fn);
    return 0;
}

}  // namespace

int32_t bh_window_autocorr_periodicity(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                                       int32_t dim, int64_t start, int32_t min_lag,
                                       double peak_threshold, void *out_strength,
                                       void *out_period, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_autocorr_periodicity", dtype, traj, n_save, n, dim,
                                   start);
    if (rc) return rc;
    if (min_lag < 0) return fail(-5, "bh_window_autocorr_periodicity: min_lag must be >= 0");
    if (n == 0) return 0;
    if (!out_strength || !out_period) return fail(-5, "bh_window_autocorr_periodicity: null output");
    const size_t elem = dtype == BH_F32 ? 4 : 8;
    int G;
    size_t lds;
    rc = series_launch(n_save - start, elem, &G, &lds, true);
    if (rc) return rc;
    const int64_t np = n * dim;
    const int64_t blocks = (np + G - 1) / G;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_autocorr_periodicity: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::autocorr_kernel<float>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const float *>(traj),
                           n_save, n, dim, start, G, min_lag, peak_threshold,
                           reinterpret_cast<float *>(out_strength), reinterpret_cast<float *>(out_period));
    else
        hipLaunchKernelGGL(bh::win::autocorr_kernel<double>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s,
                           reinterpret_cast<const double *>(traj), n_save, n, dim, start, G, min_lag,
                           peak_threshold, reinterpret_cast<double *>(out_strength),
                           reinterpret_cast<double *>(out_period));
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_spectral_ratio(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                                 int32_t dim, int64_t start, void *out_ratio, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_spectral_ratio", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (n == 0) return 0;
    if (!out_ratio) return fail(-5, "bh_window_spectral_ratio: null output");
    const size_t elem = dtype == BH_F32 ? 4 : 8;
    int G;
    size_t lds;
    rc = series_launch(n_save - start, elem, &G, &lds, false);
    if (rc) return rc;
    const int64_t np = n * dim;
    const int64_t blocks = (np + G - 1) / G;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_spectral_ratio: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::spectral_ratio_kernel<float>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const float *>(traj),
                           n_save, n, dim, start, G, reinterpret_cast<float *>(out_ratio));
    else
        hipLaunchKernelGGL(bh::win::spectral_ratio_kernel<double>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s,
                           reinterpret_cast<const double *>(traj), n_save, n, dim, start, G,
                           reinterpret_cast<double *>(out_ratio));
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_spectrum(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                           int32_t dim, int64_t start, int32_t n_coeff, double *out_coeff,
                           double *out_agg, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_spectrum", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (n_coeff < 0 || n_coeff > 4096) return fail(-5, "bh_window_spectrum: n_coeff must be in [0, 4096]");
    if (n_coeff > 0 && !out_coeff) return fail(-5, "bh_window_spectrum: null out_coeff");
    if (n_coeff == 0 && !out_agg) return fail(-5, "bh_window_spectrum: nothing to compute");
    if (n == 0) return 0;
    const int64_t w = n_save - start;
    const size_t elem = dtype == BH_F32 ? 4 : 8;
    // series rows plus (with out_agg) one row of w / 2 + 1 powers per wave
    const int64_t budget = 62 * 1024;
    const int64_t pw_bytes = out_agg ? 4 * (w / 2 + 1) * 8 : 0;
    const int64_t per = w * (int64_t)elem;
    if (per + pw_bytes > budget)
        return fail(-5, "bh_window_spectrum: window of %lld rows exceeds the %lld-byte LDS budget",
                    (long long)w, (long long)budget);
    int64_t G = (budget - pw_bytes) / per;
    if (G > 16) G = 16;
    const size_t lds = (size_t)((G * per + 15) / 16 * 16 + pw_bytes);
    const int64_t np = n * dim;
    const int64_t blocks = (np + G - 1) / G;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_spectrum: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::spectrum_kernel<float>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const float *>(traj),
                           n_save, n, dim, start, (int)G, n_coeff, out_coeff, out_agg);
    else
        hipLaunchKernelGGL(bh::win::spectrum_kernel<double>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const double *>(traj),
                           n_save, n, dim, start, (int)G, n_coeff, out_coeff, out_agg);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_autocov(int32_t dtype, const void *traj, int64_t n_save, int64_t n, int32_t dim,
                          int64_t start, int32_t max_lag, double *out, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_autocov", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (!out) return fail(-5, "bh_window_autocov: null out");
    if (max_lag < 0 || max_lag > 4095) return fail(-5, "bh_window_autocov: max_lag must be in [0, 4095]");
    if (n == 0) return 0;
    const int64_t w = n_save - start;
    const size_t elem = dtype == BH_F32 ? 4 : 8;
    int G = 0;
    size_t lds = 0;
    rc = series_launch(w, elem, &G, &lds, false);
    if (rc) return rc;
    const int64_t np = n * dim;
    const int64_t blocks = (np + G - 1) / G;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_autocov: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::autocov_kernel<float>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const float *>(traj),
                           n_save, n, dim, start, G, max_lag, out);
    else
        hipLaunchKernelGGL(bh::win::autocov_kernel<double>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const double *>(traj),
                           n_save, n, dim, start, G, max_lag, out);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_trends(int32_t dtype, const void *traj, int64_t n_save, int64_t n, int32_t dim,
                         int64_t start, int32_t want_cwt, double *out, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_trends", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (!out) return fail(-5, "bh_window_trends: null out");
    if (n == 0) return 0;
    const int64_t np = n * dim;
    const int64_t blocks = (np + 255) / 256;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_trends: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::trends_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, s,
                           reinterpret_cast<const float *>(traj), n_save, n, dim, start,
                           want_cwt ? 1 : 0, out);
    else
        hipLaunchKernelGGL(bh::win::trends_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, s,
                           reinterpret_cast<const double *>(traj), n_save, n, dim, start,
                           want_cwt ? 1 : 0, out);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_peak_counts(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                              int32_t dim, int64_t start, int32_t s0, int32_t s1, int32_t s2,
                              int32_t s3, double *out, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_peak_counts", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (!out) return fail(-5, "bh_window_peak_counts: null out");
    const int32_t sup[4] = {s0, s1, s2, s3};
    for (int k = 0; k < 4; k++)
        if (sup[k] < 0 || sup[k] > 64) return fail(-5, "bh_window_peak_counts: supports must be in [0, 64]");
    if (n == 0) return 0;
    const int64_t w = n_save - start;
    const int64_t elem = dtype == BH_F32 ? 4 : 8;
    // the block's series rows plus one list of w 16-bit candidate positions per wave
    const int64_t budget = 62 * 1024, lists = (4 * w * 2 + 15) / 16 * 16;
    auto need = [&](int64_t r) { return (4 * r * 2 + 15) / 16 * 16 + (r * elem + 15) / 16 * 16; };
    if (need(w) > budget || w > 65535) {
        int64_t wmax = budget / (8 + elem);
        while (wmax > 0 && need(wmax) > budget) wmax--;
        return fail(-5, "bh_window_peak_counts: window of %lld rows exceeds the LDS budget (at most %lld rows)",
                    (long long)w, (long long)wmax);
    }
    int64_t G = (budget - lists) / (w * elem);
    if (G > 16) G = 16;
    const size_t lds = (size_t)((G * w * elem + 15) / 16 * 16 + lists);
    const int64_t np = n * dim;
    const int64_t blocks = (np + G - 1) / G;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_peak_counts: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int4 sv = make_int4(s0, s1, s2, s3);
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::peak_counts_kernel<float>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const float *>(traj),
                           n_save, n, dim, start, (int)G, sv, out);
    else
        hipLaunchKernelGGL(bh::win::peak_counts_kernel<double>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const double *>(traj),
                           n_save, n, dim, start, (int)G, sv, out);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_perm_entropy(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                               int32_t dim, int64_t start, double *out, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_perm_entropy", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (!out) return fail(-5, "bh_window_perm_entropy: null out");
    if (n == 0) return 0;
    const int64_t w = n_save - start;
    const int64_t elem = dtype == BH_F32 ? 4 : 8;
    // the block's series rows plus one histogram of 5,910 16-bit counts per wave
    const int64_t budget = 62 * 1024, hist = 4 * (5910 / 2) * 4;
    if (hist + w * elem > budget)
        return fail(-5, "bh_window_perm_entropy: window of %lld rows exceeds the LDS budget (at most %lld rows)",
                    (long long)w, (long long)((budget - hist) / elem));
    int64_t G = (budget - hist) / (w * elem);
    if (G > 16) G = 16;
    const size_t lds = (size_t)((G * w * elem + 15) / 16 * 16 + hist);
    const int64_t np = n * dim;
    const int64_t blocks = (np + G - 1) / G;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_perm_entropy: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::perm_entropy_kernel<float>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const float *>(traj),
                           n_save, n, dim, start, (int)G, out);
    else
        hipLaunchKernelGGL(bh::win::perm_entropy_kernel<double>, dim3((unsigned)blocks),
                           dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const double *>(traj),
                           n_save, n, dim, start, (int)G, out);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_embedding(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                            int32_t dim, int64_t start, int32_t which, double *out, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_embedding", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (!out) return fail(-5, "bh_window_embedding: null out");
    if (which != BH_EMB_CORR_DIM && which != BH_EMB_LYAP_R && which != BH_EMB_LYAP_E)
        return fail(-5, "bh_window_embedding: which must be BH_EMB_CORR_DIM, BH_EMB_LYAP_R or BH_EMB_LYAP_E");
    if (n == 0) return 0;
    const int64_t w = n_save - start;
    const int64_t elem = dtype == BH_F32 ? 4 : 8;
    // the block's series rows (plus, for the correlation sum, the radii and
    // the histogram of each wave)
    const int64_t budget = 62 * 1024;
    const int64_t extra = which == BH_EMB_CORR_DIM ? 4 * 50 * 8 + 4 * 64 * 4
                          : (which == BH_EMB_LYAP_E ? 4 * 64 * 15 * 8 : 0);
    // lyap_e keys carry an orbit point in 13 bits: the budget keeps windows below 8,204 rows
    static_assert((62 * 1024 - 4 * 64 * 15 * 8) / 4 - 12 <= (int64_t)bh::win::kLeKeyMask, "lyap_e key bits");
    if (extra + w * elem > budget)
        return fail(-5, "bh_window_embedding: window of %lld rows exceeds the %lld-byte LDS budget",
                    (long long)w, (long long)budget);
    int64_t G = (budget - extra) / (w * elem);
    if (G > 16) G = 16;
    const size_t lds = (size_t)((G * w * elem + 15) / 16 * 16 + extra);
    const int64_t np = n * dim;
    const int64_t blocks = (np + G - 1) / G;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_embedding: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (which == BH_EMB_CORR_DIM) {
        if (dtype == BH_F32)
            hipLaunchKernelGGL(bh::win::corr_dim_kernel<float>, dim3((unsigned)blocks),
                               dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const float *>(traj),
                               n_save, n, dim, start, (int)G, out);
        else
            hipLaunchKernelGGL(bh::win::corr_dim_kernel<double>, dim3((unsigned)blocks),
                               dim3(bh::win::kSeriesBlock), lds, s,
                               reinterpret_cast<const double *>(traj), n_save, n, dim, start, (int)G, out);
    } else if (which == BH_EMB_LYAP_E) {
        if (dtype == BH_F32)
            hipLaunchKernelGGL(bh::win::lyap_e_kernel<float>, dim3((unsigned)blocks),
                               dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const float *>(traj),
                               n_save, n, dim, start, (int)G, out);
        else
            hipLaunchKernelGGL(bh::win::lyap_e_kernel<double>, dim3((unsigned)blocks),
                               dim3(bh::win::kSeriesBlock), lds, s,
                               reinterpret_cast<const double *>(traj), n_save, n, dim, start, (int)G, out);
    } else {
        if (dtype == BH_F32)
            hipLaunchKernelGGL(bh::win::lyap_r_kernel<float>, dim3((unsigned)blocks),
                               dim3(bh::win::kSeriesBlock), lds, s, reinterpret_cast<const float *>(traj),
                               n_save, n, dim, start, (int)G, out);
        else
            hipLaunchKernelGGL(bh::win::lyap_r_kernel<double>, dim3((unsigned)blocks),
                               dim3(bh::win::kSeriesBlock), lds, s,
                               reinterpret_cast<const double *>(traj), n_save, n, dim, start, (int)G, out);
    }
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_lempel_ziv(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                             int32_t dim, int64_t start, const int32_t *bins, int32_t n_bins,
                             double *out, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_lempel_ziv", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (!out || !bins) return fail(-5, "bh_window_lempel_ziv: null bins or out");
    if (n_bins < 1 || n_bins > 8) return fail(-5, "bh_window_lempel_ziv: n_bins must be in [1, 8]");
    int32_t bv[8] = {2, 2, 2, 2, 2, 2, 2, 2};
    for (int k = 0; k < n_bins; k++) {
        if (bins[k] < 1 || bins[k] > 255) return fail(-5, "bh_window_lempel_ziv: bins must be in [1, 255]");
        bv[k] = bins[k];
    }
    const int64_t w = n_save - start;
    if (w > 1024) return fail(-5, "bh_window_lempel_ziv: windows of at most 1,024 rows (got %lld)", (long long)w);
    if (n == 0) return 0;
    const int64_t elem = dtype == BH_F32 ? 4 : 8;
    int cap_log = 6;
    while ((1LL << cap_log) < 2 * w) cap_log++;
    // one hash table of 2^cap_log 32-bit entries and one row of w one-byte
    // symbols per bins setting, then the series rows
    const int64_t budget = 62 * 1024,
                  tables = (int64_t)n_bins * (4LL << cap_log) + ((int64_t)n_bins * w + 15) / 16 * 16;
    if (tables + w * elem > budget)
        return fail(-5, "bh_window_lempel_ziv: %d settings of a %lld-row window exceed the LDS budget",
                    n_bins, (long long)w);
    int64_t G = (budget - tables) / (w * elem);
    if (G > 4) G = 4;
    const size_t lds = (size_t)((G * w * elem + 15) / 16 * 16 + tables);
    const int64_t np = n * dim;
    const int64_t blocks = (np + G - 1) / G;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_lempel_ziv: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int4 ba = make_int4(bv[0], bv[1], bv[2], bv[3]), bb = make_int4(bv[4], bv[5], bv[6], bv[7]);
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::lempel_ziv_kernel<float>, dim3((unsigned)blocks), dim3(64), lds, s,
                           reinterpret_cast<const float *>(traj), n_save, n, dim, start, (int)G, cap_log,
                           ba, bb, n_bins, out);
    else
        hipLaunchKernelGGL(bh::win::lempel_ziv_kernel<double>, dim3((unsigned)blocks), dim3(64), lds, s,
                           reinterpret_cast<const double *>(traj), n_save, n, dim, start, (int)G, cap_log,
                           ba, bb, n_bins, out);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C++" template <class T>
static void launch_sorted(int64_t P, int64_t blocks, size_t lds, hipStream_t s, const T *traj,
                          int64_t n_save, int64_t n, int32_t dim, int64_t start, int64_t G,
                          double *out_dup, T *out_levels, T *out_median) {
    // values per lane: P / 64 (P >= 64)
    const int R = (int)(P / 64);
#define BH_SORTED_LAUNCH(RR)                                                                   \
    hipLaunchKernelGGL((bh::win::sorted_kernel<T, RR>), dim3((unsigned)blocks),                 \
                       dim3(bh::win::kSeriesBlock), lds, s, traj, n_save, n, dim, start, (int)G, \
                       out_dup, out_levels, out_median)
    switch (R) {
        case 1: BH_SORTED_LAUNCH(1); break;
        case 2: BH_SORTED_LAUNCH(2); break;
        case 4: BH_SORTED_LAUNCH(4); break;
        case 8: BH_SORTED_LAUNCH(8); break;
        case 16: BH_SORTED_LAUNCH(16); break;
        default: BH_SORTED_LAUNCH(32); break;
    }
#undef BH_SORTED_LAUNCH
}

int32_t bh_window_sorted(int32_t dtype, const void *traj, int64_t n_save, int64_t n, int32_t dim,
                         int64_t start, double *out_dup, void *out_levels, void *out_median,
                         void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_sorted", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (!out_dup) return fail(-5, "bh_window_sorted: null out_dup");
    if (n == 0) return 0;
    const int64_t w = n_save - start;
    const int64_t elem = dtype == BH_F32 ? 4 : 8;
    int64_t P = 64;
    while (P < w) P <<= 1;
    // the block's series rows plus one sorted row of P values per wave
    const int64_t budget = 62 * 1024;
    const int64_t rows = 4 * P * elem;
    hipStream_t s0 = reinterpret_cast<hipStream_t>(stream);
    if (P > 2048 || rows + w * elem > budget) {
        // longer windows: one series per block, sorted in LDS by the block
        if (P * elem > 32 * 1024)
            return fail(-5, "bh_window_sorted: window of %lld rows exceeds the LDS budget (at most %d %s rows)",
                        (long long)w, dtype == BH_F32 ? 8192 : 4096, dtype == BH_F32 ? "float32" : "float64");
        const int64_t np = n * dim;
        if (np > 2147483647LL) return fail(-5, "bh_window_sorted: too many series");
        const size_t lds = (size_t)(P * elem + 4 * 5 * 8);
        if (dtype == BH_F32)
            hipLaunchKernelGGL(bh::win::sorted_block_kernel<float>, dim3((unsigned)np), dim3(256), lds, s0,
                               reinterpret_cast<const float *>(traj), n_save, n, dim, start, (int)P, out_dup,
                               reinterpret_cast<float *>(out_levels), reinterpret_cast<float *>(out_median));
        else
            hipLaunchKernelGGL(bh::win::sorted_block_kernel<double>, dim3((unsigned)np), dim3(256), lds, s0,
                               reinterpret_cast<const double *>(traj), n_save, n, dim, start, (int)P, out_dup,
                               reinterpret_cast<double *>(out_levels), reinterpret_cast<double *>(out_median));
        BH_HIP_CHECK(hipGetLastError());
        return 0;
    }
    int64_t G = (budget - rows) / (w * elem);
    if (G > 16) G = 16;
    const size_t lds = (size_t)((G * w * elem + 15) / 16 * 16 + rows);
    const int64_t np = n * dim;
    const int64_t blocks = (np + G - 1) / G;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_sorted: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (dtype == BH_F32)
        launch_sorted<float>(P, blocks, lds, s, reinterpret_cast<const float *>(traj), n_save, n, dim,
                             start, G, out_dup, reinterpret_cast<float *>(out_levels),
                             reinterpret_cast<float *>(out_median));
    else
        launch_sorted<double>(P, blocks, lds, s, reinterpret_cast<const double *>(traj), n_save, n,
                              dim, start, G, out_dup, reinterpret_cast<double *>(out_levels),
                              reinterpret_cast<double *>(out_median));
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_change_quantiles(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                                   int32_t dim, int64_t start, const void *levels, double *out,
                                   void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_change_quantiles", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (!levels || !out) return fail(-5, "bh_window_change_quantiles: null levels or out");
    if (n == 0) return 0;
    const int64_t np = n * dim;
    const int64_t blocks = (np + 255) / 256;
    if (blocks > 2147483647LL) return fail(-5, "bh_window_change_quantiles: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::change_quantiles_kernel<float>, dim3((unsigned)blocks), dim3(256), 0,
                           s, reinterpret_cast<const float *>(traj), n_save, n, dim, start,
                           reinterpret_cast<const float *>(levels), out);
    else
        hipLaunchKernelGGL(bh::win::change_quantiles_kernel<double>, dim3((unsigned)blocks), dim3(256),
                           0, s, reinterpret_cast<const double *>(traj), n_save, n, dim, start,
                           reinterpret_cast<const double *>(levels), out);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

int32_t bh_window_features(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                           int32_t dim, int64_t start, double *out, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_series_args("bh_window_features", dtype, traj, n_save, n, dim, start);
    if (rc) return rc;
    if (n == 0) return 0;
    if (!out) return fail(-5, "bh_window_features: null output");
    const int64_t np = n * dim;
    if ((np + 255) / 256 > 2147483647LL) return fail(-5, "bh_window_features: too many series");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)((np + 255) / 256));
    if (dtype == BH_F32)
        hipLaunchKernelGGL(bh::win::window_features_kernel<float>, grid, dim3(256), 0, s,
                           reinterpret_cast<const float *>(traj), n_save, n, dim, start, out);
    else
        hipLaunchKernelGGL(bh::win::window_features_kernel<double>, grid, dim3(256), 0, s,
                           reinterpret_cast<const double *>(traj), n_save, n, dim, start, out);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // extern "C"
