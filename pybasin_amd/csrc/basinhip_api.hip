// basinhip_api.hip -- C-ABI entry points of libbasinhip.so (see include/basinhip.h).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "../../include/basinhip.h"
#include "dopri5_lane.hip"
#include "rossler_tile.hip"
#include "mlp.hip"

namespace {

thread_local char g_err[512] = "";

int32_t fail(int32_t code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define BH_HIP_CHECK(expr)                                                              \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(-6, "%s failed: %s", #expr, hipGetErrorString(e_));             \
    } while (0)

// resident lanes that fill the device for a kernel, from its occupancy
template <class K>
int32_t device_lanes(K kernel, int block, int64_t *lanes) {
    int dev = 0, cus = 0, per_cu = 0;
    BH_HIP_CHECK(hipGetDevice(&dev));
    BH_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
    if (per_cu < 1) per_cu = 1;
    *lanes = (int64_t)cus * per_cu * block;
    return 0;
}

template <class Sys, class T, bool kPeaks>
int32_t launch_lane(const bh::LaneArgs &args, int32_t max_lanes, hipStream_t stream) {
    constexpr int block = 256;
    auto kernel = bh::dopri5_lane_kernel<Sys, T, kPeaks>;
    int64_t lanes = 0;
    int32_t rc = device_lanes(kernel, block, &lanes);
    if (rc) return rc;
    // Resident lanes: every trajectory gets its own lane up to the occupancy
    // limit (more resident waves hide the latency of the dependent step
    // arithmetic); beyond that, finished lanes take the next trajectory.
    if (max_lanes > 0 && max_lanes < lanes) lanes = max_lanes;
    if (lanes > args.n) lanes = args.n;
    const int64_t blocks = (lanes + block - 1) / block;
    // lanes [0, blocks*block) start on their own index; the counter hands out the rest
    hipLaunchKernelGGL(bh::set_counter_kernel, dim3(1), dim3(64), 0, stream, args.counter,
                       (unsigned long long)(blocks * block));
    BH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(block), 0, stream, args);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

template <class Sys>
int32_t dispatch_dtype(int32_t dtype, const bh::LaneArgs &args, int32_t max_lanes,
                       hipStream_t s) {
    const bool peaks = args.peak_counts != nullptr;
    if (dtype == BH_F32)
        return peaks ? launch_lane<Sys, float, true>(args, max_lanes, s)
                     : launch_lane<Sys, float, false>(args, max_lanes, s);
    if (dtype == BH_F64)
        return peaks ? launch_lane<Sys, double, true>(args, max_lanes, s)
                     : launch_lane<Sys, double, false>(args, max_lanes, s);
    return fail(-5, "dtype must be BH_F32 or BH_F64, got %d", dtype);
}

struct SysInfo {
    int s, np;
};

int32_t check_system(int32_t sys, int32_t dim, int32_t n_params) {
    SysInfo info;
    switch (sys) {
    case BH_SYS_DECAY:
        if (dim != 1 && dim != 2) return fail(-5, "decay system supports dim 1 or 2, got %d", dim);
        info = {dim, 1};
        break;
    case BH_SYS_LINEAR2: info = {2, 4}; break;
    case BH_SYS_PENDULUM: info = {2, 3}; break;
    case BH_SYS_DUFFING: info = {2, 3}; break;
    case BH_SYS_LORENZ: info = {3, 3}; break;
    default: return fail(-1, "system id %d has no per-lane device functor", sys);
    }
    if (dim != info.s) return fail(-5, "system %d has state dim %d, got %d", sys, info.s, dim);
    if (n_params != info.np)
        return fail(-5, "system %d takes %d parameters, got %d", sys, info.np, n_params);
    return 0;
}

int32_t solve_batch_impl(int32_t system_id, int32_t dtype, const void *y0s, int64_t n, int32_t dim,
                         const double *params, int32_t n_params, int64_t params_stride, double t0,
                         const double *save_ts, int32_t n_save, double rtol, double atol,
                         int64_t max_steps, uint32_t flags, double stop_threshold,
                         int32_t steady_idx, void *out_traj, double *out_feats, void *out_final,
                         int32_t *out_status, int64_t *out_nsteps, int32_t peak_cap,
                         void *out_peak_vals, int32_t *out_peak_counts, void *workspace,
                         int64_t workspace_bytes, int32_t max_lanes, void *stream) {
    g_err[0] = 0;
    int32_t rc = check_system(system_id, dim, n_params);
    if (rc) return rc;
    if (n < 0) return fail(-5, "n must be >= 0, got %lld", (long long)n);
    if (n == 0) return 0;
    if (!y0s || !params || !save_ts) return fail(-5, "y0s, params and save_ts are required");
    if (n_save < 1) return fail(-5, "n_save must be >= 1, got %d", n_save);
    if (params_stride != 0 && params_stride < n_params)
        return fail(-5, "params_stride %lld < n_params %d", (long long)params_stride, n_params);
    if (!(rtol >= 0.0) || !(atol >= 0.0) || rtol + atol <= 0.0)
        return fail(-5, "rtol/atol must be >= 0 with rtol + atol > 0");
    if (max_steps < 1) return fail(-5, "max_steps must be >= 1");
    if (!workspace || workspace_bytes < BH_WORKSPACE_BYTES)
        return fail(-5, "workspace of at least %d bytes is required", BH_WORKSPACE_BYTES);
    if ((flags & BH_FLAG_STOP) && !(stop_threshold > 0.0))
        return fail(-5, "BH_FLAG_STOP needs stop_threshold > 0");
    if (max_lanes < 0) return fail(-5, "max_lanes must be >= 0");

    bh::LaneArgs a;
    a.y0 = y0s;
    a.params = params;
    a.p_stride = params_stride;
    a.n = n;
    a.save_t = save_ts;
    a.n_save = n_save;
    a.steady_idx = steady_idx < 0 ? 0 : steady_idx;
    a.t0 = t0;
    a.rtol = rtol;
    a.atol = atol;
    a.max_steps = max_steps < INT32_MAX ? max_steps : INT32_MAX;  // per-lane counter is 32-bit
    a.flags = flags;
    a.stop_thr = stop_threshold;
    a.traj = out_traj;
    a.feats = (steady_idx < n_save) ? out_feats : nullptr;
    a.fin = out_final;
    a.status = out_status;
    a.nsteps = out_nsteps;
    a.counter = reinterpret_cast<unsigned long long *>(workspace);
    a.peak_vals = out_peak_vals;
    a.peak_counts = out_peak_counts;
    a.peak_cap = peak_cap;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);

    switch (system_id) {
    case BH_SYS_DECAY:
        return dim == 1 ? dispatch_dtype<bh::Decay<1>>(dtype, a, max_lanes, s)
                        : dispatch_dtype<bh::Decay<2>>(dtype, a, max_lanes, s);
    case BH_SYS_LINEAR2: return dispatch_dtype<bh::Linear2>(dtype, a, max_lanes, s);
    case BH_SYS_PENDULUM: return dispatch_dtype<bh::Pendulum>(dtype, a, max_lanes, s);
    case BH_SYS_DUFFING: return dispatch_dtype<bh::Duffing>(dtype, a, max_lanes, s);
    case BH_SYS_LORENZ: return dispatch_dtype<bh::Lorenz>(dtype, a, max_lanes, s);
    default: return fail(-1, "unknown system id %d", system_id);
    }
}

}  // namespace

extern "C" {

int32_t bh_version(void) { return 10000; }

const char *bh_last_error(void) { return g_err; }

int32_t bh_solve_batch(int32_t system_id, int32_t dtype, const void *y0s, int64_t n, int32_t dim,
                       const double *params, int32_t n_params, int64_t params_stride, double t0,
                       const double *save_ts, int32_t n_save, double rtol, double atol,
                       int64_t max_steps, uint32_t flags, double stop_threshold,
                       int32_t steady_idx, void *out_traj, double *out_feats, void *out_final,
                       int32_t *out_status, int64_t *out_nsteps, void *workspace,
                       int64_t workspace_bytes, int32_t max_lanes, void *stream) {
    return solve_batch_impl(system_id, dtype, y0s, n, dim, params, n_params, params_stride, t0,
                            save_ts, n_save, rtol, atol, max_steps, flags, stop_threshold,
                            steady_idx, out_traj, out_feats, out_final, out_status, out_nsteps, 0,
                            nullptr, nullptr, workspace, workspace_bytes, max_lanes, stream);
}

int32_t bh_solve_batch_peaks(int32_t system_id, int32_t dtype, const void *y0s, int64_t n,
                             int32_t dim, const double *params, int32_t n_params,
                             int64_t params_stride, double t0, const double *save_ts,
                             int32_t n_save, double rtol, double atol, int64_t max_steps,
                             uint32_t flags, double stop_threshold, int32_t steady_idx,
                             void *out_traj, double *out_feats, void *out_final,
                             int32_t *out_status, int64_t *out_nsteps, int32_t peak_cap,
                             void *out_peak_vals, int32_t *out_peak_counts, void *workspace,
                             int64_t workspace_bytes, int32_t max_lanes, void *stream) {
    g_err[0] = 0;
    if (!out_peak_counts) return fail(-5, "out_peak_counts is required");
    if (peak_cap < 0) return fail(-5, "peak_cap must be >= 0, got %d", peak_cap);
    if (peak_cap > 0 && !out_peak_vals) return fail(-5, "peak_cap > 0 needs out_peak_vals");
    return solve_batch_impl(system_id, dtype, y0s, n, dim, params, n_params, params_stride, t0,
                            save_ts, n_save, rtol, atol, max_steps, flags, stop_threshold,
                            steady_idx, out_traj, out_feats, out_final, out_status, out_nsteps,
                            peak_cap, out_peak_vals, out_peak_counts, workspace, workspace_bytes,
                            max_lanes, stream);
}

int32_t bh_solve_rossler(int32_t dtype, const void *y0s, int64_t n, const double *params,
                         int64_t params_stride, const float *laplacian, double t0, double t_end,
                         double rtol, double atol, int64_t max_steps, uint32_t flags,
                         double stop_threshold, void *out_final, int32_t *out_status,
                         int64_t *out_nsteps, void *workspace, int64_t workspace_bytes,
                         int32_t max_tiles, void *stream) {
    g_err[0] = 0;
    namespace R = bh::ross;
    if (dtype != BH_F32) return fail(-5, "the Rossler network kernel runs in float32 only");
    if (n < 0) return fail(-5, "n must be >= 0");
    if (n == 0) return 0;
    if (!y0s || !params || !laplacian || !out_final)
        return fail(-5, "y0s, params, laplacian and out_final are required");
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
    R::Args a;
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
    a.lap = laplacian;
    a.fin = reinterpret_cast<float *>(out_final);
    a.status = out_status;
    a.nsteps = out_nsteps;
    a.counter = reinterpret_cast<unsigned long long *>(workspace);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int64_t lanes = 0;
    int32_t rc = device_lanes(R::tile_kernel, R::THREADS, &lanes);
    if (rc) return rc;
    int64_t tiles = lanes / R::THREADS;
    if (max_tiles > 0 && max_tiles < tiles) tiles = max_tiles;
    const int64_t need = (n + R::TILE - 1) / R::TILE;
    if (tiles > need) tiles = need;
    if (tiles < 1) tiles = 1;
    hipLaunchKernelGGL(bh::set_counter_kernel, dim3(1), dim3(64), 0, s, a.counter,
                       (unsigned long long)(tiles * R::TILE));
    BH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(R::tile_kernel, dim3((unsigned)tiles), dim3(R::THREADS), 0, s, a);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
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

}  // extern "C"
