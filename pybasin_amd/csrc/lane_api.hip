// lane_api.hip -- C-ABI entry points of the per-lane Dopri5 solver
// (bh_solve_batch, bh_solve_batch_peaks, bh_solve_batch_phased,
// bh_autotune_lane; declared in include/basinhip.h).
// Its own translation unit so that it can be built with the instruction
// scheduler that favours independent instructions: the step loop of a
// wave alone on its SIMD is bound by the latency of its dependent
// arithmetic.
#include <mutex>

#include "api_common.hpp"
#include "dopri5_lane.hip"

namespace {

using bhapi::device_lanes;
using bhapi::fail;
using bhapi::g_err;

template <class Sys, class T, bool kPeaks>
int32_t launch_lane(const bh::LaneArgs &args, int32_t max_lanes, hipStream_t stream) {
    constexpr int block = 256;
    int64_t lanes = 0;
    int32_t rc = device_lanes(bh::dopri5_lane_kernel<Sys, T, kPeaks, true>, block, &lanes);
    if (rc) return rc;
    // Resident lanes: every trajectory gets its own lane up to the occupancy
    // limit (more resident waves hide the latency of the dependent step
    // arithmetic); beyond that, finished lanes take the next trajectory.
    if (max_lanes > 0 && max_lanes < lanes) lanes = max_lanes;
    if (lanes > args.n) lanes = args.n;
    const int64_t blocks = (lanes + block - 1) / block;
    if (blocks * block >= args.n) {
        // every trajectory has a lane from the start: no refills
        hipLaunchKernelGGL((bh::dopri5_lane_kernel<Sys, T, kPeaks, false>), dim3((unsigned)blocks),
                           dim3(block), 0, stream, args);
    } else {
        // lanes [0, blocks*block) start on their own index; the counter hands out the rest
        hipLaunchKernelGGL(bh::set_counter_kernel, dim3(1), dim3(64), 0, stream, args.counter,
                           (unsigned long long)(blocks * block));
        BH_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL((bh::dopri5_lane_kernel<Sys, T, kPeaks, true>), dim3((unsigned)blocks),
                           dim3(block), 0, stream, args);
    }
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

// two-phase launch (args.phase_iters > 0): counter reset, phase 1 over every
// trajectory, phase 2 over at most n suspended states (lanes past the
// counter leave at once), no host synchronisation
template <class Sys, class T>
int32_t launch_phased(const bh::LaneArgs &args, hipStream_t stream) {
    constexpr int block = 256;
    const int64_t blocks = (args.n + block - 1) / block;
    if (blocks > 2147483647LL) return fail(-5, "n too large for a two-phase launch");
    hipLaunchKernelGGL(bh::set_counter_kernel, dim3(1), dim3(64), 0, stream, args.counter, 0ull);
    BH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((bh::dopri5_phase_kernel<Sys, T, 1>), dim3((unsigned)blocks), dim3(block), 0,
                       stream, args);
    BH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((bh::dopri5_phase_kernel<Sys, T, 2>), dim3((unsigned)blocks), dim3(block), 0,
                       stream, args);
    BH_HIP_CHECK(hipGetLastError());
    return 0;
}

template <class Sys>
int32_t dispatch_dtype(int32_t dtype, const bh::LaneArgs &args, int32_t max_lanes,
                       hipStream_t s) {
    const bool peaks = args.peak_counts != nullptr;
    if (args.phase_iters > 0) {
        if (dtype == BH_F32) return launch_phased<Sys, float>(args, s);
        if (dtype == BH_F64) return launch_phased<Sys, double>(args, s);
        return fail(-5, "dtype must be BH_F32 or BH_F64, got %d", dtype);
    }
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

template <class Sys>
int64_t state_bytes_of(int32_t dtype) {
    return dtype == BH_F64 ? bh::lane_state_bytes<Sys, double>() : bh::lane_state_bytes<Sys, float>();
}

// workspace of a two-phase launch: counter, trajectory index and state record
// per possibly suspended lane (-1 for an unknown system)
int64_t phased_workspace(int32_t sys, int32_t dtype, int32_t dim, int64_t n) {
    int64_t sb;
    switch (sys) {
    case BH_SYS_DECAY:
        sb = dim == 1 ? state_bytes_of<bh::Decay<1>>(dtype) : state_bytes_of<bh::Decay<2>>(dtype);
        break;
    case BH_SYS_LINEAR2: sb = state_bytes_of<bh::Linear2>(dtype); break;
    case BH_SYS_PENDULUM: sb = state_bytes_of<bh::Pendulum>(dtype); break;
    case BH_SYS_DUFFING: sb = state_bytes_of<bh::Duffing>(dtype); break;
    case BH_SYS_LORENZ: sb = state_bytes_of<bh::Lorenz>(dtype); break;
    default: return -1;
    }
    const int64_t ic_bytes = (n * 8 + 255) / 256 * 256;
    return BH_WORKSPACE_BYTES + ic_bytes + n * sb;
}

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
                         int64_t workspace_bytes, int32_t max_lanes, void *stream,
                         int32_t phase_iters = 0) {
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
    if (phase_iters < 0) return fail(-5, "phase_iters must be >= 0");
    if (phase_iters > 0) {
        if (out_peak_counts) return fail(-5, "orbit peaks are not produced by two-phase launches");
        const int64_t need = phased_workspace(system_id, dtype, dim, n);
        if (workspace_bytes < need)
            return fail(-5, "a two-phase launch needs %lld workspace bytes", (long long)need);
    }

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
    a.phase_iters = phase_iters;
    a.state_ic = nullptr;
    a.state = nullptr;
    if (phase_iters > 0) {
        char *w = reinterpret_cast<char *>(workspace);
        a.state_ic = reinterpret_cast<int64_t *>(w + BH_WORKSPACE_BYTES);
        a.state = w + BH_WORKSPACE_BYTES + (n * 8 + 255) / 256 * 256;
    }
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

// Scratch of the allocation-free two-phase entry point
// (bh_solve_batch_phased_noalloc): the refill counter, the trajectory
// indices and the suspended lane records, laid out as the caller workspace
// of bh_solve_batch_phased.  Each device holds its own copy for the life of
// the process.  In host code `scratch` names a host-side stand-in, never the
// device memory: kernels get the address hipGetSymbolAddress returns for
// the current device.
static __device__ __align__(256) unsigned char scratch[BH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per device
int32_t scratch_on_current_device(void **out) {
    int dev = 0;
    BH_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(-5, "device %d: the static scratch table covers devices 0..%d", dev,
                    kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        BH_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(-6, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

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

int64_t bh_solve_batch_phased_workspace_bytes(int32_t system_id, int32_t dtype, int32_t dim,
                                              int64_t n) {
    g_err[0] = 0;
    if (dtype != BH_F32 && dtype != BH_F64) return fail(-5, "dtype must be BH_F32 or BH_F64");
    if (n < 0) return fail(-5, "n must be >= 0");
    if (system_id == BH_SYS_DECAY && dim != 1 && dim != 2)
        return fail(-5, "decay system supports dim 1 or 2, got %d", dim);
    const int64_t bytes = phased_workspace(system_id, dtype, dim, n);
    if (bytes < 0) return fail(-1, "system id %d has no per-lane device functor", system_id);
    return bytes;
}

int32_t bh_solve_batch_phased(int32_t system_id, int32_t dtype, const void *y0s, int64_t n,
                              int32_t dim, const double *params, int32_t n_params,
                              int64_t params_stride, double t0, const double *save_ts,
                              int32_t n_save, double rtol, double atol, int64_t max_steps,
                              uint32_t flags, double stop_threshold, int32_t steady_idx,
                              void *out_traj, double *out_feats, void *out_final,
                              int32_t *out_status, int64_t *out_nsteps, int32_t phase_iters,
                              void *workspace, int64_t workspace_bytes, void *stream) {
    if (phase_iters < 1) {
        g_err[0] = 0;
        return fail(-5, "phase_iters must be >= 1, got %d", phase_iters);
    }
    return solve_batch_impl(system_id, dtype, y0s, n, dim, params, n_params, params_stride, t0,
                            save_ts, n_save, rtol, atol, max_steps, flags, stop_threshold,
                            steady_idx, out_traj, out_feats, out_final, out_status, out_nsteps, 0,
                            nullptr, nullptr, workspace, workspace_bytes, 0, stream, phase_iters);
}

int32_t bh_solve_batch_phased_noalloc(int32_t system_id, int32_t dtype, const void *y0s, int64_t n,
                                      int32_t dim, const double *params, int32_t n_params,
                                      int64_t params_stride, double t0, const double *save_ts,
                                      int32_t n_save, double rtol, double atol, int64_t max_steps,
                                      uint32_t flags, double stop_threshold, int32_t steady_idx,
                                      void *out_traj, double *out_feats, void *out_final,
                                      int32_t *out_status, int64_t *out_nsteps,
                                      int32_t phase_iters, void *stream) {
    g_err[0] = 0;
    if (phase_iters < 1) return fail(-5, "phase_iters must be >= 1, got %d", phase_iters);
    int32_t rc = check_system(system_id, dim, n_params);
    if (rc) return rc;
    if (dtype != BH_F32 && dtype != BH_F64) return fail(-5, "dtype must be BH_F32 or BH_F64");
    if (n < 0) return fail(-5, "n must be >= 0, got %lld", (long long)n);
    if (n == 0) return 0;
    const int64_t need = phased_workspace(system_id, dtype, dim, n);
    if (need > (int64_t)BH_NOALLOC_SCRATCH_BYTES)
        return fail(-5,
                    "bh_solve_batch_phased_noalloc: n = %lld needs %lld scratch bytes, the static "
                    "buffer holds %lld (at most %lld trajectories of this system and dtype; use "
                    "bh_solve_batch_phased with a workspace)",
                    (long long)n, (long long)need, (long long)BH_NOALLOC_SCRATCH_BYTES,
                    (long long)bh_solve_batch_phased_noalloc_max_n(system_id, dtype, dim));
    void *ws = nullptr;
    rc = scratch_on_current_device(&ws);
    if (rc) return rc;
    // the counter is reset by a kernel inside the launch (no memset)
    return solve_batch_impl(system_id, dtype, y0s, n, dim, params, n_params, params_stride, t0,
                            save_ts, n_save, rtol, atol, max_steps, flags, stop_threshold,
                            steady_idx, out_traj, out_feats, out_final, out_status, out_nsteps, 0,
                            nullptr, nullptr, ws, (int64_t)BH_NOALLOC_SCRATCH_BYTES, 0, stream,
                            phase_iters);
}

int64_t bh_solve_batch_phased_noalloc_max_n(int32_t system_id, int32_t dtype, int32_t dim) {
    g_err[0] = 0;
    if (dtype != BH_F32 && dtype != BH_F64) return fail(-5, "dtype must be BH_F32 or BH_F64");
    if (system_id == BH_SYS_DECAY && dim != 1 && dim != 2)
        return fail(-5, "decay system supports dim 1 or 2, got %d", dim);
    const int64_t one = phased_workspace(system_id, dtype, dim, 1);
    if (one < 0) return fail(-1, "system id %d has no per-lane device functor", system_id);
    // largest n whose workspace fits (the index array rounds up to 256 bytes)
    const int64_t per = one - BH_WORKSPACE_BYTES - 256 + 8;
    int64_t hi = ((int64_t)BH_NOALLOC_SCRATCH_BYTES - BH_WORKSPACE_BYTES) / per;
    while (hi > 0 && phased_workspace(system_id, dtype, dim, hi) > (int64_t)BH_NOALLOC_SCRATCH_BYTES) hi--;
    return hi;
}

int32_t bh_autotune_lane(int32_t system_id, int32_t dtype, const void *y0s, int64_t n,
                         int32_t dim, const double *params, int32_t n_params,
                         int64_t params_stride, double t0, const double *save_ts, int32_t n_save,
                         double rtol, double atol, int64_t max_steps, uint32_t flags,
                         double stop_threshold, int32_t steady_idx, void *out_traj,
                         double *out_feats, void *out_final, int32_t *out_status,
                         int64_t *out_nsteps, void *workspace, int64_t workspace_bytes,
                         const int32_t *caps, int32_t n_caps, const int32_t *quorums,
                         int32_t n_quorums, const int32_t *phases, int32_t n_phases, int32_t reps,
                         int32_t *best_cap, int32_t *best_quorum, int32_t *best_phase,
                         float *timings_ms, void *stream) {
    g_err[0] = 0;
    if (n_caps < 0 || n_quorums < 0 || n_phases < 0) return fail(-5, "negative candidate count");
    if ((n_caps > 0 && !caps) || (n_quorums > 0 && !quorums) || (n_phases > 0 && !phases))
        return fail(-5, "null candidate array");
    const int32_t n_refill = n_caps * n_quorums;
    if (n_refill + n_phases < 1)
        return fail(-5, "no candidate: give caps and quorums, or phases");
    if (!best_cap || !best_quorum || !best_phase)
        return fail(-5, "best_cap, best_quorum and best_phase are required");
    if (reps < 1) return fail(-5, "reps must be >= 1, got %d", reps);
    for (int32_t i = 0; i < n_caps; i++)
        if (caps[i] < 0) return fail(-5, "caps[%d] = %d: caps are max_lanes values >= 0", i, caps[i]);
    for (int32_t j = 0; j < n_quorums; j++)
        if (quorums[j] < 1 || quorums[j] > 64)
            return fail(-5, "quorums[%d] = %d: a refill quorum is 1..64", j, quorums[j]);
    for (int32_t k = 0; k < n_phases; k++)
        if (phases[k] < 1) return fail(-5, "phases[%d] = %d: a phase budget is >= 1", k, phases[k]);
    *best_cap = n_refill ? caps[0] : 0;
    *best_quorum = n_refill ? quorums[0] : 0;
    *best_phase = n_refill ? 0 : phases[0];
    if (timings_ms)
        for (int32_t k = 0; k < n_refill + n_phases; k++) timings_ms[k] = 0.0f;
    if (n == 0) return 0;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        return fail(-6, "hipEventCreate failed");
    }
    const uint32_t base = flags & ~(0x7fu << BH_REFILL_QUORUM_SHIFT);
    // one untimed launch, then reps timed ones; every launch solves the whole
    // problem (the outputs do not depend on the launch shape)
    auto time_one = [&](int32_t cap, uint32_t f, int32_t phase, float *t_min) -> int32_t {
        *t_min = INFINITY;
        for (int32_t r = 0; r <= reps; r++) {
            if (r > 0 && hipEventRecord(ev[0], s) != hipSuccess) return fail(-6, "hipEventRecord failed");
            const int32_t rc = solve_batch_impl(
                system_id, dtype, y0s, n, dim, params, n_params, params_stride, t0, save_ts, n_save,
                rtol, atol, max_steps, f, stop_threshold, steady_idx, out_traj, out_feats, out_final,
                out_status, out_nsteps, 0, nullptr, nullptr, workspace, workspace_bytes, cap, stream,
                phase);
            if (rc) return rc;
            if (r == 0) continue;
            float ms = 0.0f;
            if (hipEventRecord(ev[1], s) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess)
                return fail(-6, "timing the launch with HIP events failed");
            *t_min = fminf(*t_min, ms);
        }
        return 0;
    };
    int32_t rc = 0;
    float best = INFINITY;
    for (int32_t i = 0; i < n_caps && rc == 0; i++) {
        for (int32_t j = 0; j < n_quorums && rc == 0; j++) {
            float t = 0.0f;
            rc = time_one(caps[i], base | BH_REFILL_QUORUM(quorums[j]), 0, &t);
            if (rc) break;
            if (timings_ms) timings_ms[i * n_quorums + j] = t;
            if (t < best) {
                best = t;
                *best_cap = caps[i];
                *best_quorum = quorums[j];
                *best_phase = 0;
            }
        }
    }
    for (int32_t k = 0; k < n_phases && rc == 0; k++) {
        float t = 0.0f;
        rc = time_one(0, flags, phases[k], &t);
        if (rc) break;
        if (timings_ms) timings_ms[n_refill + k] = t;
        if (t < best) {
            best = t;
            *best_cap = 0;
            *best_quorum = 0;
            *best_phase = phases[k];
        }
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    return rc;
}

}  // extern "C"
