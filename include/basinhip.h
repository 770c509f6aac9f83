/*
 * basinhip.h -- C-ABI of libbasinhip.so, the MI355X (gfx950) Dormand-Prince
 * basin-stability integration path.
 *
 * Every array argument is DEVICE memory owned by the caller; calls are
 * asynchronous on `stream` (a hipStream_t passed as void*; NULL = the null
 * stream).  Return codes extend the reference's zigode ErrorCode
 * (src/zigode/src/c_api.zig:17-23):
 *   0 success, -1 invalid system id, -2 allocation failed, -3 solve failed,
 *   -4 (unused: thread spawn), -5 invalid argument, -6 HIP runtime error.
 * bh_last_error() returns a message for the last non-zero code of the
 * calling thread.
 */
#ifndef BASINHIP_H
#define BASINHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- identifiers ------------------------------------------------------ */
/* Right-hand sides with a device functor.  Parameter vectors follow the
   reference's TypedDict order (src/pybasin/solvers/torch_ode_system.py:57-66). */
enum {
    BH_SYS_DECAY = 0,    /* y' = p0*y, S = 1 or 2  (tests/unit/solvers/test_solver.py:16-18) */
    BH_SYS_LINEAR2 = 1,  /* y' = [[p0,p1],[p2,p3]] y, S = 2 (test system)              */
    BH_SYS_PENDULUM = 2, /* [alpha, T, K], S = 2  (case_studies/pendulum/pendulum_ode.py:23-40) */
    BH_SYS_DUFFING = 3,  /* [delta, k3, A], S = 2 (case_studies/duffing_oscillator/duffing_ode.py:26-42) */
    BH_SYS_LORENZ = 4,   /* [sigma, r, b], S = 3, freeze mask |y|inf >= 200
                            (case_studies/lorenz/lorenz_ode.py:28-64) */
    BH_SYS_ROSSLER100 = 5 /* [a, b, c, K], S = 300; see bh_solve_rossler */
};
enum { BH_F32 = 0, BH_F64 = 1 };

/* per-lane status written to out_status */
enum {
    BH_LANE_OK = 0,
    BH_LANE_MAX_STEPS = 1,    /* step budget exhausted; remaining saves hold the last state */
    BH_LANE_DT_UNDERFLOW = 2, /* t + dt == t (torchdiffeq's "underflow in dt")             */
    BH_LANE_NONFINITE = 3,    /* NaN/Inf state or error ratio                              */
    BH_LANE_DIVERGED = 4      /* |y|inf >= stop_threshold; lane stopped (freeze or Inf fill) */
};

/* flags */
#define BH_FLAG_STOP 0x1u     /* enable the per-lane divergence stop at stop_threshold    */
#define BH_FLAG_INF_FILL 0x2u /* after a stop, remaining saves are +Inf (JaxSolver Event
                                 semantics) instead of the frozen state (masked-RHS semantics) */
/* out_nsteps[i] = attempted steps | (steps whose span held a save point << 32)
   instead of the attempted steps alone (the work count of the dense output,
   for roofline accounting; the per-lane entry points) */
#define BH_FLAG_COUNT_SAVE_STEPS 0x4u
/* out_feats holds BH_NFEAT_TREND values per state instead of BH_NFEAT: the
   four statistics and the least-squares slope of the window against the
   sample position 0 .. N'-1 (linear_trend attr "slope",
   src/pybasin/ts_torch/calculators/torch_features_trend.py:11-27) */
#define BH_FLAG_TREND 0x8u
/* Refill quorum, bits 8-14 of flags (a launch-shape choice; results do not
   depend on it): finished lanes of a wave take new trajectories once at
   least q of them wait (or every lane still holding one has finished).
   q = 1 refills each lane at once (best when trajectories end at very
   different times, e.g. divergence stops); q = 64 refills whole waves (keeps
   a wave's lanes in step, so save-point work is shared).  0 = the default. */
#define BH_REFILL_QUORUM_SHIFT 8u
#define BH_REFILL_QUORUM(q) (((uint32_t)(q) & 0x7fu) << BH_REFILL_QUORUM_SHIFT)

/* number of statistics per state written to out_feats: mean, std (ddof 0),
   min, max over save indices [steady_idx, n_save) */
#define BH_NFEAT 4
/* with BH_FLAG_TREND: mean, std, min, max, slope */
#define BH_NFEAT_TREND 5
/* bytes of device workspace bh_solve_batch needs (a lane-refill counter) */
#define BH_WORKSPACE_BYTES 256

int32_t bh_version(void);
const char *bh_last_error(void);

/*
 * bh_solve_batch -- integrate n independent trajectories, one per GPU lane,
 * with per-lane adaptive Dormand-Prince 5(4) under torchdiffeq's step rules
 * (replaces the integration inside TorchDiffEqSolver._integrate,
 * src/pybasin/solvers/torchdiffeq_solver.py:97-134, and the batch entry of
 * the reference's native solver, solve_batch in src/zigode/src/c_api.zig:193-314).
 *
 *   y0s            (n, dim) in dtype
 *   params         (n, n_params) float64 with row stride params_stride,
 *                  or one shared vector when params_stride == 0
 *   t0             integration start (t_span[0]; torchdiffeq_solver.py:122-125)
 *   save_ts        (n_save) float64: the save grid in the state dtype,
 *                  promoted to float64 (base.py:113-127)
 *   rtol, atol     tolerances; max_steps: attempted-step budget per lane
 *   flags          BH_FLAG_*; stop_threshold used with BH_FLAG_STOP
 *   steady_idx     first save index of the feature window (filter_time,
 *                  feature_extractor.py:102-130); >= n_save disables features
 *   out_traj       (n_save, n, dim) in dtype, or NULL (trajectories not kept)
 *   out_feats      (n, dim, BH_NFEAT) float64 ((n, dim, BH_NFEAT_TREND) with
 *                  BH_FLAG_TREND), or NULL
 *   out_final      (n, dim) in dtype: state at the last save time, or NULL
 *   out_status     (n) int32, or NULL; out_nsteps (n) int64 attempted steps, or NULL
 *   workspace      >= BH_WORKSPACE_BYTES of device memory (refill counter)
 *   max_lanes      cap on resident lanes (0 = fill the device); lanes that
 *                  finish a trajectory take the next one (work stealing)
 */
int32_t bh_solve_batch(int32_t system_id, int32_t dtype, const void *y0s, int64_t n,
                       int32_t dim, const double *params, int32_t n_params,
                       int64_t params_stride, double t0, const double *save_ts,
                       int32_t n_save, double rtol, double atol, int64_t max_steps,
                       uint32_t flags, double stop_threshold, int32_t steady_idx,
                       void *out_traj, double *out_feats, void *out_final,
                       int32_t *out_status, int64_t *out_nsteps, void *workspace,
                       int64_t workspace_bytes, int32_t max_lanes, void *stream);

/*
 * bh_solve_batch_peaks -- bh_solve_batch plus the orbit peaks of every state
 * over the feature window [steady_idx, n_save): strict local maxima with
 * support 1 (a value greater than both neighbours; the first and last window
 * points are never peaks), in time order.  Replaces extract_orbit_data /
 * extract_peak_values on the saved trajectories (src/pybasin/utils.py:254-303,
 * src/pybasin/ts_torch/calculators/torch_features_pattern.py:64-150) with
 * peak_support = 1.
 *
 *   peak_cap         rows of out_peak_vals
 *   out_peak_vals    (peak_cap, n, dim) in dtype: peak k of (ic, state) at
 *                    [k][ic][state] for k < min(count, peak_cap); other
 *                    entries are not written
 *   out_peak_counts  (n, dim) int32: number of peaks (may exceed peak_cap;
 *                    the caller re-runs with a larger cap to get them all)
 */
int32_t bh_solve_batch_peaks(int32_t system_id, int32_t dtype, const void *y0s, int64_t n,
                             int32_t dim, const double *params, int32_t n_params,
                             int64_t params_stride, double t0, const double *save_ts,
                             int32_t n_save, double rtol, double atol, int64_t max_steps,
                             uint32_t flags, double stop_threshold, int32_t steady_idx,
                             void *out_traj, double *out_feats, void *out_final,
                             int32_t *out_status, int64_t *out_nsteps, int32_t peak_cap,
                             void *out_peak_vals, int32_t *out_peak_counts, void *workspace,
                             int64_t workspace_bytes, int32_t max_lanes, void *stream);

/*
 * bh_solve_batch_phased -- bh_solve_batch as two launches with compaction in
 * between, for batches whose trajectories end at very different times:
 * phase 1 gives every trajectory its own lane for at most phase_iters
 * attempted steps; trajectories still integrating then store their whole
 * lane state in the workspace, packed densely, and phase 2 resumes them on
 * consecutive lanes, so the long trajectories run in full waves instead of
 * a few live lanes per wave.  Results equal bh_solve_batch bit for bit (a
 * trajectory's arithmetic does not depend on the lane or launch that runs
 * it).  Arguments as bh_solve_batch, with phase_iters (>= 1) in place of
 * max_lanes; no orbit peaks; asynchronous (no host synchronisation).
 * Replaces the same reference interface as bh_solve_batch (the integration
 * inside TorchDiffEqSolver._integrate, src/pybasin/solvers/torchdiffeq_solver.py:97-134).
 *
 *   workspace   >= bh_solve_batch_phased_workspace_bytes(system_id, dtype,
 *               dim, n) device bytes (counter, and index + state record of
 *               every trajectory that may be suspended); that query returns
 *               a negative error code (bh_last_error) for an unknown system,
 *               a dtype other than BH_F32 / BH_F64 or n < 0
 */
int64_t bh_solve_batch_phased_workspace_bytes(int32_t system_id, int32_t dtype, int32_t dim,
                                              int64_t n);
int32_t bh_solve_batch_phased(int32_t system_id, int32_t dtype, const void *y0s, int64_t n,
                              int32_t dim, const double *params, int32_t n_params,
                              int64_t params_stride, double t0, const double *save_ts,
                              int32_t n_save, double rtol, double atol, int64_t max_steps,
                              uint32_t flags, double stop_threshold, int32_t steady_idx,
                              void *out_traj, double *out_feats, void *out_final,
                              int32_t *out_status, int64_t *out_nsteps, int32_t phase_iters,
                              void *workspace, int64_t workspace_bytes, void *stream);

/*
 * bh_solve_batch_phased_noalloc -- bh_solve_batch_phased that allocates
 * nothing and takes no workspace: its counter, trajectory indices and
 * suspended lane records live in a static device buffer of
 * BH_NOALLOC_SCRATCH_BYTES (64 MiB) that each GPU the library runs on holds
 * until the process ends.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The caller passes
 * the inputs and outputs as for bh_solve_batch_phased.  n is limited to
 * bh_solve_batch_phased_noalloc_max_n(system_id, dtype, dim) (about 360 k
 * Duffing or pendulum trajectories in float32); larger n is rejected with
 * -5 before any device work.  Every call on a GPU shares that GPU's buffer:
 * calls that overlap on one GPU (concurrent streams or host threads) race;
 * serialise them.  The counter is reset inside each launch by a kernel.
 */
#define BH_NOALLOC_SCRATCH_BYTES (64ll << 20)
int32_t bh_solve_batch_phased_noalloc(int32_t system_id, int32_t dtype, const void *y0s, int64_t n,
                                      int32_t dim, const double *params, int32_t n_params,
                                      int64_t params_stride, double t0, const double *save_ts,
                                      int32_t n_save, double rtol, double atol, int64_t max_steps,
                                      uint32_t flags, double stop_threshold, int32_t steady_idx,
                                      void *out_traj, double *out_feats, void *out_final,
                                      int32_t *out_status, int64_t *out_nsteps,
                                      int32_t phase_iters, void *stream);
/* largest n bh_solve_batch_phased_noalloc accepts for this system, dtype and
   dim (a negative error code for an unknown system or dtype) */
int64_t bh_solve_batch_phased_noalloc_max_n(int32_t system_id, int32_t dtype, int32_t dim);

/*
 * bh_autotune_lane -- the launch-shape tuner of the per-lane solver as one
 * call: solves the caller's problem once per candidate -- one untimed
 * launch, then `reps` launches timed with HIP events on `stream` -- and
 * returns the fastest.  Candidates: every (resident-lane cap, refill quorum)
 * pair of bh_solve_batch, and every phase budget of bh_solve_batch_phased.
 * The results do not depend on the launch shape, so on return the outputs
 * hold the problem's solution.  Blocks until the timed launches end.  The
 * reference has no counterpart (its solvers have no launch shape to
 * choose); the arguments up to workspace_bytes are those of bh_solve_batch,
 * except that the refill-quorum bits of flags are replaced by each
 * candidate's (with phase candidates the workspace must hold
 * bh_solve_batch_phased_workspace_bytes).
 *
 *   caps        (n_caps) host array of max_lanes values (0 = fill the device)
 *   quorums     (n_quorums) host array of refill quorums, each 1..64
 *   phases      (n_phases) host array of phase_iters values (>= 1), or NULL
 *               with n_phases = 0
 *   reps        timed launches per candidate (>= 1); a candidate's time is
 *               the fastest of them
 *   best_cap, best_quorum, best_phase   out: the fastest candidate (a
 *               phase winner has cap and quorum 0; a refill winner phase 0)
 *   timings_ms  (n_caps * n_quorums + n_phases) host floats: the refill
 *               pairs [cap][quorum], then the phases; or NULL
 */
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
                         float *timings_ms, void *stream);

/*
 * bh_solve_rossler -- the 100-node Rossler network (BH_SYS_ROSSLER100, state
 * layout [x_0..x_99, y_0..y_99, z_0..z_99]), 16 trajectories per workgroup,
 * coupling K (L x) as a (16 x 100) . (100 x 100) FP32 matrix-core product per
 * stage.  Replaces the integration of RosslerNetworkJaxODE
 * (case_studies/rossler_network/rossler_network_jax_ode.py:53-79) under the
 * reference's solver call (src/pybasin/solvers/base.py:129-183); only the
 * state at t_end is produced (the synchronisation features use the final
 * step, synchronization_feature_extractor.py:73-106).
 *
 *   dtype          BH_F32 only
 *   y0s            (n, 300) float32
 *   params         (n, 4) float64 [a, b, c, K] with row stride params_stride
 *                  (0 = one shared vector)
 *   laplacian      (100, 100) float32, row-major, symmetric
 *   t0, t_end      integration start and the time of the returned state
 *   out_final      (n, 300) float32; status / nsteps as in bh_solve_batch
 *   max_tiles      cap on resident workgroups (0 = fill the device)
 */
int32_t bh_solve_rossler(int32_t dtype, const void *y0s, int64_t n, const double *params,
                         int64_t params_stride, const float *laplacian, double t0, double t_end,
                         double rtol, double atol, int64_t max_steps, uint32_t flags,
                         double stop_threshold, void *out_final, int32_t *out_status,
                         int64_t *out_nsteps, void *workspace, int64_t workspace_bytes,
                         int32_t max_tiles, void *stream);

/*
 * bh_solve_rossler_edges -- as bh_solve_rossler with the coupling given by
 * the reference's own edge arrays (RosslerNetworkJaxODE(params, n, edges_i,
 * edges_j), rossler_network_jax_ode.py:45-79, with the arrays of
 * rossler_network_topology.py:26-48): (L x)_i = sum over the entries e with
 * edges_i[e] = i of (x_i - x_edges_j[e]), evaluated as deg_i x_i - sum x_j on
 * the vector units (1,600 FLOP per evaluation for the 800 entries of the
 * reference's graph, against 20,000 for the dense product).
 *
 *   edges_i, edges_j  (n_edges) HOST int32 arrays of node indices in [0, 100)
 *                     (both directions of an undirected edge are two entries);
 *                     at most 12 entries per node (-5 otherwise: use
 *                     bh_solve_rossler with the dense Laplacian)
 *   other arguments   as bh_solve_rossler
 */
int32_t bh_solve_rossler_edges(int32_t dtype, const void *y0s, int64_t n, const double *params,
                               int64_t params_stride, const int32_t *edges_i,
                               const int32_t *edges_j, int64_t n_edges, double t0, double t_end,
                               double rtol, double atol, int64_t max_steps, uint32_t flags,
                               double stop_threshold, void *out_final, int32_t *out_status,
                               int64_t *out_nsteps, void *workspace, int64_t workspace_bytes,
                               int32_t max_tiles, void *stream);

/* ---- learned surrogate (no reference counterpart: the reference has no
 * learned classifier; these kernels train and run an MLP on labels produced
 * by bh_solve_batch).  bf16 arguments are bfloat16 bit patterns. ---------- */

/* C (M x N, float32, row stride ldc) (+)= A . B on the matrix cores, with
   A(m, k) = A[m*sam + k*sak] and B(k, n) = B[k*sbk + n*sbn] in bf16 */
int32_t bh_gemm_bf16(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk,
                     int64_t sbn, float *C, int64_t ldc, int64_t M, int64_t N, int64_t K,
                     int32_t accumulate, void *stream);
/* Z + bias per column; optional copy of that pre-activation (zsave), SiLU
   (silu != 0) or identity written as bf16 (out, row stride ldo) and/or f32 */
int32_t bh_mlp_bias_act(const float *Z, int64_t M, int64_t N, const float *bias, int32_t silu,
                        float *zsave, void *out_bf16, int64_t ldo, float *out_f32, void *stream);
/* dZ = dA * SiLU'(Z), written as float32 and bf16 */
int32_t bh_mlp_silu_bwd(const float *dA, const float *Z, int64_t count, float *dZ,
                        void *dZ_bf16, void *stream);
/* two-class softmax cross-entropy: gradient of the mean loss w.r.t. the
   logits (M x 2, float32 and bf16) and the per-row loss */
int32_t bh_mlp_softmax_ce(const float *logits, const int32_t *labels, int64_t M, float *dlogits,
                          void *dlogits_bf16, float *loss, void *stream);
/* column sums of G (M x N float32) */
int32_t bh_mlp_colsum(const float *G, int64_t M, int64_t N, float *out, void *stream);
/* Adam step `step` (1-based) on float32 weights; refreshes the bf16 copy if given */
int32_t bh_mlp_adam(float *W, const float *G, float *m, float *v, int64_t count, float lr,
                    float beta1, float beta2, float eps, int64_t step, void *W_bf16, void *stream);
/* (M x 3) float32 inputs (theta, omega, T) -> (M x 16) bf16 rows
   [sin x0, cos x0, (x1-c1)s1, (x2-c2)s2, 0 * (x0-c0)s0, (x1^2/2 - cos x0)s1^2, 0, ...]
   (the raw angle column is kept at zero: the label is 2 pi periodic in x0) */
int32_t bh_mlp_prep_input(const float *x, int64_t M, float c0, float s0, float c1, float s1,
                          float c2, float s2, void *out_bf16, void *stream);
/* fused inference: (M x 3) float32 rows (theta, omega, T) -> label and max
   softmax probability.  w_packed: the four weight matrices (zero-padded to
   multiples of 32 rows and 16 columns) in fragment order [kstep][colblock]
   [lane][8] bf16, element (s, c, lane, e) = W[32 s + 8 (lane / 16) + e]
   [16 c + lane % 16]: W1 1x16, W2 8x16, W3 8x16, W4 8x1 fragments of 64
   lanes; biases: b1, b2, b3 (256 each), b4 (2) float32, concatenated; input
   scaling as bh_mlp_prep_input */
int32_t bh_mlp_infer(const float *x, int64_t M, const void *w_packed, const float *biases,
                     float c0, float s0, float c1, float s1, float c2, float s2, int32_t *label,
                     float *conf, void *stream);
/* logits (M x 2) + bias -> argmax label and max softmax probability */
int32_t bh_mlp_classify(const float *Z, const float *bias, int64_t M, int32_t *label,
                        float *conf, void *stream);

/* One training step of the surrogate on a batch, as one call: forward
   (bh_mlp_prep_input, three GEMM + bias + SiLU layers, GEMM + bias to
   logits), mean softmax cross-entropy, backward, and Adam on every weight
   and bias -- the same kernels, in the same arithmetic, as the separate
   entry points above.  No counterpart in the reference (learned surrogate).
     x (M, 3) float32, labels (M) int32 in {0, 1}; norm: 6 host floats
     (c0, s0, c1, s1, c2, s2 of bh_mlp_prep_input); hidden: width H
     W, W_bf16, b, mW, vW, mb, vb: host arrays of 4 device pointers, one per
       layer -- weights (K_in x N_out) float32 row-major with K_in 16, H, H,
       H and N_out H, H, H, 2; their bfloat16 copies (rewritten by the
       update); biases (N_out); Adam moments shaped like each
     step: Adam step number (>= 1, for the bias correction)
     loss_out (M) float32 per-row loss, or NULL
     workspace: >= bh_mlp_train_workspace_bytes(M, hidden) device bytes */
int64_t bh_mlp_train_workspace_bytes(int64_t M, int32_t hidden);
int32_t bh_mlp_train_step(const float *x, const int32_t *labels, int64_t M, const float *norm,
                          int32_t hidden, float *const *W, void *const *W_bf16, float *const *b,
                          float *const *mW, float *const *vW, float *const *mb, float *const *vb,
                          float lr, float beta1, float beta2, float eps, int64_t step,
                          float *loss_out, void *workspace, int64_t workspace_bytes,
                          void *stream);

/* Order statistics of the feature window of a kept trajectory buffer
   (n_save, n, dim) in dtype: out (n_ranks, n, dim) in dtype holds, per
   (trajectory, state), the value of 0-based rank ranks[r] (device int64,
   valid in [0, n_save - start)) among rows [start, n_save); NaN if the window
   holds a NaN or if ranks[r] lies outside the valid range.  Median = rank (N - 1) / 2 (torch.median, lower middle);
   quantiles interpolate two ranks (torch.quantile).  Replaces the windowed
   calculators of src/pybasin/ts_torch/calculators/torch_features_statistical.py
   (median 30-38, quantile 154-157). */
int32_t bh_window_order_stats(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                              int32_t dim, int64_t start, const int64_t *ranks, int32_t n_ranks,
                              void *out, void *stream);

/* Autocorrelation periodicity of each (trajectory, state) series of the
   window rows [start, n_save) of a kept trajectory buffer (n_save, n, dim) in
   dtype: out_strength and out_period (n, dim) in dtype hold the height and
   lag of the first local maximum (scipy's definition: plateaus marked at
   their midpoint) of the normalised autocorrelation at a lag >= min_lag
   above peak_threshold, 0 and 0 if there is none.  Windows of up to 15,872
   (float) / 7,936 (double) rows.  Replaces
   src/pybasin/ts_torch/calculators/torch_features_autocorrelation.py:268-331
   (defaults min_lag 2, peak_threshold 0.3). */
int32_t bh_window_autocorr_periodicity(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                                       int32_t dim, int64_t start, int32_t min_lag,
                                       double peak_threshold, void *out_strength,
                                       void *out_period, void *stream);

/* Spectral frequency ratio of each series of the same window: f2 / f1 for
   the dominant rfft bin f1 of the centred series and the strongest bin f2
   at least two bins from it; 0 when the largest power is below 1e-10.  out
   (n, dim) in dtype; same window limits.  Replaces
   src/pybasin/ts_torch/calculators/torch_features_frequency.py:390-441. */
int32_t bh_window_spectral_ratio(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                                 int32_t dim, int64_t start, void *out_ratio, void *stream);

/* Raw accumulators of the comprehensive window features: for each series of
   the window rows [start, n_save) of a kept trajectory buffer (n_save, n,
   dim) in dtype, out (BH_WF_SLOTS, n, dim) float64 holds the slots below
   (out[k][ic][state]); a window holding a non-finite value gives NaN in every
   slot.  Indices are 0-based window positions; sums run in double.  The
   named features of the reference's comprehensive set
   (src/pybasin/ts_torch/settings.py:223-340) follow from them on the host:
   sum_values, length, abs_energy, kurtosis, skewness, variation_coefficient
   (torch_features_statistical.py:20-162), absolute_sum_of_changes,
   mean_abs_change, mean_change, mean_second_derivative_central
   (torch_features_change.py), count_above / count_below (t = 0),
   count_above_mean, count_below_mean, count_value (0, 1, -1), count_in_range
   ([-1, 1], [-1e12, 0], [0, 1e12]) (torch_features_count.py),
   has_duplicate_max / _min, has_variance_larger_than_standard_deviation,
   large_standard_deviation (torch_features_boolean.py),
   first/last_location_of_maximum/minimum, index_mass_quantile
   (q = 0.1 .. 0.9 without 0.5) (torch_features_location.py),
   longest_strike_above/below_mean, number_crossing_m (m = 0, -1, 1),
   number_peaks (n = 1, 3, 5) (torch_features_pattern.py:10-108),
   autocorrelation (lags 0..9) (torch_features_autocorrelation.py:13-28),
   cid_ce (torch_features_entropy_complexity.py:182-190), c3,
   time_reversal_asymmetry_statistic (lags 1..3), energy_ratio_by_chunks (10
   segments) (torch_features_advanced.py:84-123), ratio_beyond_r_sigma and
   mean_n_absolute_max (n <= 7) (torch_features_statistical.py:165-181),
   linear_trend and linear_trend_timewise (slope, intercept, rvalue, pvalue,
   stderr) (torch_features_trend.py:11-58), partial_autocorrelation (lags
   0..9, Durbin-Levinson over the autocovariances)
   (torch_features_autocorrelation.py:31-69), binned_entropy (max_bins 10)
   (torch_features_entropy_complexity.py:58-94), symmetry_looking (with the
   median of bh_window_order_stats). */
enum {
    BH_WF_N = 0,             /* window length N'                                   */
    BH_WF_SUM = 1,           /* sum x                                              */
    BH_WF_SUMSQ = 2,         /* sum x^2                                            */
    BH_WF_MIN = 3,
    BH_WF_MAX = 4,
    BH_WF_FIRST_MAX = 5,     /* first index of the maximum                         */
    BH_WF_LAST_MAX = 6,
    BH_WF_FIRST_MIN = 7,
    BH_WF_LAST_MIN = 8,
    BH_WF_ABS_CHANGE = 9,    /* sum |x[i+1] - x[i]|                                */
    BH_WF_SQ_CHANGE = 10,    /* sum (x[i+1] - x[i])^2                              */
    BH_WF_X_FIRST = 11,      /* x[0], x[1], x[N'-2], x[N'-1]                       */
    BH_WF_X_SECOND = 12,
    BH_WF_X_SECOND_LAST = 13,
    BH_WF_X_LAST = 14,
    BH_WF_COUNT_POS = 15,    /* #(x > 0), #(x < 0), #(x == 0), #(x == 1), #(x == -1) */
    BH_WF_COUNT_NEG = 16,
    BH_WF_COUNT_ZERO = 17,
    BH_WF_COUNT_ONE = 18,
    BH_WF_COUNT_MINUS_ONE = 19,
    BH_WF_RANGE_UNIT = 20,   /* #(-1 <= x <= 1), #(-1e12 <= x <= 0), #(0 <= x <= 1e12) */
    BH_WF_RANGE_NONPOS = 21,
    BH_WF_RANGE_NONNEG = 22,
    BH_WF_CROSS_0 = 23,      /* #(i: (x[i] > m) != (x[i+1] > m)) for m = 0, -1, 1  */
    BH_WF_CROSS_M1 = 24,
    BH_WF_CROSS_P1 = 25,
    BH_WF_CHUNK = 26,        /* 10 slots: sum x^2 over [j s, (j+1) s), s = N' / 10 */
    BH_WF_C3 = 36,           /* 3 slots: sum x[i] x[i+l] x[i+2l], l = 1..3         */
    BH_WF_TREV = 39,         /* 3 slots: sum x[i+2l]^2 x[i+l] - x[i+l] x[i]^2      */
    BH_WF_SUM_ABS = 42,      /* sum |x|                                            */
    BH_WF_TOP_ABS = 43,      /* 7 slots: the largest |x|, descending (-1: none)    */
    BH_WF_PEAKS_1 = 50,      /* strict maxima of support 1, 3, 5                   */
    BH_WF_PEAKS_3 = 51,
    BH_WF_PEAKS_5 = 52,
    BH_WF_CM2 = 53,          /* sum (x - mean)^2, ^3, ^4                           */
    BH_WF_CM3 = 54,
    BH_WF_CM4 = 55,
    BH_WF_SS_TX = 56,        /* sum (i - (N'-1)/2) (x - mean)                      */
    BH_WF_COUNT_ABOVE_MEAN = 57,
    BH_WF_COUNT_BELOW_MEAN = 58,
    BH_WF_STRIKE_ABOVE = 59, /* longest run of x > mean, of x < mean               */
    BH_WF_STRIKE_BELOW = 60,
    BH_WF_COUNT_MAX = 61,    /* #(x == max), #(x == min)                           */
    BH_WF_COUNT_MIN = 62,
    BH_WF_BEYOND = 63,       /* 10 slots: #(|x - mean| > r sd), r = 0.5, 1, 1.5, 2,
                                2.5, 3, 5, 6, 7, 10 (sd: population)               */
    BH_WF_AUTOCOV = 73,      /* 9 slots: sum (x[i] - mean)(x[i+l] - mean), l = 1..9 */
    BH_WF_MASS = 82,         /* 8 slots: first i with sum_{j<=i} |x[j]| >= q sum |x|,
                                q = 0.1, 0.2, 0.3, 0.4, 0.6, 0.7, 0.8, 0.9          */
    BH_WF_BINS = 90,         /* 10 slots: binned_entropy counts, bin = trunc((x - min)
                                / (range + 1e-10) * (10 - 1e-6)) in the series dtype,
                                range 0 taken as 1                                  */
    BH_WF_SLOTS = 100
};
int32_t bh_window_features(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                           int32_t dim, int64_t start, double *out, void *stream);

/* The rfft of each series of the same window (torch.fft.rfft of the raw
   window, as the reference's fft_coefficient, spkt_welch_density and
   fft_aggregated take it; src/pybasin/ts_torch/calculators/
   torch_features_frequency.py:11-70), in double: out_coeff (n_coeff, 2, n,
   dim) float64 holds the real and imaginary parts of bins 0 .. n_coeff - 1
   (0 past the last bin w / 2); out_agg (5, n, dim) float64, or NULL, holds the
   centroid, variance, skew and kurtosis of the normalised power spectrum
   p_b = P_b / (sum P + 1e-10) over all w / 2 + 1 bins (skew over
   var^1.5 + 1e-10, kurtosis over var^2 + 1e-10) and its entropy
   -sum c log c, c = max(p_b, 1e-10) (fourier_entropy,
   torch_features_entropy_complexity.py:97-107).  0 <= n_coeff <= 4096; windows whose rows and
   (with out_agg) four power rows fit 62 KB of LDS (15,872 float32 rows,
   3,173 with out_agg). */
int32_t bh_window_spectrum(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                           int32_t dim, int64_t start, int32_t n_coeff, double *out_coeff,
                           double *out_agg, void *stream);

/* Rows of bh_window_trends' output (each (n, dim) float64). */
enum {
    BH_TR_ALT = 0,        /* 48: agg_linear_trend, (size * 4 + agg) * 4 + attr; sizes 5, 10,
                             50; aggs max, min, mean, var; attrs rvalue, intercept, slope,
                             stderr */
    BH_TR_FRIED = 48,     /* 4: friedrich_coefficients (m = 3), coefficients of x^3 .. x^0 */
    BH_TR_LANGEVIN = 52,  /* max_langevin_fixed_point (m = 3) */
    BH_TR_BENFORD = 53,   /* benford_correlation */
    BH_TR_CWT = 54,       /* 60: cwt_coefficients, width index * 15 + coeff; widths 2, 5, 10,
                             20; coeff 0..14 */
    BH_TR_SLOTS = 114
};

/* Chunk trends, polynomial fits, leading digits and wavelet coefficients of
   each series of the same window, one lane per series in double (out
   (BH_TR_SLOTS or BH_TR_CWT rows, n, dim) float64, rows as above):
   agg_linear_trend (src/pybasin/ts_torch/calculators/torch_features_trend.py:
   60-83; chunks of 5 / 10 / 50 samples, their max / min / mean / population
   variance fitted against the chunk index; fewer than two chunks give 0);
   friedrich_coefficients and max_langevin_fixed_point with m = 3
   (torch_features_dynamical.py:87-177: the cubic's normal equations + 1e-6 I
   solved with partial pivoting; the largest x_i where the fitted cubic
   changes sign to x_{i+1}, 0 if none); benford_correlation
   (torch_features_advanced.py:21-78; the digit floor(|x| / 10^floor(log10
   |x|)) in double, counted when 1..9); with want_cwt, cwt_coefficients
   (torch_features_frequency.py:71-122: Ricker widths 2, 5, 10, 20,
   coefficients 0..14, reflect padding, replicate for windows of at most 4 w
   rows, 0 past the window).  Non-finite windows give NaN. */
int32_t bh_window_trends(int32_t dtype, const void *traj, int64_t n_save, int64_t n, int32_t dim,
                         int64_t start, int32_t want_cwt, double *out, void *stream);

/* Peak counts of each series of the same window: out (12, n, dim) float64,
   rows 0..3 = the positions strictly greater than their s_r neighbours on
   each side (number_peaks with n = s_r, src/pybasin/ts_torch/calculators/
   torch_features_pattern.py:64-105; s_r in [0, 64], 0 counts nothing), rows
   4..11 = the positions equal to the maximum of the 2 w + 1 samples about
   them for w = 1..8 (the per-width counts of number_cwt_peaks, :154-190).
   Windows up to 5,290 float32 or 3,968 float64 rows; non-finite windows give
   NaN. */
int32_t bh_window_peak_counts(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                              int32_t dim, int64_t start, int32_t s0, int32_t s1, int32_t s2,
                              int32_t s3, double *out, void *stream);

/* permutation_entropy (tau = 1) of each series of the same window for the
   embedding dimensions 3..7 (torch_features_entropy_complexity.py:11-53):
   out (5, n, dim) float64, the entropy of the counts of the reference's
   pattern ids (sum_k argsort_k (d-1-k)!, equal values in index order) over
   ln d!; 0 for windows shorter than d, NaN for non-finite windows.  Windows
   up to 4,052 float32 or 2,026 float64 rows. */
int32_t bh_window_perm_entropy(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                               int32_t dim, int64_t start, double *out, void *stream);

/* Delay-embedding estimates of each series of the same window, in double
   (out (n, dim) float64 unless stated; NaN for non-finite windows):
   BH_EMB_CORR_DIM: correlation_dimension with emb_dim 4, lag 1, 50 radii
   (src/pybasin/ts_torch/torch_corr_dim.py:44-124: the log-log slope of the
   correlation sum over radii log-spaced in [0.1 sd, 0.5 sd]);
   BH_EMB_LYAP_R: lyapunov_r with emb_dim 10, lag 1, trajectory_len 20
   (torch_lyapunov_r.py:20-150: Rosenstein's divergence slope, nearest
   neighbours at a temporal separation above the mean period; divide by tau);
   BH_EMB_LYAP_E: lyapunov_e with emb_dim 10, matrix_dim 4, min_nb 8,
   min_tsep 0 (torch_lyapunov_e.py:15-150: Eckmann's four exponents, out
   (4, n, dim), inf for non-finite windows and for exponents nothing was
   accumulated for; divide by tau).
   Windows that fit the 62 KB LDS budget. */
enum { BH_EMB_CORR_DIM = 0, BH_EMB_LYAP_R = 1, BH_EMB_LYAP_E = 2 };
int32_t bh_window_embedding(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                            int32_t dim, int64_t start, int32_t which, double *out, void *stream);

/* lempel_ziv_complexity of each series of the same window for n_bins (1..8)
   bins settings bins[k] (host array, 1..255 each;
   src/pybasin/ts_torch/calculators/torch_features_entropy_complexity.py:
   140-180): the window cut into bins[k] symbols between its min and max (in
   the series dtype), parsed into phrases (each the shortest string at its
   position not seen before; a tail of seen prefixes is not counted); out
   (n_bins, n, dim) float64 = phrase count / (w / (ln w + 1e-10)), NaN for
   non-finite windows.  Windows of at most 1,024 rows (five settings fit the
   LDS budget at any such window). */
int32_t bh_window_lempel_ziv(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                             int32_t dim, int64_t start, const int32_t *bins, int32_t n_bins,
                             double *out, void *stream);

/* Autocovariances of each series of the same window: out (max_lag + 1, n,
   dim) float64 holds C_k = sum_{i < w - k} (x_i - mean)(x_{i+k} - mean) for
   k = 0 .. max_lag (0 for k >= w), in double: the linear autocorrelation of
   the centred window that agg_autocorrelation and ar_coefficient form by FFT
   (src/pybasin/ts_torch/calculators/torch_features_autocorrelation.py:72-96,
   torch_features_trend.py:86-131).  0 <= max_lag <= 4095; windows that fit
   the 62 KB LDS budget. */
int32_t bh_window_autocov(int32_t dtype, const void *traj, int64_t n_save, int64_t n, int32_t dim,
                          int64_t start, int32_t max_lag, double *out, void *stream);

/* Statistics of the sorted values of each series of the same window (sorted
   in LDS): out_dup (5, n, dim) float64 = the adjacent equal pairs D of the
   sorted series, the values that lie in groups of equal values R, the number
   of such groups, the sum of the values in groups and the sum of one value per
   group, in double (has_duplicate, ratio_value_number_to_time_series_length
   and the four reoccurrence calculators: src/pybasin/ts_torch/calculators/
   torch_features_boolean.py:11-19, torch_features_reocurrance.py:11-161).
   out_levels (6, n, dim) in the series dtype, or NULL: torch.quantile of the
   window at 0, 0.2, 0.4, 0.6, 0.8, 1 (linear interpolation in the series
   dtype), the corridor bounds of bh_window_change_quantiles.  out_median
   (n, dim) in the series dtype, or NULL: the lower middle value of the
   sorted window (x.median(dim=0), torch_features_statistical.py:30-33).
   Non-finite windows give NaN.  Windows up to 8,192 float32 or 4,096 float64
   rows (past 2,048 / 1,024 rows one block sorts each series). */
int32_t bh_window_sorted(int32_t dtype, const void *traj, int64_t n_save, int64_t n, int32_t dim,
                         int64_t start, double *out_dup, void *out_levels, void *out_median,
                         void *stream);

/* change_quantiles (torch_features_change.py:20-139) of each series of the
   same window for the 15 corridors ql < qh of the levels 0, 0.2, .., 1 in the
   order (0, .2), (0, .4), .., (0, 1), (.2, .4), .., (.8, 1), given their
   bounds `levels` (6, n, dim) in the series dtype (bh_window_sorted's
   out_levels): out (15, 4, n, dim) float64 = mean and population variance of
   the changes x_{i+1} - x_i whose two samples lie in [level(ql), level(qh)],
   then of their absolute values (0 where none does; NaN where a level is
   NaN). */
int32_t bh_window_change_quantiles(int32_t dtype, const void *traj, int64_t n_save, int64_t n,
                                   int32_t dim, int64_t start, const void *levels, double *out,
                                   void *stream);

/* ---- label layer (replaces the sklearn calls of the reference's
   HDBSCANClusterer, src/pybasin/predictors/hdbscan_clusterer.py:12-148, and
   the 1-NN template classifiers of the case studies) --------------------- */

/* k nearest reference rows of every query row (Euclidean, double; rows of F
   <= 32 features, row-major): out_d (nq, k) distances ascending, out_i (nq, k)
   reference indices; equal distances keep the lower index.  1 <= k <= 32.
   sklearn NearestNeighbors.kneighbors / KNeighborsClassifier (brute force). */
int32_t bh_knn(const double *Q, int64_t nq, const double *R, int64_t nr, int32_t F, int32_t k,
               double *out_d, int64_t *out_i, void *stream);

/* device workspace bh_hdbscan_mst and bh_core_distances need for n rows */
int64_t bh_hdbscan_workspace_bytes(int64_t n);

/* HDBSCAN core distances: out_core[i] = distance from row i of X (n, F) to
   its k-th nearest row, the row itself included (the value equals column
   k - 1 of bh_knn(X, X)).  Rows should be in spatial order (e.g. along a
   Z-order curve): tiles of 64 consecutive rows are skipped by their bounding
   boxes, so the cost falls far below n^2 when consecutive rows are close.
   1 <= k <= min(32, n).  Replaces the core-distance query of sklearn
   HDBSCAN.fit (sklearn/cluster/_hdbscan/hdbscan.py, _hdbscan_brute /
   _hdbscan_prims: NearestNeighbors(min_samples).kneighbors), reached from
   HDBSCANClusterer.fit_predict (src/pybasin/predictors/hdbscan_clusterer.py:50-69). */
int32_t bh_core_distances(const double *X, int64_t n, int32_t F, int32_t k, double *out_core,
                          void *workspace, int64_t workspace_bytes, void *stream);

/* Minimum spanning tree of the HDBSCAN mutual-reachability graph
   max(core_a, core_b, d(a, b)) of X (n, F) given the core distances (the
   distance to the min_samples-th nearest row counting the row itself, e.g.
   column min_samples - 1 of bh_knn(X, X)): n - 1 edges (out_a, out_b,
   out_w) in device memory.  Boruvka rounds; SYNCHRONOUS (the host merges
   components between rounds).  Replaces the spanning tree inside sklearn
   HDBSCAN.fit (mst_from_data_matrix, sklearn/cluster/_hdbscan/_linkage.pyx). */
int32_t bh_hdbscan_mst(const double *X, int64_t n, int32_t F, const double *core, int64_t *out_a,
                       int64_t *out_b, double *out_w, void *workspace, int64_t workspace_bytes,
                       void *stream);

/* HOST arrays: flat HDBSCAN labels from the spanning tree -- single-linkage
   hierarchy (built once), condensed with each of the n_sizes
   min_cluster_sizes, excess-of-mass selection without the root
   (allow_single_cluster = False, cluster_selection_epsilon = 0): labels
   (n_sizes, n) int32, -1 noise, clusters 0.. in increasing condensed-tree
   id; n_clusters (n_sizes) the numbers selected (or NULL). */
int32_t bh_hdbscan_labels(const int64_t *edge_a, const int64_t *edge_b, const double *edge_w,
                          int64_t n, const int32_t *min_cluster_sizes, int32_t n_sizes,
                          int32_t *labels, int32_t *n_clusters);

/* Silhouette value of every row (rows sorted by label 0..n_labels-1, label
   c occupying [starts[c], starts[c+1])): s = (b - a) / max(a, b), 0 for a
   label of one row.  sklearn.metrics.silhouette_score = mean of s_out
   (replaces the silhouette_score call of _find_optimal_min_cluster_size,
   hdbscan_clusterer.py:112-148).  Optional weights (n) with label_counts
   (n_labels) = their sums per label: row i then stands for weights[i] equal
   rows (collapse duplicates to cut the pair count); the score is the
   weights-weighted mean of s_out.  Both NULL: every row counts once.
   sums: (n_labels, n) float64 device scratch for the per-label distance sums. */
int32_t bh_silhouette(const double *X_sorted, int64_t n, int32_t F, const int32_t *labels_sorted,
                      const int64_t *starts, int32_t n_labels, const double *weights,
                      const double *label_counts, double *sums, double *s_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* BASINHIP_H */
