/*
 * dopri5_oracle.c -- CPU restatement of the reference's Dormand-Prince 5(4)
 * integration path.  TEST INFRASTRUCTURE ONLY: used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product (pybasin_amd) never links, loads or calls this library.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, plain C11 + pthreads).
 *
 * Entry points (all arrays are host memory, doubles, row-major):
 *   or_solve(...)            batch solve, mode = per-lane torchdiffeq,
 *                            shared-dt torchdiffeq, or zigode semantics
 *   or_set_rossler_edges()   graph for the Rossler-network RHS
 *   or_rhs(...)              one RHS evaluation (for RHS unit tests)
 *
 * References (paths relative to the reference repository):
 *   src/zigode/src/solver/dopri5_generic.zig:16-321   tableau, controller,
 *                                                    DPS_C_MID, Hairer h0
 *   src/zigode/src/c_api.zig:193-314                 batch over ICs, threads
 *   src/pybasin/solvers/torchdiffeq_solver.py:97-134 anchor + odeint call
 *   src/pybasin/solvers/base.py:113-183              save grid, IC-major P
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum {
    OR_SYS_DECAY = 0,
    OR_SYS_LINEAR2 = 1,
    OR_SYS_PENDULUM = 2,
    OR_SYS_DUFFING = 3,
    OR_SYS_LORENZ = 4,
    OR_SYS_ROSSLER = 5,
};
enum {
    OR_OK = 0,
    OR_ERR_ALLOC = -2,
    OR_ERR_SOLVE = -3,
    OR_ERR_THREAD = -4,
    OR_ERR_ARG = -5,
    OR_ERR_MAX_STEPS = 1,
    OR_ERR_DT_UNDERFLOW = 2,
    OR_ERR_NONFINITE = 3,
};
enum { OR_MODE_TD_LANE = 0, OR_MODE_TD_SHARED = 1, OR_MODE_ZIG = 2 };

#define OR_ROSSLER_N 100
#define OR_ROSSLER_E 400
static int g_edges_i[2 * OR_ROSSLER_E], g_edges_j[2 * OR_ROSSLER_E];

/* rossler_network_topology.py:26-48: both directions of every edge, in
   list order (i->j, then j->i). */
int or_set_rossler_edges(const int32_t *ei, const int32_t *ej, int n_edges)
{
    if (n_edges != OR_ROSSLER_E) return OR_ERR_ARG;
    for (int e = 0; e < n_edges; e++) {
        g_edges_i[2 * e] = ei[e];
        g_edges_j[2 * e] = ej[e];
        g_edges_i[2 * e + 1] = ej[e];
        g_edges_j[2 * e + 1] = ei[e];
    }
    return OR_OK;
}

#define REAL float
#define SFX _f32
#define SIN sinf
#define COS cosf
#define FABS fabsf
#define POW powf
#define NEXTAFTER nextafterf
#include "dopri5_body.h"
#undef REAL
#undef SFX
#undef SIN
#undef COS
#undef FABS
#undef POW
#undef NEXTAFTER

#define REAL double
#define SFX _f64
#define SIN sin
#define COS cos
#define FABS fabs
#define POW pow
#define NEXTAFTER nextafter
#include "dopri5_body.h"

static int sys_dim_ok(int sys, int S)
{
    switch (sys) {
    case OR_SYS_DECAY: return S >= 1;
    case OR_SYS_LINEAR2: case OR_SYS_PENDULUM: case OR_SYS_DUFFING: return S == 2;
    case OR_SYS_LORENZ: return S == 3;
    case OR_SYS_ROSSLER: return S == 3 * OR_ROSSLER_N;
    default: return 0;
    }
}

typedef struct {
    int sys, bits, mode, S, n_save;
    int64_t n, ps, max_steps, start, end;
    const double *y0, *p, *save_ts;
    double t0, t1, rtol, atol;
    double *out;
    int32_t *status;
    int64_t *nsteps;
    int n_p;
} job_t;

/* one row through the chosen driver; converts double inputs to REAL */
static void solve_row(const job_t *J, int64_t r)
{
    const int S = J->S;
    const int64_t n = J->n;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)J->n_save * S);
    int rc = OR_ERR_ALLOC;
    int64_t steps = 0, acc = 0;
    if (!tmp) goto out;
    if (J->bits == 32) {
        float y0[OR_ROSSLER_N * 3], p[16];
        for (int i = 0; i < S; i++) y0[i] = (float)J->y0[r * S + i];
        for (int i = 0; i < J->n_p; i++) p[i] = (float)J->p[r * J->ps + i];
        if (J->mode == OR_MODE_ZIG)
            rc = zig_solve_f32(J->sys, S, y0, p, J->t0, J->t1, J->save_ts, J->n_save, J->rtol,
                               J->atol, J->max_steps, tmp, S, &steps);
        else
            rc = td_solve_f32(J->sys, 1, S, y0, p, 0, J->t0, J->save_ts, J->n_save, J->rtol,
                              J->atol, J->max_steps, tmp, &steps, &acc);
    } else {
        double y0[OR_ROSSLER_N * 3], p[16];
        for (int i = 0; i < S; i++) y0[i] = J->y0[r * S + i];
        for (int i = 0; i < J->n_p; i++) p[i] = J->p[r * J->ps + i];
        if (J->mode == OR_MODE_ZIG)
            rc = zig_solve_f64(J->sys, S, y0, p, J->t0, J->t1, J->save_ts, J->n_save, J->rtol,
                               J->atol, J->max_steps, tmp, S, &steps);
        else
            rc = td_solve_f64(J->sys, 1, S, y0, p, 0, J->t0, J->save_ts, J->n_save, J->rtol,
                              J->atol, J->max_steps, tmp, &steps, &acc);
    }
    for (int s = 0; s < J->n_save; s++)
        for (int i = 0; i < S; i++)
            J->out[((int64_t)s * n + r) * S + i] = tmp[(int64_t)s * S + i];
out:
    free(tmp);
    if (J->status) J->status[r] = rc;
    if (J->nsteps) J->nsteps[r] = steps;
}

static void *worker(void *arg)
{
    job_t *J = (job_t *)arg;
    for (int64_t r = J->start; r < J->end; r++) solve_row(J, r);
    return NULL;
}

/*
 * Batch solve.  y0: (n, S); params: (n, n_p) with row stride ps (ps = 0 for
 * one shared parameter vector); save_ts: (n_save) doubles (the caller's save
 * grid, already rounded to the state dtype); t0: integration start (the
 * reference's t_span[0]); t1: end time for zigode mode (its dt clamp).
 * out: (n_save, n, S).  status/nsteps: (n) or NULL.
 * mode 1 (shared dt) integrates the whole batch as one tensor and ignores
 * n_threads; modes 0 and 2 split rows over n_threads pthreads.
 */
int or_solve(int sys, int bits, int mode, int64_t n, int S, const double *y0, const double *params,
             int n_p, int64_t ps, double t0, double t1, const double *save_ts, int n_save,
             double rtol, double atol, int64_t max_steps, double *out, int32_t *status,
             int64_t *nsteps, int n_threads)
{
    if (!sys_dim_ok(sys, S) || n <= 0 || n_save <= 0 || n_p > 16 || (bits != 32 && bits != 64))
        return OR_ERR_ARG;
    if (mode == OR_MODE_TD_SHARED) {
        int64_t m = n * S;
        int rc;
        int64_t steps = 0, acc = 0;
        if (bits == 32) {
            float *y = (float *)malloc(sizeof(float) * (size_t)m);
            float *p = (float *)malloc(sizeof(float) * (size_t)n * n_p + 4);
            if (!y || !p) { free(y); free(p); return OR_ERR_ALLOC; }
            for (int64_t i = 0; i < m; i++) y[i] = (float)y0[i];
            for (int64_t r = 0; r < n; r++)
                for (int i = 0; i < n_p; i++) p[r * n_p + i] = (float)params[r * ps + i];
            rc = td_solve_f32(sys, n, S, y, p, n_p, t0, save_ts, n_save, rtol, atol, max_steps,
                              out, &steps, &acc);
            free(y);
            free(p);
        } else {
            double *p = (double *)malloc(sizeof(double) * (size_t)n * n_p + 8);
            if (!p) return OR_ERR_ALLOC;
            for (int64_t r = 0; r < n; r++)
                for (int i = 0; i < n_p; i++) p[r * n_p + i] = params[r * ps + i];
            rc = td_solve_f64(sys, n, S, y0, p, n_p, t0, save_ts, n_save, rtol, atol, max_steps,
                              out, &steps, &acc);
            free(p);
        }
        for (int64_t r = 0; r < n; r++) {
            if (status) status[r] = rc;
            if (nsteps) nsteps[r] = steps;
        }
        return rc < 0 ? rc : OR_OK;
    }
    if (mode != OR_MODE_TD_LANE && mode != OR_MODE_ZIG) return OR_ERR_ARG;
    if (n_threads < 1) n_threads = 1;
    if (n_threads > n) n_threads = (int)n;
    job_t base = {sys, bits, mode, S, n_save, n, ps, max_steps, 0, n, y0, params, save_ts,
                  t0, t1, rtol, atol, out, status, nsteps, n_p};
    if (n_threads == 1) {
        worker(&base);
        return OR_OK;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
    job_t *jobs = (job_t *)malloc(sizeof(job_t) * n_threads);
    if (!th || !jobs) { free(th); free(jobs); return OR_ERR_ALLOC; }
    int64_t chunk = (n + n_threads - 1) / n_threads;
    int spawned = 0, rc = OR_OK;
    for (int w = 0; w < n_threads; w++) {
        jobs[w] = base;
        jobs[w].start = w * chunk;
        jobs[w].end = (w + 1) * chunk < n ? (w + 1) * chunk : n;
        if (jobs[w].start >= jobs[w].end) break;
        if (pthread_create(&th[w], NULL, worker, &jobs[w]) != 0) { rc = OR_ERR_THREAD; break; }
        spawned++;
    }
    for (int w = 0; w < spawned; w++) pthread_join(th[w], NULL);
    free(th);
    free(jobs);
    return rc;
}

/* one RHS evaluation (REAL = float when bits == 32) for RHS unit tests */
int or_rhs(int sys, int bits, int S, double t, const double *y, const double *p, int n_p,
           double *dy)
{
    if (!sys_dim_ok(sys, S) || n_p > 16) return OR_ERR_ARG;
    if (bits == 32) {
        float yy[OR_ROSSLER_N * 3] = {0}, dd[OR_ROSSLER_N * 3], pp[16] = {0};
        for (int i = 0; i < S; i++) yy[i] = (float)y[i];
        for (int i = 0; i < n_p; i++) pp[i] = (float)p[i];
        rhs_f32(sys, (float)t, yy, dd, pp, S);
        for (int i = 0; i < S; i++) dy[i] = dd[i];
    } else {
        double pp[16] = {0};
        for (int i = 0; i < n_p; i++) pp[i] = p[i];
        rhs_f64(sys, t, y, dy, pp, S);
    }
    return OR_OK;
}
