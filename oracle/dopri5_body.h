/*
 * Oracle body, included twice by dopri5_oracle.c with REAL = float / double
 * and SFX = _f32 / _f64.  TEST INFRASTRUCTURE ONLY: this file is the CPU
 * checker for the HIP path; nothing in the product links or calls it.
 *
 * Two Dormand-Prince 5(4) drivers are restated here:
 *
 *   td_solve   torchdiffeq 0.2.5 `dopri5` as called by the reference's
 *              TorchDiffEqSolver (src/pybasin/solvers/torchdiffeq_solver.py:97-134):
 *              one dt shared by the whole batch, error = RMS over all B*S
 *              elements, state arithmetic in REAL, time-like values in double,
 *              stage times cast to REAL with the "previous representable value"
 *              perturbation on the c=1 stages, no clamp at the last save time,
 *              power-form evaluation of the DPS_C_MID quartic.  Called with one
 *              row it is the per-lane variant the HIP kernel implements.
 *              torchdiffeq itself is an external dependency (uv.lock:3587-3588,
 *              version 0.2.5) that is not vendored in the reference; its
 *              algorithm is restated from the published package and from the
 *              reference's own restatement in
 *              src/zigode/src/solver/dopri5_generic.zig:16-321.
 *
 *   zig_solve  the reference's native per-IC solver,
 *              src/zigode/src/solver/dopri5_generic.zig:77-242 (solve_into) and
 *              :270-321 (initialStepSize): per-IC RMS error, dt clamped to t1,
 *              Horner evaluation of the same quartic, 1e-12 save tolerance,
 *              tail fill with the last state.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* one RHS evaluation for one row; t is already in REAL (torchdiffeq casts the
   time to the state dtype before calling the ODE, torch_ode_system.py:69-79). */
static void FN(rhs)(int sys, REAL t, const REAL *y, REAL *dy, const REAL *p, int S)
{
    switch (sys) {
    case OR_SYS_DECAY: /* tests/unit/solvers/test_solver.py:16-18: p[...,0:1] * y */
        for (int i = 0; i < S; i++) dy[i] = p[0] * y[i];
        break;
    case OR_SYS_LINEAR2: /* y' = M y, M = [[p0, p1], [p2, p3]] (test system) */
        dy[0] = p[0] * y[0] + p[1] * y[1];
        dy[1] = p[2] * y[0] + p[3] * y[1];
        break;
    case OR_SYS_PENDULUM: /* case_studies/pendulum/pendulum_ode.py:32-40 */
        dy[0] = y[1];
        dy[1] = -p[0] * y[1] + p[1] - p[2] * SIN(y[0]);
        break;
    case OR_SYS_DUFFING: /* case_studies/duffing_oscillator/duffing_ode.py:35-42 */
        dy[0] = y[1];
        dy[1] = -p[0] * y[1] - p[1] * (y[0] * y[0] * y[0]) + p[2] * COS(t);
        break;
    case OR_SYS_LORENZ: { /* case_studies/lorenz/lorenz_ode.py:40-64 (freeze mask) */
        REAL m = FABS(y[0]);
        if (FABS(y[1]) > m) m = FABS(y[1]);
        if (FABS(y[2]) > m) m = FABS(y[2]);
        REAL mask = (m < (REAL)200) ? (REAL)1 : (REAL)0;
        dy[0] = (p[0] * (y[1] - y[0])) * mask;
        dy[1] = (p[1] * y[0] - y[0] * y[2] - y[1]) * mask;
        dy[2] = (y[0] * y[1] - p[2] * y[2]) * mask;
        break;
    }
    case OR_SYS_ROSSLER: { /* rossler_network_jax_ode.py:63-79, edge scatter-add */
        const int N = OR_ROSSLER_N;
        const REAL *x = y, *yy = y + N, *z = y + 2 * N;
        REAL coup[OR_ROSSLER_N];
        for (int i = 0; i < N; i++) coup[i] = 0;
        for (int e = 0; e < 2 * OR_ROSSLER_E; e++) {
            int i = g_edges_i[e], j = g_edges_j[e];
            coup[i] += x[i] - x[j];
        }
        for (int i = 0; i < N; i++) {
            dy[i] = -yy[i] - z[i] - p[3] * coup[i];
            dy[N + i] = x[i] + p[0] * yy[i];
            dy[2 * N + i] = p[1] + z[i] * (x[i] - p[2]);
        }
        break;
    }
    default:
        for (int i = 0; i < S; i++) dy[i] = 0;
    }
}

static void FN(rhs_batch)(int sys, REAL t, const REAL *y, REAL *dy, const REAL *p,
                          int64_t p_stride, int64_t n, int S)
{
    for (int64_t r = 0; r < n; r++)
        FN(rhs)(sys, t, y + r * S, dy + r * S, p + r * p_stride, S);
}

/* RMS over m elements, accumulated in double (torch's `_rms_norm`,
   tensor.abs().pow(2).mean().sqrt(), returns the state dtype). */
static REAL FN(rms)(const REAL *v, int64_t m)
{
    double acc = 0.0;
    for (int64_t i = 0; i < m; i++) acc += (double)v[i] * (double)v[i];
    return (REAL)sqrt(acc / (double)m);
}

/* torchdiffeq `_select_initial_step` (Hairer, Norsett & Wanner II.4), done over
   the whole batch tensor; mirrored per IC in dopri5_generic.zig:270-321. */
static double FN(td_initial_step)(int sys, double t0, const REAL *y0, const REAL *f0,
                                  const REAL *p, int64_t ps, int64_t n, int S,
                                  double rtol, double atol, REAL *w1, REAL *w2, REAL *w3)
{
    const int64_t m = n * S;
    const REAL ratol = (REAL)atol, rrtol = (REAL)rtol;
    for (int64_t i = 0; i < m; i++) {
        REAL scale = ratol + FABS(y0[i]) * rrtol;
        w1[i] = y0[i] / scale;
        w2[i] = f0[i] / scale;
    }
    REAL d0 = FN(rms)(w1, m), d1 = FN(rms)(w2, m);
    REAL h0 = (d0 < (REAL)1e-5 || d1 < (REAL)1e-5) ? (REAL)1e-6 : (REAL)0.01 * d0 / d1;
    h0 = FABS(h0);
    for (int64_t i = 0; i < m; i++) w1[i] = y0[i] + h0 * f0[i];
    FN(rhs_batch)(sys, (REAL)(t0 + (double)h0), w1, w3, p, ps, n, S);
    for (int64_t i = 0; i < m; i++) {
        REAL scale = ratol + FABS(y0[i]) * rrtol;
        w2[i] = (w3[i] - f0[i]) / scale;
    }
    REAL d2 = FABS(FN(rms)(w2, m) / h0);
    REAL h1;
    if (d1 <= (REAL)1e-15 && d2 <= (REAL)1e-15) {
        h1 = h0 * (REAL)1e-3;
        if (h1 < (REAL)1e-6) h1 = (REAL)1e-6;
    } else {
        REAL dm = d1 > d2 ? d1 : d2;
        h1 = POW((REAL)0.01 / dm, (REAL)(1.0 / 5.0));
    }
    h1 = FABS(h1);
    REAL h = (REAL)100 * h0;
    if (h1 < h) h = h1;
    return (double)h;
}

/* torchdiffeq `_optimal_step_size` (safety 0.9, ifactor 10, dfactor 0.2,
   order 5); same rule as dopri5_generic.zig:228-235. */
#ifndef OR_TD_NEXT_DT
#define OR_TD_NEXT_DT
static double td_next_dt(double dt, double ratio)
{
    if (ratio == 0.0) return dt * 10.0;
    double dfac = ratio < 1.0 ? 1.0 : 0.2;
    double f = 0.9 / pow(ratio, 1.0 / 5.0);
    if (f < dfac) f = dfac;
    if (f > 10.0) f = 10.0;
    return dt * f;
}
#endif

/*
 * torchdiffeq driver over a batch of n rows (shared dt).  save_ts are the
 * REAL save times promoted to double (torchdiffeq: t.to(float64)); if
 * save_ts[0] > t0 the caller's anchor semantics apply (integration starts at
 * t0, torchdiffeq_solver.py:122-125).  out: (n_save, n, S) in double.
 * Returns 0 or an OR_ERR_* code; *nsteps counts attempted steps.
 */
static int FN(td_solve)(int sys, int64_t n, int S, const REAL *y0in, const REAL *p,
                        int64_t ps, double t0, const double *save_ts, int n_save,
                        double rtol, double atol, int64_t max_steps, double *out,
                        int64_t *nsteps_out, int64_t *naccept_out)
{
    const int64_t m = n * S;
    REAL *buf = (REAL *)malloc(sizeof(REAL) * (size_t)m * 16);
    if (!buf) return OR_ERR_ALLOC;
    REAL *y = buf, *k[7], *ys = buf + 8 * m, *y1 = buf + 9 * m, *yerr = buf + 10 * m;
    REAL *w1 = buf + 11 * m, *w2 = buf + 12 * m, *w3 = buf + 13 * m;
    REAL *ca = buf + 14 * m, *cb = buf + 15 * m; /* interp coefficients a, b */
    REAL *cc = w3;                               /* reused after the initial step */
    for (int j = 0; j < 7; j++) k[j] = buf + (1 + j) * m;
    /* interp state: [e, d, c, b, a] for the last accepted step */
    REAL *ce = (REAL *)malloc(sizeof(REAL) * (size_t)m * 2);
    if (!ce) { free(buf); return OR_ERR_ALLOC; }
    REAL *cd = ce + m;

    memcpy(y, y0in, sizeof(REAL) * (size_t)m);
    FN(rhs_batch)(sys, (REAL)t0, y, k[0], p, ps, n, S);
    double dt = FN(td_initial_step)(sys, t0, y, k[0], p, ps, n, S, rtol, atol, w1, w2, w3);
    /* before the first step the interpolant is the constant y0 */
    for (int64_t i = 0; i < m; i++) { ce[i] = y[i]; cd[i] = 0; cc[i] = 0; cb[i] = 0; ca[i] = 0; }
    double ti0 = t0, ti1 = t0; /* interval of the current interpolant */
    int64_t steps = 0, accepted = 0;
    int rc = OR_OK;

    static const double A[6][6] = {
        {1.0 / 5.0},
        {3.0 / 40.0, 9.0 / 40.0},
        {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
        {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
        {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
        {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
    static const double C[6] = {1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
    static const double E[7] = {35.0 / 384.0 - 1951.0 / 21600.0, 0.0,
                                500.0 / 1113.0 - 22642.0 / 50085.0, 125.0 / 192.0 - 451.0 / 720.0,
                                -2187.0 / 6784.0 - -12231.0 / 42400.0, 11.0 / 84.0 - 649.0 / 6300.0,
                                -1.0 / 60.0};
    static const double MID[7] = {6025192743.0 / 30085553152.0 / 2.0, 0.0,
                                  51252292925.0 / 65400821598.0 / 2.0,
                                  -2691868925.0 / 45128329728.0 / 2.0,
                                  187940372067.0 / 1594534317056.0 / 2.0,
                                  -1776094331.0 / 19743644256.0 / 2.0,
                                  11237099.0 / 235043384.0 / 2.0};

    for (int si = 0; si < n_save; si++) {
        const double ts = save_ts[si];
        while (ts > ti1) {
            if (steps >= max_steps) { rc = OR_ERR_MAX_STEPS; goto done; }
            const double tA = ti1, tB = ti1 + dt;
            if (!(tB > tA)) { rc = OR_ERR_DT_UNDERFLOW; goto done; }
            for (int64_t i = 0; i < m; i++)
                if (!isfinite((double)y[i])) { rc = OR_ERR_NONFINITE; goto done; }
            const REAL tf = (REAL)tA, dtf = (REAL)dt, t1f = (REAL)tB;
            const REAL t1prev = NEXTAFTER(t1f, t1f - (REAL)1);
            for (int st = 0; st < 6; st++) {
                REAL coef[6];
                for (int j = 0; j <= st; j++) coef[j] = (REAL)A[st][j] * dtf;
                for (int64_t i = 0; i < m; i++) {
                    REAL acc = 0;
                    for (int j = 0; j <= st; j++) acc += k[j][i] * coef[j];
                    ys[i] = y[i] + acc;
                }
                REAL tst = (C[st] == 1.0) ? t1prev : tf + (REAL)C[st] * dtf;
                FN(rhs_batch)(sys, tst, ys, k[st + 1], p, ps, n, S);
            }
            /* y1 is the last stage input (c_sol == beta[-1] for Dormand-Prince) */
            memcpy(y1, ys, sizeof(REAL) * (size_t)m);
            REAL ecoef[7];
            for (int j = 0; j < 7; j++) ecoef[j] = dtf * (REAL)E[j];
            const REAL ratol = (REAL)atol, rrtol = (REAL)rtol;
            for (int64_t i = 0; i < m; i++) {
                REAL e = 0;
                for (int j = 0; j < 7; j++) e += k[j][i] * ecoef[j];
                REAL a0 = FABS(y[i]), a1 = FABS(y1[i]);
                REAL tol = ratol + rrtol * (a0 > a1 ? a0 : a1);
                yerr[i] = e / tol;
            }
            const double ratio = (double)FN(rms)(yerr, m);
            steps++;
            if (ratio <= 1.0) {
                /* torchdiffeq `_interp_fit` with DPS_C_MID */
                REAL mcoef[7];
                for (int j = 0; j < 7; j++) mcoef[j] = dtf * (REAL)MID[j];
                for (int64_t i = 0; i < m; i++) {
                    REAL acc = 0;
                    for (int j = 0; j < 7; j++) acc += k[j][i] * mcoef[j];
                    REAL ymid = y[i] + acc;
                    REAL f0 = k[0][i], f1 = k[6][i], a0 = y[i], a1 = y1[i];
                    ca[i] = (REAL)2 * dtf * (f1 - f0) - (REAL)8 * (a1 + a0) + (REAL)16 * ymid;
                    cb[i] = dtf * ((REAL)5 * f0 - (REAL)3 * f1) + (REAL)18 * a0 + (REAL)14 * a1 - (REAL)32 * ymid;
                    cc[i] = dtf * (f1 - (REAL)4 * f0) - (REAL)11 * a0 - (REAL)5 * a1 + (REAL)16 * ymid;
                    cd[i] = dtf * f0;
                    ce[i] = a0;
                }
                memcpy(y, y1, sizeof(REAL) * (size_t)m);
                memcpy(k[0], k[6], sizeof(REAL) * (size_t)m);
                ti0 = tA;
                ti1 = tB;
                accepted++;
            }
            dt = td_next_dt(dt, ratio);
        }
        /* torchdiffeq `_interp_evaluate`: x in double, cast to REAL, power form */
        double xd = (ti1 > ti0) ? (ts - ti0) / (ti1 - ti0) : 0.0;
        REAL x = (REAL)xd, x2 = x * x, x3 = x2 * x, x4 = x3 * x;
        double *o = out + (int64_t)si * m;
        for (int64_t i = 0; i < m; i++) {
            REAL v = ce[i] + x * cd[i];
            v = v + x2 * cc[i];
            v = v + x3 * cb[i];
            v = v + x4 * ca[i];
            o[i] = (double)v;
        }
    }
done:
    if (nsteps_out) *nsteps_out = steps;
    if (naccept_out) *naccept_out = accepted;
    free(ce);
    free(buf);
    return rc;
}

/* Zig reference driver, one IC (dopri5_generic.zig:77-242). out: (n_save, S) */
static int FN(zig_solve)(int sys, int S, const REAL *y0, const REAL *p, double t0, double t1,
                         const double *save_at, int n_save, double rtol, double atol,
                         int64_t max_steps, double *out, int64_t out_stride, int64_t *nsteps_out)
{
    REAL *buf = (REAL *)malloc(sizeof(REAL) * (size_t)S * 16);
    if (!buf) return OR_ERR_ALLOC;
    REAL *y = buf, *k[7], *ys = buf + 8 * S, *ynew = buf + 9 * S, *yerr = buf + 10 * S;
    REAL *ia = buf + 11 * S, *ib = buf + 12 * S, *ic = buf + 13 * S, *tmp = buf + 14 * S,
         *ftmp = buf + 15 * S;
    for (int j = 0; j < 7; j++) k[j] = buf + (1 + j) * S;
    int k1 = 0, k7 = 6;
    memcpy(y, y0, sizeof(REAL) * (size_t)S);
    double t = t0;
    int si = 0;
    FN(rhs)(sys, (REAL)t, y, k[k1], p, S);
    /* initialStepSize (dopri5_generic.zig:270-321), in double like the Zig code */
    double d0 = 0, d1 = 0;
    for (int i = 0; i < S; i++) {
        double sc = atol + fabs((double)y0[i]) * rtol;
        double r0 = (double)y0[i] / sc, r1 = (double)k[k1][i] / sc;
        d0 += r0 * r0;
        d1 += r1 * r1;
    }
    d0 = sqrt(d0 / S);
    d1 = sqrt(d1 / S);
    double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    for (int i = 0; i < S; i++) tmp[i] = (REAL)((double)y0[i] + h0 * (double)k[k1][i]);
    FN(rhs)(sys, (REAL)(t0 + h0), tmp, ftmp, p, S);
    double d2 = 0;
    for (int i = 0; i < S; i++) {
        double sc = atol + fabs((double)y0[i]) * rtol;
        double r = ((double)ftmp[i] - (double)k[k1][i]) / sc;
        d2 += r * r;
    }
    d2 = sqrt(d2 / S) / h0;
    double hh1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3)
                                               : pow(0.01 / fmax(d1, d2), 0.2);
    double dt = fmin(100.0 * h0, hh1);

    const double a21 = 1.0 / 5.0, a31 = 3.0 / 40.0, a32 = 9.0 / 40.0, a41 = 44.0 / 45.0,
                 a42 = -56.0 / 15.0, a43 = 32.0 / 9.0, a51 = 19372.0 / 6561.0,
                 a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0, a54 = -212.0 / 729.0,
                 a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                 a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0, a71 = 35.0 / 384.0,
                 a73 = 500.0 / 1113.0, a74 = 125.0 / 192.0, a75 = -2187.0 / 6784.0,
                 a76 = 11.0 / 84.0;
    const double c2 = 0.2, c3 = 0.3, c4 = 0.8, c5 = 8.0 / 9.0;
    const double e1 = 35.0 / 384.0 - 1951.0 / 21600.0, e3 = 500.0 / 1113.0 - 22642.0 / 50085.0,
                 e4 = 125.0 / 192.0 - 451.0 / 720.0, e5 = -2187.0 / 6784.0 - -12231.0 / 42400.0,
                 e6 = 11.0 / 84.0 - 649.0 / 6300.0, e7 = -1.0 / 60.0;
    const double m1 = 6025192743.0 / 30085553152.0 / 2.0, m3 = 51252292925.0 / 65400821598.0 / 2.0,
                 m4 = -2691868925.0 / 45128329728.0 / 2.0,
                 m5 = 187940372067.0 / 1594534317056.0 / 2.0,
                 m6 = -1776094331.0 / 19743644256.0 / 2.0, m7 = 11237099.0 / 235043384.0 / 2.0;
    int64_t step_count = 0;
    while (t < t1 && step_count < max_steps) {
        step_count++;
        if (t + dt > t1) dt = t1 - t;
        if (dt <= 0) break;
        REAL *K1 = k[k1], *K2 = k[1], *K3 = k[2], *K4 = k[3], *K5 = k[4], *K6 = k[5], *K7 = k[k7];
        const REAL h = (REAL)dt;
        for (int i = 0; i < S; i++) ys[i] = y[i] + (h * (REAL)a21) * K1[i];
        FN(rhs)(sys, (REAL)(t + c2 * dt), ys, K2, p, S);
        for (int i = 0; i < S; i++) ys[i] = y[i] + h * ((REAL)a31 * K1[i] + (REAL)a32 * K2[i]);
        FN(rhs)(sys, (REAL)(t + c3 * dt), ys, K3, p, S);
        for (int i = 0; i < S; i++)
            ys[i] = y[i] + h * ((REAL)a41 * K1[i] + (REAL)a42 * K2[i] + (REAL)a43 * K3[i]);
        FN(rhs)(sys, (REAL)(t + c4 * dt), ys, K4, p, S);
        for (int i = 0; i < S; i++)
            ys[i] = y[i] + h * ((REAL)a51 * K1[i] + (REAL)a52 * K2[i] + (REAL)a53 * K3[i] +
                                (REAL)a54 * K4[i]);
        FN(rhs)(sys, (REAL)(t + c5 * dt), ys, K5, p, S);
        for (int i = 0; i < S; i++)
            ys[i] = y[i] + h * ((REAL)a61 * K1[i] + (REAL)a62 * K2[i] + (REAL)a63 * K3[i] +
                                (REAL)a64 * K4[i] + (REAL)a65 * K5[i]);
        FN(rhs)(sys, (REAL)(t + dt), ys, K6, p, S);
        for (int i = 0; i < S; i++)
            ynew[i] = y[i] + h * ((REAL)a71 * K1[i] + (REAL)a73 * K3[i] + (REAL)a74 * K4[i] +
                                  (REAL)a75 * K5[i] + (REAL)a76 * K6[i]);
        FN(rhs)(sys, (REAL)(t + dt), ynew, K7, p, S);
        for (int i = 0; i < S; i++)
            yerr[i] = h * ((REAL)e1 * K1[i] + (REAL)e3 * K3[i] + (REAL)e4 * K4[i] +
                           (REAL)e5 * K5[i] + (REAL)e6 * K6[i] + (REAL)e7 * K7[i]);
        double err = 0;
        for (int i = 0; i < S; i++) {
            double sc = atol + rtol * fmax(fabs((double)y[i]), fabs((double)ynew[i]));
            double r = (double)yerr[i] / sc;
            err += r * r;
        }
        err = sqrt(err / S);
        if (err <= 1.0) {
            double tnew = t + dt;
            if (si < n_save && save_at[si] <= tnew + 1e-12) {
                for (int i = 0; i < S; i++) {
                    REAL ymid = y[i] + h * ((REAL)m1 * K1[i] + (REAL)m3 * K3[i] + (REAL)m4 * K4[i] +
                                            (REAL)m5 * K5[i] + (REAL)m6 * K6[i] + (REAL)m7 * K7[i]);
                    ia[i] = (REAL)2 * h * (K7[i] - K1[i]) - (REAL)8 * (ynew[i] + y[i]) + (REAL)16 * ymid;
                    ib[i] = h * ((REAL)5 * K1[i] - (REAL)3 * K7[i]) + (REAL)18 * y[i] +
                            (REAL)14 * ynew[i] - (REAL)32 * ymid;
                    ic[i] = h * (K7[i] - (REAL)4 * K1[i]) - (REAL)11 * y[i] - (REAL)5 * ynew[i] +
                            (REAL)16 * ymid;
                }
                while (si < n_save && save_at[si] <= tnew + 1e-12) {
                    double *o = out + (int64_t)si * out_stride;
                    if (save_at[si] <= t + 1e-12) {
                        for (int i = 0; i < S; i++) o[i] = (double)y[i];
                    } else {
                        REAL x = (REAL)((save_at[si] - t) / dt);
                        for (int i = 0; i < S; i++)
                            o[i] = (double)(y[i] + x * (h * K1[i] + x * (ic[i] + x * (ib[i] + x * ia[i]))));
                    }
                    si++;
                }
            }
            memcpy(y, ynew, sizeof(REAL) * (size_t)S);
            t = tnew;
            int tmpi = k1;
            k1 = k7;
            k7 = tmpi;
        }
        double dfac = err <= 1.0 ? 1.0 : 0.2;
        double factor = err == 0.0 ? 10.0 : fmin(10.0, fmax(dfac, 0.9 * pow(err, -0.2)));
        dt *= factor;
    }
    for (; si < n_save; si++) {
        double *o = out + (int64_t)si * out_stride;
        for (int i = 0; i < S; i++) o[i] = (double)y[i];
    }
    if (nsteps_out) *nsteps_out = step_count;
    free(buf);
    return OR_OK;
}

#undef FN
#undef CAT
#undef CAT2
