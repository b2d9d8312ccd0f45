/*
 * luna_oracle.c -- CPU oracle for the propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against (tests/, __graft_entry__.smoke and bench.py's cpu_baseline
 * leg).  It is never linked into, loaded by, or called from the product path.
 *
 * It restates, in plain C99 with the reference's order of floating-point
 * operations (built with -ffp-contract=off, as Julia does not fuse a*b+c):
 *   - FFTW's unnormalised r2c / c2r / c2c conventions used by
 *     src/NonlinearRHS.jl:16-53 and src/Luna.jl:366-370 (own radix-2 FFT with
 *     long-double twiddles; c2r ignores imag of the DC and Nyquist bins as FFTW does);
 *   - Maths.cumtrapz!            src/Maths.jl:299-327
 *   - Maths.CSpline evaluation   src/Maths.jl:771-775, 840-848
 *   - makePPTaccel lookup        src/Ionisation.jl:247-271
 *   - Kerr_field / Kerr_env      src/Nonlinear.jl:8-10, 24-34, 47-49, 63-73
 *   - PlasmaCumtrapz (scalar)    src/Nonlinear.jl:121-133, 164-177
 *   - RamanPolarEnv              src/Nonlinear.jl:201-333 (h(t) given as input)
 *   - RamanPolarField (thg)      src/Nonlinear.jl:183-247, 284-333 (h(t) given as input)
 *   - TransModeAvg               src/NonlinearRHS.jl:362-399
 *   - TransRadial                src/NonlinearRHS.jl:466-474 (QDHT as a dense product)
 *   - RK45.solve / PreconStepper src/RK45.jl:27-89, 128-309, 343-406, dopri.jl:1-28
 *   - Luna.run stepfun filter    src/Luna.jl:358-391
 *   - MemoryOutput/GridCondition src/Output.jl:65-81, 490-502
 *   - gradient fill: Capillary.gradient src/Capillary.jl:306-340, densityspline
 *     src/PhysData.jl:719-723, neff(mode, εco, nwg) src/Capillary.jl:216-234,
 *     make_linop src/LinearOps.jl:337-358 with conj_clamp :254, make_prop!(linop!)
 *     src/RK45.jl:283-296, norm_mode_average src/NonlinearRHS.jl:375-387
 *   - Stats.default / collect_stats, mode-averaged set, src/Stats.jl:18-258,
 *     409-412, 421-468, 476-532; Fields.energyfuncs src/Fields.jl:510-534;
 *     Maths.moment/level_xings/linterpx src/Maths.jl:89-96, 162-252;
 *     NumericalIntegration.jl 0.3 SimpsonEven (not in the reference tree:
 *     the extended Simpson rule h/48·(17,59,43,49,48,…,48,49,43,59,17))
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/luna_hip.h"

typedef double _Complex cplx;

#define NLSCALE_EPS0 8.854187817620389e-12
static const double E_CHARGE = 1.6021766208e-19;
static const double M_ELEC = 9.10938356e-31;

/* ------------------------------------------------------------------------ */
/* FFT: iterative radix-2 complex transform, unnormalised.                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    cplx *w; /* w[k] = exp(-2πi k/n), k < n */
    int *rev;
} fft_plan;

#define MAX_PLANS 32
static fft_plan g_plans[MAX_PLANS];
static int g_nplans = 0;

static const fft_plan *get_plan(int n) {
    for (int i = 0; i < g_nplans; i++)
        if (g_plans[i].n == n) return &g_plans[i];
    if (g_nplans >= MAX_PLANS) return NULL;
    fft_plan *p = &g_plans[g_nplans];
    p->n = n;
    p->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    p->rev = (int *)malloc(sizeof(int) * (size_t)n);
    const long double tau = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n; k++) {
        long double a = tau * (long double)k / (long double)n;
        p->w[k] = (double)cosl(a) - I * (double)sinl(a);
    }
    int lg = 0;
    while ((1 << lg) < n) lg++;
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++)
            if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        p->rev[i] = r;
    }
    g_nplans++;
    return p;
}

static inline cplx cmul(cplx a, cplx b) {
    return (creal(a) * creal(b) - cimag(a) * cimag(b)) + I * (creal(a) * cimag(b) + cimag(a) * creal(b));
}

/* sign = -1: forward (FFTW_FORWARD), +1: backward.  Unnormalised. */
static void fft_c2c(int n, cplx *x, int sign) {
    const fft_plan *p = get_plan(n);
    for (int i = 0; i < n; i++) {
        int r = p->rev[i];
        if (r > i) {
            cplx t = x[i];
            x[i] = x[r];
            x[r] = t;
        }
    }
    for (int m = 1; m < n; m <<= 1) {
        int step = n / (2 * m);
        for (int s = 0; s < n; s += 2 * m) {
            for (int j = 0; j < m; j++) {
                cplx w = p->w[j * step];
                if (sign > 0) w = conj(w);
                cplx a = x[s + j];
                cplx b = cmul(w, x[s + j + m]);
                x[s + j] = a + b;
                x[s + j + m] = a - b;
            }
        }
    }
}

/* r2c, n even: out[0..n/2] */
static void rfft(int n, const double *x, cplx *out, cplx *work) {
    int M = n / 2;
    for (int j = 0; j < M; j++) work[j] = x[2 * j] + I * x[2 * j + 1];
    fft_c2c(M, work, -1);
    const fft_plan *p = get_plan(n);
    for (int k = 0; k <= M; k++) {
        cplx zk = work[k % M];
        cplx zc = conj(work[(M - k) % M]);
        cplx E = 0.5 * (zk + zc);
        cplx O = -0.5 * I * (zk - zc);
        cplx w = (k < n) ? p->w[k] : 1.0;
        out[k] = E + cmul(w, O);
    }
}

/* c2r unnormalised (FFTW_BACKWARD), ignores imag of X[0] and X[n/2]. */
static void irfft_unnorm(int n, const cplx *Xin, double *x, cplx *work) {
    int M = n / 2;
    const fft_plan *p = get_plan(n);
    for (int k = 0; k < M; k++) {
        cplx Xk = Xin[k];
        cplx Xm = Xin[M - k];
        if (k == 0) {
            Xk = creal(Xk);
            Xm = creal(Xm);
        }
        cplx A = Xk + conj(Xm);
        cplx B = cmul(Xk - conj(Xm), conj(p->w[k]));
        work[k] = A + I * B;
    }
    fft_c2c(M, work, +1);
    for (int j = 0; j < M; j++) {
        x[2 * j] = creal(work[j]);
        x[2 * j + 1] = cimag(work[j]);
    }
}

/* ------------------------------------------------------------------------ */
/* Maths                                                                     */
/* ------------------------------------------------------------------------ */
/* cumtrapz!(out, y, dx), src/Maths.jl:323-327 */
static void cumtrapz(int n, const double *y, double dx, double *out) {
    out[0] = 0.0;
    for (int i = 1; i < n; i++) out[i] = out[i - 1] + (0.5 * (y[i - 1] + y[i])) * dx;
}

static void cumtrapz_c(int n, const cplx *y, double dx, cplx *out) {
    out[0] = 0.0;
    for (int i = 1; i < n; i++) out[i] = out[i - 1] + (0.5 * (y[i - 1] + y[i])) * dx;
}

/* CSpline call with the uniform index function, src/Maths.jl:771-775, 840-848 */
static double cspline_eval(int n, const double *x, const double *y, const double *D, double x0) {
    double xmin = x[0], xmax = x[n - 1];
    int i; /* 1-based */
    if (x0 <= xmin)
        i = 2;
    else if (x0 >= xmax)
        i = n;
    else
        i = (int)ceil((x0 - xmin) / (xmax - xmin) * (double)(n - 1)) + 1;
    int ic = i - 1, im = i - 2; /* 0-based */
    if (x0 == x[ic]) return y[ic];
    if (x0 == x[im]) return y[im];
    double t = (x0 - x[im]) / (x[ic] - x[im]);
    return (y[im] + D[im] * t + (3 * (y[ic] - y[im]) - 2 * D[im] - D[ic]) * (t * t) +
            (2 * (y[im] - y[ic]) + D[im] + D[ic]) * (t * t * t));
}

/* makePPTaccel ionrate!, src/Ionisation.jl:256-269 */
static int ppt_rate(const lh_desc *d, int n, const double *E, double *out) {
    double emin = d->ppt_x[0], emax = d->ppt_x[d->ppt_n - 1];
    for (int i = 0; i < n; i++) {
        double aE = fabs(E[i]);
        if (aE < emin)
            out[i] = 0.0;
        else if (aE > emax)
            return LH_E_PPT_RANGE;
        else if (!(aE <= emax))
            return LH_E_NONFINITE; /* Julia: ceil(Int, NaN) throws InexactError */
        else
            out[i] = exp(cspline_eval(d->ppt_n, d->ppt_x, d->ppt_y, d->ppt_d, aE));
    }
    return LH_OK;
}

/* ------------------------------------------------------------------------ */
/* Workspace                                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n_to, n_r, n_omega;
    double *Et, *Pt, *rate, *frac, *phase, *J, *P; /* real buffers [n_to*n_r] */
    cplx *Etc, *Ptc, *Eo, *work, *E2, *hw, *tmp;    /* complex buffers */
    int hw_ready; /* hω depends only on h(t): the reference rebuilds it every call
                     (Nonlinear.jl:311-312); computing it once gives identical values */
} rhs_ws;

static void ws_init(rhs_ws *w, const lh_desc *d) {
    size_t nto = (size_t)d->n_to, nr = (size_t)d->n_r;
    w->n_to = d->n_to;
    w->n_r = d->n_r;
    w->n_omega = d->n_omega;
    w->Et = (double *)calloc(nto * nr, sizeof(double));
    w->Pt = (double *)calloc(nto * nr, sizeof(double));
    w->rate = (double *)calloc(nto, sizeof(double));
    w->frac = (double *)calloc(nto, sizeof(double));
    w->phase = (double *)calloc(nto, sizeof(double));
    w->J = (double *)calloc(nto, sizeof(double));
    w->P = (double *)calloc(nto, sizeof(double));
    w->Etc = (cplx *)calloc(nto * nr, sizeof(cplx));
    w->Ptc = (cplx *)calloc(nto * nr, sizeof(cplx));
    w->Eo = (cplx *)calloc((nto + 2) * nr, sizeof(cplx));
    w->work = (cplx *)calloc(2 * nto + 2, sizeof(cplx));
    w->E2 = (cplx *)calloc(2 * nto, sizeof(cplx));
    w->hw = (cplx *)calloc(2 * nto, sizeof(cplx));
    w->tmp = (cplx *)calloc(2 * nto, sizeof(cplx));
    w->hw_ready = 0;
}

static void ws_free(rhs_ws *w) {
    free(w->Et); free(w->Pt); free(w->rate); free(w->frac); free(w->phase);
    free(w->J); free(w->P); free(w->Etc); free(w->Ptc); free(w->Eo);
    free(w->work); free(w->E2); free(w->hw); free(w->tmp);
}

/* ------------------------------------------------------------------------ */
/* Responses on one real column (Et_to_Pt!, src/NonlinearRHS.jl:121-126)     */
/* ------------------------------------------------------------------------ */
static void raman_field(const lh_desc *d, rhs_ws *w, const double *E, double *Pt);

static int responses_real(const lh_desc *d, rhs_ws *w, const double *E, double *Pt) {
    int n = d->n_to;
    double rho = d->density;
    for (int i = 0; i < n; i++) Pt[i] = 0.0;
    if (d->kerr == LH_KERR_FIELD) {
        double fac = rho * NLSCALE_EPS0 * d->gamma3;
        for (int i = 0; i < n; i++) Pt[i] = Pt[i] + fac * (E[i] * E[i] * E[i]);
    }
    if (d->plasma) {
        /* PlasmaScalar!, src/Nonlinear.jl:121-133 */
        double dt = d->dt_o;
        double e_ratio = E_CHARGE * E_CHARGE / M_ELEC;
        int st = ppt_rate(d, n, E, w->rate);
        if (st) return st;
        cumtrapz(n, w->rate, dt, w->frac);
        for (int i = 0; i < n; i++) w->frac[i] = (d->preionfrac + 1) - exp(-w->frac[i]);
        for (int i = 0; i < n; i++) w->phase[i] = (w->frac[i] * e_ratio) * E[i];
        cumtrapz(n, w->phase, dt, w->J);
        for (int i = 0; i < n; i++)
            if (fabs(E[i]) > 0) w->J[i] += ((d->ionpot * w->rate[i]) * (1 - w->frac[i])) / E[i];
        cumtrapz(n, w->J, dt, w->P);
        for (int i = 0; i < n; i++) Pt[i] = Pt[i] + rho * w->P[i];
    }
    if (d->raman) raman_field(d, w, E, Pt);
    return LH_OK;
}

/* RamanPolarField call with thg = true, src/Nonlinear.jl:183-247, 284-333:
 * E.^2 on the doubled grid, FT, ×hω·δt, inverse, Pout = ρ·E·P (R.r(ht, ρ) =
 * raman_h).  The reference's rfft/irfft pair on real data is the c2c pair of
 * the same length restricted to real input. */
static void raman_field(const lh_desc *d, rhs_ws *w, const double *E, double *Pt) {
    int n = d->n_to, n2 = 2 * d->n_to;
    double rho = d->density;
    for (int i = 0; i < n2; i++) w->E2[i] = 0.0;
    for (int i = 0; i < n; i++) w->E2[i] = E[i] * E[i];
    if (!w->hw_ready) {
        for (int i = 0; i < n2; i++) w->hw[i] = (i < n) ? d->raman_h[i] : 0.0;
        fft_c2c(n2, w->hw, -1);
        w->hw_ready = 1;
    }
    fft_c2c(n2, w->E2, -1);
    for (int i = 0; i < n2; i++) w->tmp[i] = cmul(w->hw[i], w->E2[i]) * d->dt_o;
    fft_c2c(n2, w->tmp, +1);
    double s = 1.0 / (double)n2;
    for (int i = 0; i < n; i++) {
        double P = creal(w->tmp[i]) * s;
        Pt[i] = Pt[i] + (rho * E[i]) * P;
    }
}

/* RamanPolarEnv call, src/Nonlinear.jl:282-333, with R.r(ht, ρ) = raman_h. */
static void raman_env(const lh_desc *d, rhs_ws *w, const cplx *E, cplx *Pt) {
    int n = d->n_to, n2 = 2 * d->n_to;
    double rho = d->density;
    for (int i = 0; i < n2; i++) w->E2[i] = 0.0;
    for (int i = 0; i < n; i++) w->E2[i] = 0.5 * (creal(E[i]) * creal(E[i]) + cimag(E[i]) * cimag(E[i]));
    if (!w->hw_ready) {
        for (int i = 0; i < n2; i++) w->hw[i] = (i < n) ? d->raman_h[i] : 0.0;
        fft_c2c(n2, w->hw, -1);
        w->hw_ready = 1;
    }
    fft_c2c(n2, w->E2, -1);
    for (int i = 0; i < n2; i++) w->tmp[i] = cmul(w->hw[i], w->E2[i]) * d->dt_o;
    fft_c2c(n2, w->tmp, +1);
    double s = 1.0 / (double)n2;
    for (int i = 0; i < n; i++) {
        cplx P = w->tmp[i] * s;
        cplx pout = cmul(rho * E[i], P);
        Pt[i] = Pt[i] + pout;
    }
}

static void responses_env(const lh_desc *d, rhs_ws *w, const cplx *E, cplx *Pt) {
    int n = d->n_to;
    for (int i = 0; i < n; i++) Pt[i] = 0.0;
    if (d->kerr == LH_KERR_ENV) {
        double fac = d->density * NLSCALE_EPS0 * d->gamma3;
        for (int i = 0; i < n; i++) {
            double a2 = creal(E[i]) * creal(E[i]) + cimag(E[i]) * cimag(E[i]);
            double s = ((0.75 * fac) * a2);
            Pt[i] = Pt[i] + (s * creal(E[i]) + I * (s * cimag(E[i])));
        }
    }
    if (d->raman) raman_env(d, w, E, Pt);
}

/* ------------------------------------------------------------------------ */
/* Transforms                                                                */
/* ------------------------------------------------------------------------ */
/* X[:, i] = scale * sum_j T[i, j] X[:, j] for complex columns of nto samples
 * (Hankel.jl's ldiv!/mul! on a ComplexF64 array: the real matrix applied to
 * real and imaginary parts), through tmp; sum in increasing j as for the
 * field grid. */
static void qdht_cols_c(const lh_desc *d, cplx *X, cplx *tmp, double scale) {
    int nto = d->n_to, nr = d->n_r;
    for (int i = 0; i < nr; i++)
        for (int t = 0; t < nto; t++) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < nr; j++) {
                double m = d->qdht[(size_t)i * nr + j];
                cplx x = X[(size_t)j * nto + t];
                sr += m * creal(x);
                si += m * cimag(x);
            }
            tmp[(size_t)i * nto + t] = (sr * scale) + I * (si * scale);
        }
    memcpy(X, tmp, sizeof(cplx) * (size_t)nto * nr);
}

static int rhs_eval(const lh_desc *d, rhs_ws *w, cplx *nl, const cplx *Ew) {
    int N = d->n_omega, nto = d->n_to, nr = d->n_r;
    int st = LH_OK;
    if (d->grid_kind == LH_GRID_REAL) {
        int No = nto / 2 + 1;
        double sc_t = (double)(No - 1) / (double)(N - 1);
        double sc_f = (double)(N - 1) / (double)(No - 1);
        double inv_n = 1.0 / (double)nto;
        /* to_time! per column, src/NonlinearRHS.jl:16-23 */
        for (int c = 0; c < nr; c++) {
            cplx *Eo = w->Eo + (size_t)c * No;
            for (int i = 0; i < No; i++) Eo[i] = 0.0;
            for (int i = 0; i < N; i++) Eo[i] = sc_t * Ew[(size_t)c * N + i];
            double *Et = w->Et + (size_t)c * nto;
            irfft_unnorm(nto, Eo, Et, w->work);
            for (int i = 0; i < nto; i++) Et[i] = inv_n * Et[i];
        }
        if (d->transform == LH_TRANS_MODEAVG) {
            /* Eto /= nlscale*sqrt(aeff(z)), src/NonlinearRHS.jl:364 */
            for (int i = 0; i < nto; i++) w->Et[i] = w->Et[i] / d->field_div;
        } else {
            /* ldiv!(Eto, QDHT, Eto): k→r, Eto[:, i] = (K/R) Σ_j T[i,j] Eto[:, j] */
            double *tmp = w->Pt;
            for (int i = 0; i < nr; i++)
                for (int t = 0; t < nto; t++) {
                    double s = 0.0;
                    for (int j = 0; j < nr; j++) s += d->qdht[(size_t)i * nr + j] * w->Et[(size_t)j * nto + t];
                    tmp[(size_t)i * nto + t] = s * d->qdht_inv;
                }
            memcpy(w->Et, tmp, sizeof(double) * (size_t)nto * nr);
        }
        for (int c = 0; c < nr; c++) {
            st = responses_real(d, w, w->Et + (size_t)c * nto, w->Pt + (size_t)c * nto);
            if (st) return st;
        }
        for (int c = 0; c < nr; c++)
            for (int i = 0; i < nto; i++) w->Pt[(size_t)c * nto + i] *= d->to_win[i];
        if (d->transform == LH_TRANS_RADIAL) {
            /* mul!(Pto, QDHT, Pto): r→k */
            double *tmp = w->Et;
            for (int i = 0; i < nr; i++)
                for (int t = 0; t < nto; t++) {
                    double s = 0.0;
                    for (int j = 0; j < nr; j++) s += d->qdht[(size_t)i * nr + j] * w->Pt[(size_t)j * nto + t];
                    tmp[(size_t)i * nto + t] = s * d->qdht_fwd;
                }
            memcpy(w->Pt, tmp, sizeof(double) * (size_t)nto * nr);
        }
        /* to_freq!, src/NonlinearRHS.jl:39-45 */
        for (int c = 0; c < nr; c++) {
            rfft(nto, w->Pt + (size_t)c * nto, w->Eo, w->work);
            for (int i = 0; i < N; i++) nl[(size_t)c * N + i] = sc_f * w->Eo[i];
        }
    } else {
        /* envelope: complex to_time!/to_freq!, src/NonlinearRHS.jl:25-32, 47-53 */
        int No = nto, h = N / 2;
        double sc_t = (double)No / (double)N;
        double sc_f = (double)N / (double)No;
        double inv_n = 1.0 / (double)nto;
        for (int c = 0; c < nr; c++) {
            cplx *Et = w->Etc + (size_t)c * nto;
            for (int i = 0; i < No; i++) Et[i] = 0.0;
            const cplx *src = Ew + (size_t)c * N;
            for (int i = 0; i < h; i++) Et[i] = sc_t * src[i];
            for (int i = 0; i < h; i++) Et[No - 1 - i] = sc_t * src[N - 1 - i];
            fft_c2c(nto, Et, +1);
            for (int i = 0; i < nto; i++) Et[i] = inv_n * Et[i];
            if (d->transform == LH_TRANS_MODEAVG)
                for (int i = 0; i < nto; i++) Et[i] = Et[i] / d->field_div;
        }
        /* TransRadial on the envelope grid (src/NonlinearRHS.jl:457-459, 466-474):
         * the same real QDHT matrix acts on the real and imaginary parts */
        if (d->transform == LH_TRANS_RADIAL) qdht_cols_c(d, w->Etc, w->Ptc, d->qdht_inv);
        for (int c = 0; c < nr; c++) {
            responses_env(d, w, w->Etc + (size_t)c * nto, w->Ptc + (size_t)c * nto);
            cplx *Pt = w->Ptc + (size_t)c * nto;
            for (int i = 0; i < nto; i++) Pt[i] = Pt[i] * d->to_win[i];
        }
        if (d->transform == LH_TRANS_RADIAL) qdht_cols_c(d, w->Ptc, w->Etc, d->qdht_fwd);
        for (int c = 0; c < nr; c++) {
            cplx *Pt = w->Ptc + (size_t)c * nto;
            fft_c2c(nto, Pt, -1);
            cplx *dst = nl + (size_t)c * N;
            for (int i = 0; i < h; i++) dst[i] = sc_f * Pt[i];
            for (int i = 0; i < h; i++) dst[N - 1 - i] = sc_f * Pt[No - 1 - i];
        }
    }
    /* normalisation */
    const cplx *nf = (const cplx *)d->norm;
    if (d->transform == LH_TRANS_MODEAVG) {
        for (int i = 0; i < N; i++)
            if (d->sidx[i]) nl[i] = cmul(nl[i], nf[i]);
        for (int i = 0; i < N; i++)
            if (d->sidx[i]) nl[i] = nl[i] * d->omega_win[i];
    } else {
        for (size_t i = 0; i < (size_t)N * nr; i++) nl[i] = cmul(nl[i], nf[i]);
    }
    return st;
}

/* ------------------------------------------------------------------------ */
/* RK45 (PreconStepper), src/RK45.jl                                         */
/* ------------------------------------------------------------------------ */
static const double B_[6][6] = {
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0},
    {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}};
static const double nodes_[6] = {1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1, 1};
static const double b5_[7] = {5179.0 / 57600, 0, 7571.0 / 16695, 393.0 / 640, -92097.0 / 339200, 187.0 / 2100, 1.0 / 40};
static const double b4_[7] = {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84, 0};
static const double interpC_[4][7] = {
    {1, 0, 0, 0, 0, 0, 0},
    {-183.0 / 64, 0, 1500.0 / 371, -125.0 / 32, 9477.0 / 3392, -11.0 / 7, 3.0 / 2},
    {37.0 / 12, 0, -1000.0 / 159, 125.0 / 12, -729.0 / 106, 11.0 / 3, -4},
    {-145.0 / 128, 0, 1000.0 / 371, -375.0 / 64, 25515.0 / 6784, -55.0 / 28, 5.0 / 2}};

typedef int (*orc_rhs_fn)(void *user, cplx *out, const cplx *y, double t);

typedef void (*orc_linop_fn)(void *user, double t, cplx *L);

typedef struct {
    int n;
    const cplx *linop;
    orc_linop_fn linop_at; /* non-constant operator: linop!(out, t) (NULL: constant) */
    void *luser;
    cplx *Lz;
    double lastt2;
    orc_rhs_fn f;
    void *user;
    cplx *y, *yn, *yi, *yerr, *ybar, *ks[7];
    double t, tn, dt, dtn, rtol, atol, safety, max_dt, min_dt, err, errlast;
    int ok;
    int64_t rhs_evals;
} stepper;

/* make_prop!(linop::AbstractArray), src/RK45.jl:271-281: y .*= exp(linop*h);
 * make_prop!(linop!), :283-296: the operator is evaluated at t2 (the later
 * time, also when propagating backwards) whenever t2 changes. */
static void prop(stepper *s, cplx *y, double t1, double t2, int bwd) {
    double h = bwd ? (t1 - t2) : (t2 - t1);
    const cplx *Lv = s->linop;
    if (s->linop_at) {
        if (s->lastt2 != t2) s->linop_at(s->luser, t2, s->Lz);
        s->lastt2 = t2;
        Lv = s->Lz;
    }
    for (int i = 0; i < s->n; i++) {
        cplx L = Lv[i];
        double zr = creal(L) * h, zi = cimag(L) * h;
        cplx e;
        if (zi == 0.0)
            e = exp(zr) + I * zi;
        else {
            double er = exp(zr);
            e = er * cos(zi) + I * (er * sin(zi));
        }
        y[i] = cmul(y[i], e);
    }
}

/* make_fbar!, src/RK45.jl:299-309 */
static int fbar(stepper *s, cplx *out, const cplx *ybar, double t1, double t2) {
    memcpy(s->ybar, ybar, sizeof(cplx) * (size_t)s->n);
    prop(s, s->ybar, t1, t2, 0);
    int st = s->f(s->user, out, s->ybar, t2);
    s->rhs_evals++;
    if (st) return st;
    prop(s, out, t1, t2, 1);
    return LH_OK;
}

static double weaknorm(int n, const cplx *yerr, const cplx *y, const cplx *yn, double rtol, double atol) {
    double sy = 0, syn = 0, syerr = 0;
    for (int i = 0; i < n; i++) {
        sy += creal(y[i]) * creal(y[i]) + cimag(y[i]) * cimag(y[i]);
        syn += creal(yn[i]) * creal(yn[i]) + cimag(yn[i]) * cimag(yn[i]);
        syerr += creal(yerr[i]) * creal(yerr[i]) + cimag(yerr[i]) * cimag(yerr[i]);
    }
    double errwt = fmax(fmax(sqrt(sy), sqrt(syn)), atol);
    return sqrt(syerr) / rtol / errwt;
}

static void stepcontrol_pi(stepper *s) {
    const double beta1 = 3.0 / 5 / 5, beta2 = -1.0 / 5 / 5, eps = 0.8;
    if (s->ok) {
        if (s->errlast == 0) s->errlast = s->err;
        double fac;
        if (s->err == 0)
            fac = 1.5;
        else
            fac = s->safety * pow(eps / s->err, beta1) * pow(eps / s->errlast, beta2);
        s->dtn = fac * s->dt;
        s->errlast = s->err;
    } else {
        if (!isfinite(s->err))
            s->dtn = s->dt / 2;
        else
            s->dtn = s->dt * fmax(0.1, s->safety * pow(s->err, -1.0 / 5));
    }
    if (s->dtn > s->max_dt)
        s->dtn = s->max_dt;
    else if (s->dtn < s->min_dt) {
        s->dtn = s->min_dt;
        s->ok = 1;
    }
}

static int stepper_init(stepper *s, int n, const cplx *linop, orc_rhs_fn f, void *user, const cplx *y0,
                        double t, double dt, double rtol, double atol, double safety, double max_dt,
                        double min_dt, orc_linop_fn linop_at, void *luser) {
    memset(s, 0, sizeof(*s));
    s->n = n;
    s->linop = linop;
    s->linop_at = linop_at;
    s->luser = luser;
    s->lastt2 = -INFINITY; /* typemin(Float64) */
    if (linop_at) s->Lz = (cplx *)calloc((size_t)n, sizeof(cplx));
    s->f = f;
    s->user = user;
    size_t b = sizeof(cplx) * (size_t)n;
    s->y = (cplx *)malloc(b); s->yn = (cplx *)malloc(b); s->yi = (cplx *)malloc(b);
    s->yerr = (cplx *)malloc(b); s->ybar = (cplx *)malloc(b);
    for (int i = 0; i < 7; i++) s->ks[i] = (cplx *)calloc((size_t)n, sizeof(cplx));
    memcpy(s->y, y0, b);
    memcpy(s->yn, y0, b);
    s->t = s->tn = t;
    s->dt = s->dtn = dt;
    s->rtol = rtol; s->atol = atol; s->safety = safety; s->max_dt = max_dt; s->min_dt = min_dt;
    return fbar(s, s->ks[0], y0, t, t);
}

static void stepper_free(stepper *s) {
    free(s->y); free(s->yn); free(s->yi); free(s->yerr); free(s->ybar); free(s->Lz);
    for (int i = 0; i < 7; i++) free(s->ks[i]);
}

/* step!, src/RK45.jl:167-192, evaluate!(::PreconStepper) :209-223 */
static int step(stepper *s) {
    int n = s->n;
    memcpy(s->y, s->yn, sizeof(cplx) * (size_t)n);
    prop(s, s->ks[0], s->t, s->tn, 0);
    s->dt = s->dtn;
    s->t = s->tn;
    for (int ii = 0; ii < 6; ii++) {
        memcpy(s->yn, s->y, sizeof(cplx) * (size_t)n);
        for (int jj = 0; jj <= ii; jj++) {
            if (B_[ii][jj] == 0) continue;
            double c = s->dt * B_[ii][jj];
            for (int i = 0; i < n; i++) s->yn[i] = s->yn[i] + c * s->ks[jj][i];
        }
        int st = fbar(s, s->ks[ii + 1], s->yn, s->t, s->t + nodes_[ii] * s->dt);
        if (st) return st;
    }
    memcpy(s->yn, s->y, sizeof(cplx) * (size_t)n);
    for (int jj = 0; jj < 7; jj++) {
        if (b5_[jj] == 0) continue;
        double c = s->dt * b5_[jj];
        for (int i = 0; i < n; i++) s->yn[i] = s->yn[i] + c * s->ks[jj][i];
    }
    for (int i = 0; i < n; i++) s->yerr[i] = 0.0;
    for (int ii = 0; ii < 7; ii++) {
        double e = b5_[ii] - b4_[ii];
        if (e == 0) continue;
        for (int i = 0; i < n; i++) s->yerr[i] = s->yerr[i] + (s->dt * s->ks[ii][i]) * e;
    }
    s->err = weaknorm(n, s->yerr, s->y, s->yn, s->rtol, s->atol);
    s->ok = s->err <= 1;
    stepcontrol_pi(s);
    if (s->ok) {
        s->tn = s->t + s->dt;
        memcpy(s->ks[0], s->ks[6], sizeof(cplx) * (size_t)n);
    } else {
        memcpy(s->yn, s->y, sizeof(cplx) * (size_t)n);
    }
    prop(s, s->yn, s->t, s->tn, 0);
    return LH_OK;
}

/* interpolate(::PreconStepper, ti), src/RK45.jl:249-268.  Writes into out. */
static void interpolate(stepper *s, double ti, cplx *out) {
    int n = s->n;
    if (ti == s->t) {
        memcpy(out, s->y, sizeof(cplx) * (size_t)n);
        return;
    }
    if (ti == s->tn) {
        memcpy(out, s->yn, sizeof(cplx) * (size_t)n);
        return;
    }
    double sg = (ti - s->t) / s->dt;
    double sp[4] = {sg, pow(sg, 2), pow(sg, 3), pow(sg, 4)};
    double b[7];
    for (int j = 0; j < 7; j++) {
        double acc = sp[0] * interpC_[0][j];
        for (int p = 1; p < 4; p++) acc = acc + sp[p] * interpC_[p][j];
        b[j] = acc;
    }
    for (int i = 0; i < n; i++) s->yi[i] = 0.0;
    for (int j = 0; j < 7; j++)
        for (int i = 0; i < n; i++) s->yi[i] = s->yi[i] + s->ks[j][i] * b[j];
    for (int i = 0; i < n; i++) out[i] = s->y[i] + s->dt * s->yi[i];
    prop(s, out, s->t, ti, 0);
}

/* ------------------------------------------------------------------------ */
/* Exported API                                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    const lh_desc *d;
    rhs_ws w;
    int status;
    const lh_gradient_desc *g; /* gradient fill or NULL */
    cplx *nf;                  /* [n_omega] norm at the RHS's z */
} rhs_user;

static void grad_constants(const lh_desc *d, const lh_gradient_desc *g, double z, cplx *L, cplx *nf, double *rho);

/* t(nl, Eω, z): with a gradient fill, ρ(z) and norm!(nl, z) at the call's z. */
static int rhs_cb(void *user, cplx *out, const cplx *y, double t) {
    rhs_user *u = (rhs_user *)user;
    if (!u->g) return rhs_eval(u->d, &u->w, out, y);
    lh_desc dz = *u->d;
    grad_constants(u->d, u->g, t, NULL, u->nf, &dz.density);
    dz.norm = (const double *)u->nf;
    return rhs_eval(&dz, &u->w, out, y);
}

static void grad_linop_cb(void *user, double t, cplx *L) {
    rhs_user *u = (rhs_user *)user;
    grad_constants(u->d, u->g, t, L, NULL, NULL);
}

int orc_fft_c2c(int n, double *inout, int sign) {
    fft_c2c(n, (cplx *)inout, sign);
    return 0;
}

int orc_rfft(int n, const double *x, double *out) {
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)(n / 2 + 1));
    rfft(n, x, (cplx *)out, work);
    free(work);
    return 0;
}

int orc_irfft_unnorm(int n, const double *X, double *x) {
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)(n / 2 + 1));
    irfft_unnorm(n, (const cplx *)X, x, work);
    free(work);
    return 0;
}

int orc_cumtrapz(int n, const double *y, double dx, double *out) {
    cumtrapz(n, y, dx, out);
    return 0;
}

int orc_cspline_eval(int n, const double *x, const double *y, const double *D, int m, const double *x0,
                     double *out) {
    for (int i = 0; i < m; i++) out[i] = cspline_eval(n, x, y, D, x0[i]);
    return 0;
}

int orc_ppt_rate(const lh_desc *d, int n, const double *E, double *out) { return ppt_rate(d, n, E, out); }

/* One field-response evaluation on a real time series (Et_to_Pt!). */
int orc_responses_real(const lh_desc *d, const double *E, double *Pt) {
    rhs_ws w;
    ws_init(&w, d);
    int st = responses_real(d, &w, E, Pt);
    ws_free(&w);
    return st;
}

int orc_rhs(const lh_desc *d, double *nl, const double *Ew, double z) {
    (void)z;
    rhs_ws w;
    ws_init(&w, d);
    int st = rhs_eval(d, &w, (cplx *)nl, (const cplx *)Ew);
    ws_free(&w);
    return st;
}

/* Generic RK45.solve_precon with the RK45-internal output (output=true),
 * src/RK45.jl:27-89, as exercised by test/test_rk45.jl. */
typedef int (*orc_py_rhs)(void *user, double *out, const double *y, double t);

typedef struct {
    orc_py_rhs f;
    void *user;
} py_user;

static int py_rhs(void *user, cplx *out, const cplx *y, double t) {
    py_user *u = (py_user *)user;
    return u->f(u->user, (double *)out, (const double *)y, t);
}

int orc_solve_precon(int n, const double *linop, const double *y0, double t0, double dt0, double tmax,
                     double rtol, double atol, orc_py_rhs f, void *user, int outputN, double *tout,
                     double *yout, int64_t *nsteps) {
    py_user pu = {f, user};
    stepper s;
    int st = stepper_init(&s, n, (const cplx *)linop, py_rhs, &pu, (const cplx *)y0, t0, dt0, rtol, atol, 0.9,
                          INFINITY, 0.0, NULL, NULL);
    if (st) return st;
    for (int i = 0; i < outputN; i++) {
        /* range(t, stop=tmax, length=outputN) */
        long double a = (long double)t0, b = (long double)tmax;
        tout[i] = (double)((a * (long double)(outputN - 1 - i) + b * (long double)i) / (long double)(outputN - 1));
    }
    cplx *Y = (cplx *)yout;
    memcpy(Y, s.y, sizeof(cplx) * (size_t)n);
    int saved = 1;
    int64_t steps = 0;
    int repeated = 0;
    while (s.tn <= tmax) {
        st = step(&s);
        if (st) break;
        steps++;
        if (s.ok) {
            while (saved < outputN && tout[saved] < s.tn) {
                interpolate(&s, tout[saved], Y + (size_t)saved * n);
                saved++;
            }
            repeated = 0;
        } else {
            repeated++;
            if (repeated > 10) {
                st = LH_E_REPEAT_LIMIT;
                break;
            }
        }
    }
    *nsteps = steps;
    stepper_free(&s);
    return st;
}

/* ------------------------------------------------------------------------ */
/* Per-step statistics                                                       */
/* ------------------------------------------------------------------------ */
#define ORC_C_LIGHT 299792458.0
#define ORC_PI 3.141592653589793

/* NumericalIntegration.integrate(x, y, SimpsonEven()) on a uniform grid. */
static double simpson_even(int n, double dx, const double *y) {
    double r = (17 * y[0] + 59 * y[1] + 43 * y[2] + 49 * y[3] + 49 * y[n - 4] + 43 * y[n - 3] + 59 * y[n - 2] +
                17 * y[n - 1]) / 48;
    for (int i = 4; i < n - 4; i++) r += y[i];
    return dx * r;
}

/* Maths.linterpx, src/Maths.jl:248-252 */
static double linterpx(double x1, double x2, double y1, double y2, double val) {
    double slope = (y2 - y1) / (x2 - x1);
    double icpt = y1 - slope * x1;
    return (val - icpt) / slope;
}

/* Maths.level_xings(x, y; method=:linear, minmax), src/Maths.jl:162-214,
 * returning fwhm = |right - left| (src/Maths.jl:128-131); NaN where the
 * reference's try/catch returns NaN (no crossing, or a neighbour off the grid). */
static double fwhm_linear(int n, const double *x, const double *y, int minmax_max) {
    int maxidx = 0;
    for (int i = 1; i < n; i++)
        if (y[i] > y[maxidx]) maxidx = i; /* argmax: first maximum */
    double val = 0.5 * y[maxidx];
    double xmax = x[maxidx];
    int lefti = -1, righti = -1;
    if (!minmax_max) {
        for (int i = 0; i < n; i++)
            if (x[i] < xmax && y[i] < val) lefti = i; /* findlast */
        for (int i = 0; i < n; i++)
            if (x[i] > xmax && y[i] < val) { righti = i; break; } /* findfirst */
        if (lefti < 0 || righti < 0) return NAN;
        double left = linterpx(x[lefti], x[lefti + 1], y[lefti], y[lefti + 1], val);
        double right = linterpx(x[righti - 1], x[righti], y[righti - 1], y[righti], val);
        return fabs(right - left);
    }
    for (int i = 0; i < n; i++)
        if (x[i] < xmax && y[i] > val) { lefti = i; break; } /* findfirst */
    for (int i = 0; i < n; i++)
        if (x[i] > xmax && y[i] > val) righti = i; /* findlast */
    if (lefti < 1 || righti < 0 || righti + 1 >= n) return NAN; /* BoundsError → NaN */
    double left = linterpx(x[lefti - 1], x[lefti], y[lefti - 1], y[lefti], val);
    double right = linterpx(x[righti], x[righti + 1], y[righti], y[righti + 1], val);
    return fabs(right - left);
}

/* collect_stats(Eω, z, dz) with the Stats.default functions of a mode-averaged
 * propagation (src/Stats.jl:454-532): analytic!, then ω0, energy, peakpower,
 * fwhm_t, density, peakintensity(aeff), electrondensity (plasma), zdz!, and
 * energy_window for each of nwin windows (win: [nwin][n_omega] window values;
 * energyfun_ω(Eω .* window), src/Stats.jl:52-79).
 * t: grid.t [n_t]; aeff: Aeff; rec[LH_NSTAT]. */
static double energy_of(const lh_desc *d, const double *y) {
    /* Fields.energyfuncs energy_ω of |Eω|² samples y, src/Fields.jl:510-534 */
    const int N = d->n_omega;
    const double *w = d->omega;
    if (d->grid_kind == LH_GRID_REAL) {
        double prefac = 2 * ORC_PI / (w[N - 1] * w[N - 1]);
        return prefac * simpson_even(N, w[1] - w[0], y);
    }
    double dw = w[1] - w[0];
    double Dw = (double)N * dw;
    double prefac = 2 * ORC_PI * dw / (Dw * Dw);
    double s = 0.0;
    for (int i = 0; i < N; i++) s += y[i];
    return prefac * s;
}

static int step_stats(const lh_desc *d, const double *t, double aeff, int nwin, const double *win, const cplx *Ew,
                      double z, double dz, double *rec) {
    const int N = d->n_omega, nt = d->n_t;
    const double *w = d->omega;
    double *y = (double *)malloc(sizeof(double) * (size_t)(N > nt ? N : nt));
    cplx *Et = (cplx *)calloc((size_t)nt, sizeof(cplx));
    double *out = (double *)malloc(sizeof(double) * (size_t)nt);
    double *frac = (double *)malloc(sizeof(double) * (size_t)nt);
    int st = LH_OK;
    /* ω0: Maths.moment(ω, abs2.(Eω)), src/Stats.jl:18-25, src/Maths.jl:89-96 */
    double sxy = 0.0, sy = 0.0;
    for (int i = 0; i < N; i++) {
        y[i] = creal(Ew[i]) * creal(Ew[i]) + cimag(Ew[i]) * cimag(Ew[i]);
        sxy += w[i] * y[i];
        sy += y[i];
    }
    rec[LH_STAT_OMEGA0] = sxy / sy;
    rec[LH_STAT_ENERGY] = energy_of(d, y);
    /* energy_window: energyfun_ω(Eω .* window) */
    for (int q = 0; q < LH_STAT_MAX_WINDOWS; q++) {
        rec[LH_STAT_WINDOW0 + q] = NAN;
        if (q >= nwin) continue;
        double *yw = (double *)malloc(sizeof(double) * (size_t)N);
        for (int i = 0; i < N; i++) {
            cplx e = Ew[i] * win[(size_t)q * N + i];
            yw[i] = creal(e) * creal(e) + cimag(e) * cimag(e);
        }
        rec[LH_STAT_WINDOW0 + q] = energy_of(d, yw);
        free(yw);
    }
    /* analytic field: plan_analytic, src/Stats.jl:421-448 */
    if (d->grid_kind == LH_GRID_REAL) {
        for (int i = 0; i < N - 1; i++) Et[i] = 2 * Ew[i]; /* copyto_fft! */
    } else {
        for (int i = 0; i < N; i++) Et[i] = Ew[i];
    }
    fft_c2c(nt, Et, +1);
    double inv = 1.0 / (double)nt;
    for (int i = 0; i < nt; i++) Et[i] = inv * creal(Et[i]) + I * (inv * cimag(Et[i]));
    /* peakpower, fwhm_t, peakintensity: src/Stats.jl:85-99, 150-154, 182-199 */
    for (int i = 0; i < nt; i++) y[i] = creal(Et[i]) * creal(Et[i]) + cimag(Et[i]) * cimag(Et[i]);
    double pmax = y[0];
    for (int i = 1; i < nt; i++)
        if (y[i] > pmax) pmax = y[i];
    rec[LH_STAT_PEAKPOWER] = pmax;
    rec[LH_STAT_FWHM_T_MIN] = fwhm_linear(nt, t, y, 0);
    rec[LH_STAT_FWHM_T_MAX] = fwhm_linear(nt, t, y, 1);
    rec[LH_STAT_DENSITY] = d->density;
    rec[LH_STAT_PEAKINTENSITY] = pmax / aeff;
    /* electrondensity(grid, ionrate, dfun, aeff), src/Stats.jl:219-241 */
    rec[LH_STAT_ELECTRONDENSITY] = NAN;
    rec[LH_STAT_PEAK_IONISATION_RATE] = NAN;
    if (d->grid_kind == LH_GRID_REAL && d->plasma) {
        double ediv = sqrt(NLSCALE_EPS0 * ORC_C_LIGHT * aeff / 2);
        for (int i = 0; i < nt; i++) y[i] = creal(Et[i]) / ediv; /* real(Eto ./ ediv) */
        st = ppt_rate(d, nt, y, out);
        if (st == LH_OK) {
            double ratemax = out[0];
            for (int i = 1; i < nt; i++)
                if (out[i] > ratemax) ratemax = out[i];
            cumtrapz(nt, out, t[1] - t[0], frac);
            double fend = 1 - exp(-frac[nt - 1]);
            rec[LH_STAT_ELECTRONDENSITY] = fend * d->density;
            rec[LH_STAT_PEAK_IONISATION_RATE] = ratemax;
        }
    }
    rec[LH_STAT_Z] = z;
    rec[LH_STAT_DZ] = dz;
    free(y); free(Et); free(out); free(frac);
    return st;
}

/* ------------------------------------------------------------------------ */
/* Gradient fill                                                             */
/* ------------------------------------------------------------------------ */
/* p(z) of Capillary.gradient, src/Capillary.jl:306-340 */
static double grad_pressure(const lh_gradient_desc *g, double z) {
    const double *Z = g->Z, *P = g->P;
    int n = g->n_points;
    if (g->two_point) {
        double L = Z[1], p0 = P[0], p1 = P[1];
        return z > L ? p1 : z <= 0 ? p0 : sqrt(p0 * p0 + z / L * (p1 * p1 - p0 * p0));
    }
    if (z <= Z[0]) return P[0];
    if (z >= Z[n - 1]) return P[n - 1];
    int i = 0;
    for (int j = 0; j < n; j++)
        if (Z[j] < z) i = j; /* findlast(x -> x < z, Z) */
    return sqrt(P[i] * P[i] + (z - Z[i]) / (Z[i + 1] - Z[i]) * (P[i + 1] * P[i + 1] - P[i] * P[i]));
}

/* principal square root of x + iy */
static cplx csqrt_orc(double x, double y) {
    if (x == 0.0 && y == 0.0) return 0.0 + I * y;
    double t = sqrt((fabs(x) + hypot(x, y)) * 0.5);
    if (x >= 0.0) return t + I * (y / (2.0 * t));
    return fabs(y) / (2.0 * t) + I * copysign(t, y);
}

/* neff(mode, εco, nwg), src/Capillary.jl:216-234 */
static cplx grad_neff(const lh_gradient_desc *g, double eco, double nwr, double nwi) {
    cplx n;
    if (g->model == 0)
        n = csqrt_orc(eco - nwr, -nwi);
    else
        n = (1 + (eco - 1) / 2 - nwr) + I * (-nwi);
    if (!g->loss) n = creal(n);
    return n;
}

static const double FD8_[9] = {1.0 / 280, -4.0 / 105, 1.0 / 5, -4.0 / 5, 0.0, 4.0 / 5, -1.0 / 5, 4.0 / 105, -1.0 / 280};

/* The z-dependent constants at z: ρ(z) = densityspline(p(z)); the operator
 * linop!(out, z) of make_linop (src/LinearOps.jl:337-358):
 *   out = -im*(ω/c*conj_clamp(neff(εco(z), nwg), ω) - ω*β1(z)) on sidx, 0 elsewhere,
 *   β1(z) = dβ/dω at ω0 with β = ω/c*real(neff(mode, ω; z)) (8th-order stencil);
 * and norm_mode_average's factor pre/β(ω, z)*sqrt(aeff) (src/NonlinearRHS.jl:375-387).
 * L, nf, rho may be NULL. */
/* radius of the taper at z: linear a0 + (a1 - a0)·clamp(z/L, 0, 1), or the
 * piecewise-linear profile through (taper_Z, taper_A), constant outside */
static double taper_a(const lh_gradient_desc *g, double z) {
    if (g->taper == 2) {
        const double *Z = g->taper_Z, *A = g->taper_A;
        int n = g->taper_n;
        if (z <= Z[0]) return A[0];
        if (z >= Z[n - 1]) return A[n - 1];
        int i = 0;
        for (int j = 0; j < n - 1; j++)
            if (Z[j] < z) i = j; /* the segment [Z[i], Z[i+1]] holding z */
        return A[i] + (A[i + 1] - A[i]) * ((z - Z[i]) / (Z[i + 1] - Z[i]));
    }
    double f = z / g->taper_L;
    f = f < 0 ? 0 : (f > 1 ? 1 : f);
    return g->taper_a0 + (g->taper_a1 - g->taper_a0) * f;
}

/* neff_wg(m, ω; z) at radius a from U = u_nm/k and V = v_n/k
 * (src/Capillary.jl:188-214): (u/(ka))²(1 - i v_n/(ka))² for :full,
 * c²u²/(2ω²a²) [+ i c³u² v_n/(a³ω³) with loss] for :reduced. */
static cplx taper_nwg(const lh_gradient_desc *g, double U, cplx V, double a) {
    double q = U / a;
    if (g->model == 0) {
        cplx t = 1 - I * V / a;
        return (q * q) * (t * t);
    }
    cplx r = q * q / 2;
    if (g->loss) r += I * (q * q) * (V / a);
    return r;
}

static void grad_constants(const lh_desc *d, const lh_gradient_desc *g, double z, cplx *L, cplx *nf, double *rho) {
    double r = cspline_eval(g->dens_n, g->dens_p, g->dens_rho, g->dens_d, grad_pressure(g, z));
    if (rho) *rho = r;
    int N = d->n_omega;
    const cplx *nwg0 = (const cplx *)g->nwg, *pre = (const cplx *)g->pre;
    const double a = g->taper ? taper_a(g, z) : 0.0;
    const double sqrt_aeff = g->taper ? a * g->sqrt_aeff_intg : g->sqrt_aeff;
    cplx *nwg = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    for (int i = 0; i < N; i++)
        nwg[i] = g->taper ? taper_nwg(g, g->wgU[i], ((const cplx *)g->wgV)[i], a) : nwg0[i];
    if (L) {
        double acc = 0.0;
        for (int j = 0; j < 9; j++) {
            const double *f = g->fd + 4 * j;
            double co = sqrt(1 + f[1] * r);
            cplx nwj = f[2] + I * f[3];
            if (g->taper) {
                const double *t = g->fdt + 3 * j;
                nwj = taper_nwg(g, t[0], t[1] + I * t[2], a);
            }
            double nr = creal(grad_neff(g, co * co, creal(nwj), cimag(nwj)));
            if (g->model == 0 && nr < 1e-3) nr = 1e-3; /* neff(mode, ω; z)'s clamp */
            acc += FD8_[j] * (f[0] / ORC_C_LIGHT * nr);
        }
        double b1 = acc / g->fd_h;
        for (int i = 0; i < N; i++) {
            L[i] = 0.0;
            if (!d->sidx[i]) continue;
            double w = d->omega[i];
            double co = sqrt(1 + g->gamma[i] * r); /* coren(ω; z) */
            cplx n = grad_neff(g, co * co, creal(nwg[i]), cimag(nwg[i]));
            /* conj_clamp, src/LinearOps.jl:254 */
            double lim = 3000 * ORC_C_LIGHT / w;
            double ncr = creal(n) < 1e-3 ? 1e-3 : creal(n);
            double nci = cimag(n) < 0 ? 0.0 : (cimag(n) > lim ? lim : cimag(n));
            double wc = w / ORC_C_LIGHT;
            double ar = wc * ncr - w * b1, ai = wc * (-nci);
            L[i] = ai + I * (-ar); /* -im*(a) */
        }
    }
    if (nf) {
        for (int i = 0; i < N; i++) {
            nf[i] = 1.0;
            if (!d->sidx[i]) continue;
            double w = d->omega[i];
            double co = sqrt(1 + g->gamma[i] * r);
            double beta = w / ORC_C_LIGHT * creal(grad_neff(g, co * co, creal(nwg[i]), cimag(nwg[i])));
            nf[i] = (creal(pre[i]) / beta * sqrt_aeff) + I * (cimag(pre[i]) / beta * sqrt_aeff);
        }
    }
    free(nwg);
}

/* Luna.run + solve_precon + stepfun filter + GridCondition saves. */
static int solve_impl(const lh_desc *d, double *Ew_inout, const lh_rk_opts *o, double *saves, double *zsaves,
                      lh_stats *stats, const double *tgrid, double aeff, int nwin, const double *win, double *rec,
                      int cap, int *nrec, const lh_gradient_desc *g) {
    if (nrec) *nrec = 0;
    rhs_user u;
    memset(&u, 0, sizeof(u));
    u.d = d;
    u.g = g;
    ws_init(&u.w, d);
    int N = d->n_omega, nr = d->n_r, n = N * nr;
    if (g) u.nf = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    cplx *lin = (cplx *)d->linop;
    stepper s;
    memset(stats, 0, sizeof(*stats));
    int st = stepper_init(&s, n, lin, rhs_cb, &u, (const cplx *)Ew_inout, o->z0, o->init_dz, o->rtol, o->atol,
                          o->safety, o->max_dz, o->min_dz, g ? grad_linop_cb : NULL, &u);
    int nt = d->n_t;
    double *Et = (double *)malloc(sizeof(double) * (size_t)nt);
    cplx *Etc = (cplx *)malloc(sizeof(cplx) * (size_t)nt);
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)(nt / 2 + 2));
    double *grid = (double *)malloc(sizeof(double) * (size_t)(o->save_n > 0 ? o->save_n : 1));
    /* GridCondition(tmin, tmax, saveN): the output's own tmin (src/Output.jl:490-496) */
    for (int i = 0; i < o->save_n; i++) {
        long double a = (long double)o->save_z0, b = (long double)o->zmax;
        grid[i] = (o->save_n == 1) ? o->save_z0
                                   : (double)((a * (long double)(o->save_n - 1 - i) + b * (long double)i) /
                                              (long double)(o->save_n - 1));
    }
    /* a resumed run continues the saves at the cache's count (src/Output.jl:170-176) */
    int saved = o->save_first, repeated = 0;
    const long long max_steps = o->max_steps > 0 ? o->max_steps : 500000;
    while (st == LH_OK && s.tn <= o->zmax) {
        if (stats->steps >= max_steps) {
            st = LH_E_MAXSTEPS;
            break;
        }
        st = step(&s);
        if (st) break;
        stats->steps++;
        if (s.ok) {
            stats->accepted++;
            /* stepfun: src/Luna.jl:366-372 */
            cplx *yn = s.yn;
            for (int c = 0; c < nr; c++) {
                cplx *col = yn + (size_t)c * N;
                for (int i = 0; i < N; i++) col[i] = col[i] * d->omega_win[i];
                if (d->grid_kind == LH_GRID_REAL) {
                    irfft_unnorm(nt, col, Et, work);
                    double inv = 1.0 / (double)nt;
                    for (int i = 0; i < nt; i++) Et[i] = inv * Et[i];
                    for (int i = 0; i < nt; i++) Et[i] *= d->t_win[i];
                    rfft(nt, Et, col, work);
                } else {
                    memcpy(Etc, col, sizeof(cplx) * (size_t)N);
                    fft_c2c(nt, Etc, +1);
                    double inv = 1.0 / (double)nt;
                    for (int i = 0; i < nt; i++) Etc[i] = inv * Etc[i];
                    for (int i = 0; i < nt; i++) Etc[i] = Etc[i] * d->t_win[i];
                    fft_c2c(nt, Etc, -1);
                    memcpy(col, Etc, sizeof(cplx) * (size_t)N);
                }
            }
            /* MemoryOutput: statistics of the filtered state, src/Output.jl:65-67 */
            if (rec && stats->accepted <= cap) {
                lh_desc dz = *d; /* density(z) at tn */
                double aeff_z = aeff; /* aeff(z) of a taper */
                if (g) grad_constants(d, g, s.tn, NULL, NULL, &dz.density);
                if (g && g->taper) {
                    double sa = taper_a(g, s.tn) * g->sqrt_aeff_intg;
                    aeff_z = sa * sa;
                }
                st = step_stats(&dz, tgrid, aeff_z, nwin, win, s.yn, s.tn, s.dtn,
                                rec + (size_t)(stats->accepted - 1) * LH_NSTAT);
                *nrec = (int)stats->accepted;
            }
            /* MemoryOutput + GridCondition: src/Output.jl:65-81, 497-502 */
            while (saved < o->save_n && grid[saved] <= s.tn) {
                if (saves) interpolate(&s, grid[saved], (cplx *)saves + (size_t)saved * n);
                if (zsaves) zsaves[saved] = grid[saved];
                saved++;
            }
            repeated = 0;
            /* an interruption after this accepted step (the erroring stepfun of
               test/test_output.jl:226-231) */
            if (o->zstop > 0.0 && s.tn > o->zstop && s.tn <= o->zmax) st = LH_E_STOPPED;
        } else {
            stats->rejected++;
            repeated++;
            if (repeated > o->repeat_limit) st = LH_E_REPEAT_LIMIT;
        }
    }
    memcpy(Ew_inout, s.yn, sizeof(cplx) * (size_t)n);
    stats->rhs_evals = s.rhs_evals;
    stats->status = st;
    stats->saved = saved;
    stats->z = s.tn;
    stats->dz = s.dtn;
    free(Et); free(Etc); free(work); free(grid); free(u.nf);
    stepper_free(&s);
    ws_free(&u.w);
    return st;
}

int orc_solve(const lh_desc *d, double *Ew_inout, const lh_rk_opts *o, double *saves, double *zsaves,
              lh_stats *stats) {
    return solve_impl(d, Ew_inout, o, saves, zsaves, stats, NULL, 0.0, 0, NULL, NULL, 0, NULL, NULL);
}

/* Propagation with a gradient fill (make_linop + make_prop!(linop!)). */
int orc_solve_grad(const lh_desc *d, const lh_gradient_desc *g, double *Ew_inout, const lh_rk_opts *o,
                   double *saves, double *zsaves, lh_stats *stats) {
    return solve_impl(d, Ew_inout, o, saves, zsaves, stats, NULL, 0.0, 0, NULL, NULL, 0, NULL, g);
}

/* One RHS t(nl, Eω, z) with a gradient fill. */
int orc_rhs_grad(const lh_desc *d, const lh_gradient_desc *g, double *nl, const double *Ew, double z) {
    rhs_user u;
    memset(&u, 0, sizeof(u));
    u.d = d;
    u.g = g;
    ws_init(&u.w, d);
    u.nf = (cplx *)malloc(sizeof(cplx) * (size_t)d->n_omega);
    int st = rhs_cb(&u, (cplx *)nl, (const cplx *)Ew, z);
    free(u.nf);
    ws_free(&u.w);
    return st;
}

/* The gradient fill's constants at z: operator, norm factor, density. */
int orc_grad_constants(const lh_desc *d, const lh_gradient_desc *g, double z, double *L, double *nf, double *rho) {
    grad_constants(d, g, z, (cplx *)L, (cplx *)nf, rho);
    return LH_OK;
}

/* Same with per-step statistics: rec[cap][LH_NSTAT] receives the records of
 * the first cap accepted steps, *nrec their count. */
int orc_solve_stats(const lh_desc *d, double *Ew_inout, const lh_rk_opts *o, double *saves, double *zsaves,
                    lh_stats *stats, const double *t, double aeff, int nwin, const double *win, double *rec, int cap,
                    int *nrec) {
    return solve_impl(d, Ew_inout, o, saves, zsaves, stats, t, aeff, nwin, win, rec, cap, nrec, NULL);
}

/* Same with a gradient fill (and taper): ρ(z), the operator, the
 * normalisation and the statistics' Aeff at each record's z. */
int orc_solve_grad_stats(const lh_desc *d, const lh_gradient_desc *g, double *Ew_inout, const lh_rk_opts *o,
                         double *saves, double *zsaves, lh_stats *stats, const double *t, double aeff, int nwin,
                         const double *win, double *rec, int cap, int *nrec) {
    return solve_impl(d, Ew_inout, o, saves, zsaves, stats, t, aeff, nwin, win, rec, cap, nrec, g);
}

/* One record of a given state (the stats closure called directly). */
int orc_step_stats(const lh_desc *d, const double *t, double aeff, int nwin, const double *win, const double *Ew,
                   double z, double dz, double *rec) {
    if (nwin < 0 || nwin > LH_STAT_MAX_WINDOWS || (nwin > 0 && !win)) return LH_E_ARG;
    return step_stats(d, t, aeff, nwin, win, (const cplx *)Ew, z, dz, rec);
}

/* Maths.fwhm(x, y; method=:linear, minmax), src/Maths.jl:128-131. */
int orc_fwhm(int n, const double *x, const double *y, int minmax_max, double *out) {
    *out = fwhm_linear(n, x, y, minmax_max);
    return LH_OK;
}
