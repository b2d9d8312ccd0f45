/*
 * luna_hip.h -- C ABI of the MI355X propagation library (libluna_hip.so).
 *
 * The library replaces the hot path of Luna.jl's propagation:
 *   - the nonlinear transform callable  transform(nl, Eω, z)
 *       TransModeAvg  src/NonlinearRHS.jl:362-373  (field and envelope grids)
 *       TransRadial   src/NonlinearRHS.jl:466-474
 *   - the driver + stepper  Luna.run  src/Luna.jl:358-391  →  RK45.solve_precon
 *       src/RK45.jl:19-25, 27-89, 128-309, 343-406 (tableau src/dopri.jl:1-28)
 *
 * Conventions: complex arrays are interleaved (re, im) doubles, i.e. the memory
 * layout of Julia's ComplexF64 / C's double _Complex.  Multi-column arrays are
 * column-major with ω fastest: element (iω, ir) lives at iω + n_omega*ir.
 * Every entry point returns an lh_status code; lh_strerror() describes it.
 * Host pointers handed to lh_create are copied; the caller keeps ownership.
 */
#ifndef LUNA_HIP_H
#define LUNA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference raises Julia exceptions for the same events:
 * PPT field out of range  src/Ionisation.jl:262-265
 * repeated step rejection src/RK45.jl:71-77
 * bad arguments           src/Nonlinear.jl:113, src/Interface.jl:586       */
enum {
    LH_OK = 0,
    LH_E_ARG = 1,          /* invalid argument / unsupported size            */
    LH_E_HIP = 2,          /* HIP runtime error                              */
    LH_E_PPT_RANGE = 3,    /* |E| above the PPT table maximum                */
    LH_E_REPEAT_LIMIT = 4, /* more than repeat_limit consecutive rejections  */
    LH_E_NONFINITE = 5,    /* non-finite state                               */
    LH_E_NOMEM = 6,        /* device allocation failed                       */
    LH_E_NODEVICE = 7,     /* no HIP device available                        */
    LH_E_MAXSTEPS = 8,     /* step-attempt cap (lh_rk_opts.max_steps) reached:
                              no counterpart in the reference, whose loop is
                              unbounded (src/RK45.jl:55-89)                  */
    LH_E_STOPPED = 9,      /* stopped at lh_rk_opts.zstop before zmax: the
                              analogue of a run interrupted after an accepted
                              step (the cached-continuation test,
                              test/test_output.jl:203-276); resume from the
                              returned state, z, dz and saved count          */
    LH_E_TOO_LARGE = 10,   /* the problem does not fit the fixed working
                              memory of lh_solve_noalloc                     */
    LH_E_HOST_WAIT = 11    /* lh_solve_stream with every_step: the host fell
                              behind the device by more than the ring of
                              per-step slots (the run was stopped)            */
};

enum { LH_GRID_REAL = 0, LH_GRID_ENV = 1 };           /* Grid.RealGrid / Grid.EnvGrid */
enum { LH_TRANS_MODEAVG = 0, LH_TRANS_RADIAL = 1 };   /* TransModeAvg / TransRadial  */
enum { LH_KERR_NONE = 0, LH_KERR_FIELD = 1, LH_KERR_ENV = 2 };

/* Everything a transform closure captures in the reference
 * (src/NonlinearRHS.jl:323-399, 406-524; src/Nonlinear.jl:24-177, 201-333). */
typedef struct lh_desc {
    int32_t grid_kind;        /* LH_GRID_REAL or LH_GRID_ENV                     */
    int32_t transform;        /* LH_TRANS_MODEAVG or LH_TRANS_RADIAL             */
    int32_t n_omega;          /* Nω: rows of the state                           */
    int32_t n_t;              /* coarse time samples (stepfun filter grid)       */
    int32_t n_to;             /* oversampled time samples (RHS grid)             */
    int32_t n_r;              /* radial columns (1 for mode-averaged)            */
    double dt_o;              /* δto, the RHS time step                          */
    const double *omega;      /* [n_omega] ω                                     */
    const uint8_t *sidx;      /* [n_omega] real-frequency mask                   */
    const double *omega_win;  /* [n_omega] ωwin                                 */
    const double *t_win;      /* [n_t]   twin  (stepfun filter)                  */
    const double *to_win;     /* [n_to]  towin (RHS apodisation)                 */
    const double *linop;      /* complex [n_omega*n_r] constant linear operator  */
    const double *norm;       /* complex [n_omega*n_r] normalisation factor:
                                 mode-average: (pre/β)·√Aeff (applied on sidx,
                                 then ×ωwin on sidx); radial: ωwin·(-iω)/(2·norm)
                                 applied on every bin                            */
    double field_div;         /* mode-average: Eto /= nlscale·√Aeff(z); radial 1 */
    double density;           /* ρ (constant in z)                               */
    int32_t kerr;             /* LH_KERR_*                                       */
    int32_t plasma;           /* 1: PlasmaCumtrapz with the PPT spline           */
    double gamma3;            /* γ3 (or χ3 for the GNLSE)                        */
    int32_t ppt_n;            /* spline nodes                                    */
    int32_t raman;            /* 1: RamanPolarEnv (envelope grid only)           */
    const double *ppt_x;      /* [ppt_n] uniform node positions E                */
    const double *ppt_y;      /* [ppt_n] log(rate)                               */
    const double *ppt_d;      /* [ppt_n] CSpline derivatives                     */
    double ionpot;            /* ionisation potential (J)                        */
    double preionfrac;        /* pre-ionisation fraction                         */
    const double *raman_h;    /* [n_to] windowed h(t) on tt = 0..n_to-1 (×δto)   */
    const double *qdht;       /* [n_r*n_r] QDHT matrix T, row-major T[i*n_r+j]:
                                 T_ij = (2/S) J0(j_i j_j/S) / J1(j_j)^2 (Hankel.jl
                                 QDHT, not symmetric); radial transform only     */
    double qdht_fwd;          /* r→k scale R/K                                   */
    double qdht_inv;          /* k→r scale K/R                                   */
} lh_desc;

/* RK45 / Luna.run options (src/Luna.jl:358-362 defaults; src/RK45.jl:19-28). */
typedef struct lh_rk_opts {
    double z0;                /* 0.0                                             */
    double zmax;              /* grid.zmax                                       */
    double rtol;              /* 1e-6                                            */
    double atol;              /* 1e-10                                           */
    double safety;            /* 0.9                                             */
    double max_dz;            /* zmax/2                                          */
    double min_dz;            /* 0                                               */
    double init_dz;           /* 1e-4                                            */
    int32_t repeat_limit;     /* 10                                              */
    int32_t save_n;           /* 201 (MemoryOutput GridCondition points)         */
    /* The save grid is range(save_z0, zmax, save_n): GridCondition(tmin, tmax,
     * saveN) takes the output's tmin, not the solver's z0 (src/Output.jl:490-496). */
    double save_z0;           /* 0                                               */
    int32_t save_first;       /* save points already written: a run resumed from a
                                 cache continues GridCondition at this index, as
                                 HDF5Output's `saved` (src/Output.jl:170-176,
                                 473-487; Luna.run src/Luna.jl:374-379) (0)      */
    int32_t every_step;       /* lh_solve_stream: 1 hands every accepted step to
                                 the callback, as the reference calls its output
                                 at each (src/Luna.jl:366-372, src/Output.jl:
                                 65-81); 0: only steps that save (0)            */
    double zstop;             /* > 0: end with LH_E_STOPPED after the first
                                 accepted step with tn > zstop (its window filter
                                 and saves done); 0: run to zmax                 */
    int64_t max_steps;        /* cap on step attempts, LH_E_MAXSTEPS when reached;
                                 0: 500000                                       */
} lh_rk_opts;

/* Per-run counters (src/RK45.jl:38-82). */
typedef struct lh_stats {
    int64_t steps;            /* step attempts                                   */
    int64_t accepted;
    int64_t rejected;
    int64_t rhs_evals;
    int32_t status;           /* lh status code of this run/instance             */
    int32_t saved;            /* saves written                                   */
    double z;                 /* final tn                                        */
    double dz;                /* final dtn                                       */
} lh_stats;

typedef struct lh_ctx lh_ctx;

const char *lh_strerror(int status);
int lh_device_count(int *count);

/* Create a transform context: uploads the constants of *desc to the current
 * HIP device.  Replaces TransModeAvg(...) / TransRadial(...) construction,
 * src/NonlinearRHS.jl:340-350, 428-460, and Luna.setup src/Luna.jl:115-152,239-279.
 * Radial: field grid (Nto 2048-16384, Kerr_field/PPT plasma) or envelope grid
 * (Nt = Nto 256-8192, Kerr_env), Nr a multiple of 64 up to 4096; other
 * configurations give LH_E_ARG. */
int lh_create(lh_ctx **ctx, const lh_desc *desc);
int lh_destroy(lh_ctx *ctx);

/* One RHS evaluation on host arrays: nl = transform(Eω, z).
 * Replaces (t::TransModeAvg)(nl, Eω, z) src/NonlinearRHS.jl:362-373 and
 * (t::TransRadial)(nl, Eω, z) src/NonlinearRHS.jl:466-474.
 * nl, Ew: complex [n_omega*n_r].  Synchronous. */
int lh_rhs(lh_ctx *ctx, double *nl, const double *Ew, double z);

/* Same on device pointers, enqueued on stream (hipStream_t, NULL = default).
 * Errors from inside the kernel are reported by lh_ctx_status after a sync. */
int lh_rhs_device(lh_ctx *ctx, void *nl_dev, const void *Ew_dev, double z, void *stream);
int lh_ctx_status(lh_ctx *ctx);

/* Full propagation: Luna.run + RK45.solve_precon with the per-step ω/t window
 * filter (src/Luna.jl:366-372) and MemoryOutput/GridCondition saves
 * (src/Output.jl:65-81, 490-502).  The whole adaptive loop runs on the device.
 * Ew_inout: complex [n_omega*n_r], initial field in, final state out.
 * saves: complex [save_n][n_omega*n_r] (may be NULL); zsaves: [save_n] (may be NULL). */
int lh_solve(lh_ctx *ctx, double *Ew_inout, const lh_rk_opts *opts,
             double *saves, double *zsaves, lh_stats *stats);

/* Per-step output while the device solve runs: the output callable of
 * Luna.run's stepfun, output(y, z, dz, interpolant) (src/Luna.jl:366-372,
 * src/Output.jl:65-81, 262-302), for outputs other than a fixed save buffer
 * (HDF5Output, user callbacks, statistics evaluated on the host).  The kernel
 * publishes every accepted step that writes a GridCondition save (with
 * lh_rk_opts.every_step, every accepted step); the host hands each publication
 * to cb on the calling thread while the kernel goes on.  One call carries the
 * saves written since the previous call, the state, z and dz of the latest
 * published step (HDF5Output's cache: t, dt, y, saved) and the statistics
 * records of the steps since the previous call.  Without every_step a
 * publication the host has not taken before the next one is folded into it
 * (its saves and records are still handed over); with every_step each is
 * handed over on its own from a ring of page-locked slots (64 MiB, 64 to 4096
 * states) that the kernel fills without waiting: a host that falls behind by
 * the whole ring stops the run with LH_E_HOST_WAIT.  Pointers are valid
 * during the call only. */
typedef struct lh_step_out {
    int32_t first_save;       /* index of the first save in this call           */
    int32_t n_saves;          /* saves in this call (>= 1; >= 0 with every_step) */
    const double *saves;      /* complex [n_saves][n_omega*n_r]                 */
    const double *zsaves;     /* [n_saves] their positions                      */
    double z;                 /* tn of the latest published step                */
    double dz;                /* dtn after it                                   */
    const double *y;          /* complex [n_omega*n_r]: the filtered state at z */
    int64_t accepted;         /* accepted steps up to that step                 */
    int32_t first_record;     /* index of the first statistics record          */
    int32_t n_records;        /* records in this call                           */
    const double *records;    /* [n_records][LH_NSTAT]                          */
} lh_step_out;

/* Return nonzero to stop the run (as an error thrown by the reference's
 * output or stepfun ends the propagation): nothing more is handed over, the
 * solve ends after its next published step with LH_E_STOPPED. */
typedef int (*lh_output_cb)(void *user, const lh_step_out *out);

/* lh_solve with the saves and records streamed to cb instead of filled into
 * buffers; same trajectory, bit for bit.  Mode-averaged contexts with a
 * constant fill, and radial contexts (whose steps the host replays one at a
 * time here, copying each published state out of device memory); LH_E_ARG for
 * z-dependent fills, whose outputs are filled from lh_solve's buffers after
 * the run. */
int lh_solve_stream(lh_ctx *ctx, double *Ew_inout, const lh_rk_opts *opts, lh_output_cb cb, void *user,
                    lh_stats *stats);

/* Batched solve: n independent instances (each its own grid, constants, step
 * size control and status) in one persistent launch per grid class, one
 * workgroup per instance.  Radial contexts (whose one instance fills the GPU)
 * are solved one after another by a graph-replayed per-step launch sequence.
 * Replaces the per-point loop of Scans.runscan src/Scans.jl:275-430.
 * Ew_inout[i]: instance i's field; saves[i] may be NULL; stats[i] per instance.
 * Returns LH_OK when the launch ran; per-instance failures are in stats[i].status. */
int lh_batch_solve(lh_ctx *const *ctxs, int n, double *const *Ew_inout,
                   const lh_rk_opts *opts, double *const *saves, double *const *zsaves,
                   lh_stats *stats);

/* lh_batch_solve with options per instance, opts[i] for ctxs[i] (its own z
 * range, tolerances, save grid, stop and cap): the batch of Scans.runscan's
 * points when they differ in length or output (src/Scans.jl:275-289). */
int lh_batch_solve_opts(lh_ctx *const *ctxs, int n, double *const *Ew_inout, const lh_rk_opts *opts,
                        double *const *saves, double *const *zsaves, lh_stats *stats);

/* Device-resident batch path (no counterpart in the reference; it splits
 * lh_batch_solve so a sweep can be rerun on inputs already in HBM):
 * lh_set_state uploads an instance's initial field once; lh_batch_solve_resident
 * solves every instance from its resident initial field and leaves the final
 * state and saves in device memory; lh_get_state / lh_get_saves copy them out. */
int lh_set_state(lh_ctx *ctx, const double *Ew);
int lh_batch_solve_resident(lh_ctx *const *ctxs, int n, const lh_rk_opts *opts, int want_saves,
                            lh_stats *stats);
int lh_get_state(lh_ctx *ctx, double *Ew);
int lh_get_saves(lh_ctx *ctx, double *saves, double *zsaves);

/* z-dependent gas fill of a constant-radius capillary (pressure gradient),
 * mode-averaged field grid: Capillary.gradient (src/Capillary.jl:306-340)
 * gives p(z) and ρ(z) = densityspline(p(z)) (src/PhysData.jl:719-723), the
 * core index √(1 + γ(λ)ρ(z)); the linear operator is then re-evaluated at
 * every new z, as LinearOps.make_linop's linop!(out, z) (src/LinearOps.jl:
 * 337-358) inside make_prop!(linop!) (src/RK45.jl:283-296), and the RHS uses
 * ρ(z) and norm_mode_average's pre/β(ω, z)·√Aeff (src/NonlinearRHS.jl:375-387). */
typedef struct lh_gradient_desc {
    int32_t n_points;         /* pressure-profile nodes (>= 2)                   */
    int32_t two_point;        /* 1: gradient(gas, L, p0, p1), Z = {0, L};
                                 0: gradient(gas, Z, P), the multi-point form    */
    const double *Z;          /* [n_points] positions (m), increasing            */
    const double *P;          /* [n_points] pressures (bar)                      */
    int32_t dens_n;           /* density spline nodes (uniform in pressure)      */
    int32_t model;            /* Marcatili model: 0 :full, 1 :reduced            */
    const double *dens_p;     /* [dens_n] range(Pmin, Pmax, dens_n)              */
    const double *dens_rho;   /* [dens_n] density at dens_p                      */
    const double *dens_d;     /* [dens_n] CSpline derivatives                    */
    int32_t loss;             /* 1: complex neff (MarcatiliMode loss = true)     */
    int32_t taper;            /* 1: linear taper a(z) = a0 + (a1 - a0)·clamp(z/L,
                                 0, 1); 2: piecewise-linear a(z) through the
                                 nodes (taper_Z, taper_A), constant outside
                                 (MarcatiliMode with radius a(z),
                                 src/Capillary.jl:60-81, 202-214, 265-296);
                                 0: constant radius                               */
    const double *gamma;      /* [n_omega] sellmeier_gas γ(λ) per particle       */
    const double *nwg;        /* complex [n_omega] neff_wg(mode, ω), z-independent */
    const double *pre;        /* complex [n_omega] norm_mode_average's pre(ω)    */
    double sqrt_aeff;         /* √Aeff                                           */
    double fd_h;              /* β1 stencil step h                               */
    const double *fd;         /* [9][4] β1 = dβ/dω at ω0 by the 8th-order central
                                 stencil: per point ω_j = ω0 + (j-4)h, γ(ω_j),
                                 neff_wg(ω_j) (re, im)                            */
    /* taper only (otherwise ignored): neff_wg(ω, a) = (U/a)²(1 - iV/a)² (:full)
     * or (U/a)²/2 + i(U/a)²V/a (:reduced, the second term with loss), with
     * U = u_nm/k and V = v_n/k, k = ω/c (src/Capillary.jl:188-214)            */
    double taper_a0, taper_a1; /* radius at z <= 0 and z >= taper_L (m)         */
    double taper_L;           /* taper length (m)                                */
    double sqrt_aeff_intg;    /* √(Aeff/a²): Aeff(z) = a(z)²·aeff_intg           */
    const double *wgU;        /* [n_omega] U                                      */
    const double *wgV;        /* complex [n_omega] V                              */
    const double *fdt;        /* [9][3] U, Re V, Im V at the stencil points       */
    int32_t taper_n;          /* taper = 2: nodes (>= 2)                          */
    int32_t taper_reserved;   /* 0                                               */
    const double *taper_Z;    /* [taper_n] positions (m), increasing              */
    const double *taper_A;    /* [taper_n] radii (m)                              */
} lh_gradient_desc;

/* Attach a gradient fill to a mode-averaged field-grid context (g = NULL:
 * constant fill again).  Host arrays are copied. */
int lh_set_gradient(lh_ctx *ctx, const lh_gradient_desc *g);

/* PPT ionisation-rate table (set-up of PlasmaCumtrapz): ionrate_fun_PPT
 * (src/Ionisation.jl:317-390, φ :403-429) at n field values on the GPU, as
 * makePPTcache tabulates it (src/Ionisation.jl:185-200) before makePPTaccel
 * fits its log-rate spline.  One thread per field value. */
typedef struct lh_ppt_desc {
    double ionpot;            /* ionisation potential (J)                        */
    double lambda0;           /* wavelength (m)                                   */
    double Z;                 /* ion charge (PhysData.quantum_numbers)            */
    int32_t l;                /* orbital angular momentum: 0 or 1                 */
    int32_t msum;             /* 1: sum over m = -l..l (default); 0: m = 0 only   */
    double dalpha;            /* polarisability difference Δα (Stark shift), or 0 */
    double alpha_ion_au;      /* ion polarisability (a.u., dipole correction), or 0 */
    double sum_tol;           /* converge_series rtol (1e-6)                      */
    double occupancy;         /* 2                                                */
    int32_t cycle_average;    /* 0                                                */
    int32_t reserved;         /* 0                                                */
    double cnl;               /* Cnl, or 0: 2^2ns/(ns Γ(ns+ls+1) Γ(ns-ls))        */
    double c, au_energy, au_time, au_efield;  /* the constants PhysData uses      */
} lh_ppt_desc;

/* rate[i] = PPT rate at E[i] > 0 (synchronous, on the current device);
 * ms (may be NULL) receives the kernel's duration. */
int lh_ppt_table(const lh_ppt_desc *p, int n, const double *E, double *rate, double *ms);

/* Per-step statistics of a mode-averaged propagation: the set that
 * Stats.default builds (src/Stats.jl:476-532) and collect_stats evaluates
 * (src/Stats.jl:454-468) from the analytic field (plan_analytic,
 * src/Stats.jl:421-448) on every accepted step (MemoryOutput,
 * src/Output.jl:65-67), computed on the device right after the step's window
 * filter.  One record of LH_NSTAT doubles per accepted step:            */
enum {
    LH_STAT_Z = 0,                 /* z   (zdz!, src/Stats.jl:409-412)          */
    LH_STAT_DZ = 1,                /* dz: the next step size                    */
    LH_STAT_OMEGA0 = 2,            /* ω0: first moment of |Eω|², :18-25         */
    LH_STAT_ENERGY = 3,            /* energy, :32-43 with Fields.energyfuncs    */
    LH_STAT_PEAKPOWER = 4,         /* max |Et|² of the analytic field, :85-99   */
    LH_STAT_PEAKINTENSITY = 5,     /* peakpower / Aeff, :150-154                */
    LH_STAT_FWHM_T_MIN = 6,        /* Maths.fwhm(t, |Et|², :linear, :min), :182-199 */
    LH_STAT_FWHM_T_MAX = 7,        /* same, minmax = :max                       */
    LH_STAT_ELECTRONDENSITY = 8,   /* ρ·(1 - exp(-∫rate dt)), :219-241; NaN without plasma */
    LH_STAT_PEAK_IONISATION_RATE = 9, /* max rate, same function; NaN without plasma */
    LH_STAT_DENSITY = 10,          /* ρ (density(dfun), :325-329)               */
    LH_STAT_WINDOW0 = 11,          /* energy in window w = 0..3 (Stats.default's */
    LH_STAT_WINDOW1 = 12,          /* `windows`: energy_λ/energy_window,         */
    LH_STAT_WINDOW2 = 13,          /* :52-79): Σ_k window_w[w][k]·|Eω_k|²;       */
    LH_STAT_WINDOW3 = 14,          /* NaN where no window is set                 */
    LH_NSTAT = 15
};
#define LH_STAT_MAX_WINDOWS 4

/* What the statistics need beyond the transform's own constants. */
typedef struct lh_stats_desc {
    int32_t capacity;         /* records kept per run: the first `capacity`
                                 accepted steps                                   */
    int32_t n_windows;        /* energy windows, 0..LH_STAT_MAX_WINDOWS          */
    const double *energy_w;   /* [n_omega] energy = Σ_k energy_w[k]·|Eω_k|²: the
                                 quadrature of Fields.energyfuncs' energy_ω
                                 (src/Fields.jl:510-534) with its prefactor      */
    const double *t;          /* [n_t] grid.t, the coarse time axis (fwhm_t)     */
    double aeff;              /* effective area Aeff (constant in z)             */
    const double *window_w;   /* [n_windows][n_omega]: energy_w·window² of each
                                 window (energyfun_ω(Eω .* window)); NULL if none */
} lh_stats_desc;

/* Switch per-step statistics on for later solves of ctx (sd = NULL: off).
 * Mode-averaged contexts only (LH_E_ARG for radial ones).  Host arrays are
 * copied. */
int lh_set_step_stats(lh_ctx *ctx, const lh_stats_desc *sd);

/* One record for a given state: the stats closure called directly,
 * stats(Eω, z, dz) (src/Stats.jl:460-467).  Ew: complex [n_omega]; rec: [LH_NSTAT].
 * Synchronous; requires lh_set_step_stats. */
int lh_stats_eval(lh_ctx *ctx, const double *Ew, double z, double dz, double *rec);

/* Records of the last solve of ctx: copies min(kept, capacity) records into
 * rec[capacity][LH_NSTAT] and sets *n to that count; kept is the smaller of
 * the run's accepted steps and the capacity given to lh_set_step_stats. */
int lh_get_step_stats(lh_ctx *ctx, double *rec, int capacity, int *n);

/* Measurement hook (no counterpart in the reference): HIP-event durations of
 * the persistent launches of the last batch, one per grid-size class
 * (field grid Nto = 16384, 8192, 4096; envelope grid Nto = 8192, 4096; all
 * radial contexts together; envelope grid Nto = 16384; field grid Nto =
 * 32768), with the number of instances in each.  nclass must be >= 8. */
int lh_last_launch_times(double *ms, int *n_inst, int nclass);

/* Page-locked host memory for the states and saves a sweep copies back (no
 * counterpart in the reference, whose arrays live in host memory already):
 * device-to-host copies into it run at DMA speed and touch no unmapped pages.
 * lh_host_alloc: *ptr = bytes of pinned memory on the current device's host
 * (LH_E_NOMEM if refused); lh_host_free releases it (NULL is a no-op). */
int lh_host_alloc(size_t bytes, void **ptr);
int lh_host_free(void *ptr);

/* Plans of the radial RHS and their wisdom.  Counterpart of the reference's
 * FFTW planning: Luna.setup plans its transforms with settings["fftw_flag"]
 * (FFTW.PATIENT by default, src/Luna.jl:13-14, 239-258) between
 * Utils.loadFFTwisdom and Utils.saveFFTwisdom (src/Utils.jl:84-108).  A plan
 * is one variant of the radial RHS launch sequence: the radix plan of the
 * per-column transforms (0: radix 8, radix 16 from 8192 points; 1: radix 16
 * throughout) and the FP64 MFMA tiles per wave along t of the two Hankel
 * GEMMs (1, 2 or 4).  Every plan computes the same RHS; the radix plan moves
 * its last bits.  Radial contexts only (LH_E_ARG otherwise): the
 * mode-averaged grids run one persistent kernel per size class with a fixed
 * radix plan.
 *   lh_get_plan / lh_set_plan  the context's plan (a set drops the captured
 *                              step graph, re-captured on the next solve)
 *   lh_autotune   times every plan on one RHS of the field Ew (host, complex
 *                 [n_omega·n_r]; NULL: a zero field), reps timed launches
 *                 each after one untimed; installs the fastest plan whose RHS
 *                 agrees with the default plan's to 1e-12 (relative L2) in
 *                 the context and in the wisdom; tried[0..min(cap,n)-1]
 *                 receives every plan with its time (default plan first) and
 *                 *n_tried their count (6)
 *   lh_wisdom_export  the wisdom as text ("luna-hip-wisdom 1", then one line
 *                 "radial n_to n_r fft_radix_plan gemm_tiles rhs_us" per
 *                 shape); *len = bytes needed with the NUL; buf may be NULL
 *   lh_wisdom_import  merges such text (LH_E_ARG and nothing merged if any
 *                 line is malformed); contexts created afterwards for a shape
 *                 in the wisdom take its plan
 *   lh_wisdom_forget  empties the wisdom */
typedef struct lh_plan {
    int fft_radix_plan;
    int gemm_tiles;
    double rhs_us; /* measured microseconds per RHS (lh_autotune), else 0 */
    int agrees;    /* 1 if the plan's RHS agreed with the default plan's */
} lh_plan;
int lh_get_plan(const lh_ctx *ctx, lh_plan *plan);
int lh_set_plan(lh_ctx *ctx, const lh_plan *plan);
int lh_autotune(lh_ctx *ctx, const double *Ew, int reps, lh_plan *tried, int cap, int *n_tried);
int lh_wisdom_export(char *buf, size_t cap, size_t *len);
int lh_wisdom_import(const char *text);
int lh_wisdom_forget(void);

/* Allocation-free solve.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same device solve as
 * lh_solve on one mode-averaged context (field or envelope grid), with the
 * solver's working memory -- the instance record, the run counters, the save
 * grid, the second state buffer, the seven stage vectors and the envelope
 * Raman stash -- in one fixed buffer of LH_NOALLOC_SCRATCH_BYTES per GPU,
 * which the library holds until the process ends.  The call allocates nothing
 * (device or host): set the context up beforehand (lh_create, and
 * lh_set_step_stats / lh_set_gradient if used).
 *   Ew_dev    device pointer, complex [n_omega]: initial field in, final state out
 *   saves_dev device pointer, complex [save_n][n_omega], or NULL (no saves)
 *   zsaves    host array [save_n] receiving the save positions, or NULL
 *   stream    hipStream_t to run on (NULL: the context's stream); the call
 *             returns once that stream has finished the solve
 * Returns the run's status as lh_solve does; LH_E_TOO_LARGE when the working
 * memory for n_omega and save_n exceeds LH_NOALLOC_SCRATCH_BYTES (the field
 * grids and the envelope grids up to Nt = 8192 fit with up to 30000 saves;
 * the Nt = 16384 envelope grid does not), LH_E_ARG for radial contexts.
 * Every call on one GPU uses that GPU's buffer: calls that overlap in time on
 * one GPU race and must not be issued. */
#define LH_NOALLOC_SCRATCH_BYTES (1536u * 1024u)
int lh_solve_noalloc(lh_ctx *ctx, void *Ew_dev, const lh_rk_opts *opts, void *saves_dev, double *zsaves,
                     lh_stats *stats, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LUNA_HIP_H */
