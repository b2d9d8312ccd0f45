// lh_kernels.hip -- MI355X (gfx950) kernels and C ABI for the propagation hot path.
//
// Design: one workgroup (512 threads, 8 wave64s) owns one instance.  A right-hand-side
// evaluation keeps the whole oversampled time series in LDS:
//   stage prologue (ȳ = y + Δz Σ B k, ·exp(LΔ)) → c2r pre-pack → inverse
//   Stockham FFT → Kerr + PPT spline + three cumtrapz block scans + towin →
//   forward FFT → r2c post-step, crop, normalisation → stage epilogue
//   (·exp(-LΔ)) written to HBM.
// The adaptive Dormand–Prince loop (controller, window filter, dense-output
// saves) runs inside the same workgroup, so a whole propagation is one launch
// and a batch of independent instances is one launch of many workgroups.
//
// Reference semantics (file:line in the reference tree):
//   TransModeAvg              src/NonlinearRHS.jl:362-387
//   to_time!/to_freq!         src/NonlinearRHS.jl:16-53
//   Kerr_field, PlasmaCumtrapz src/Nonlinear.jl:8-10, 24-34, 88-177
//   makePPTaccel + CSpline    src/Ionisation.jl:247-271, src/Maths.jl:771-775, 840-848
//   cumtrapz!                 src/Maths.jl:299-327
//   PreconStepper/step!/evaluate!/interpolate, weaknorm, stepcontrolPI!
//                             src/RK45.jl:128-309, 343-406; tableau src/dopri.jl
//   Luna.run stepfun filter   src/Luna.jl:366-372
//   GridCondition saves       src/Output.jl:65-81, 490-502
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/luna_hip.h"
#include "lh_device.h"

#define LH_E_MAXSTEPS_INTERNAL LH_E_REPEAT_LIMIT
static const long long LH_MAX_STEPS = 500000;  // hard cap on step attempts per instance

// --------------------------------------------------------------------------
// Device-side instance description
// --------------------------------------------------------------------------
struct DInst {
    int grid_kind, transform, n_omega, n_t, n_to, n_r;
    int kerr, plasma, raman, ppt_n;
    double field_div, density, kerr_fac, ionpot, preion_p1, e_ratio;
    double dt_o, inv_nto, inv_nt, sc_t, sc_f, ppt_emin, ppt_emax;
    const double* omega_win;
    const uint8_t* sidx;
    const double* t_win;
    const double* to_win;
    const double2* linop;
    const double2* norm;
    const double4* ppt;  // (x, y, D, 0) per node
    // solver state
    double2* Y;
    double2* YN;
    double2* K[7];
    double2* saves;      // [save_n][n_state] or null
    const double* zgrid; // [save_n]
    lh_stats* stats;
    double z0, zmax, rtol, atol, safety, max_dz, min_dz, init_dz;
    int repeat_limit, save_n;
    // single-RHS I/O
    const double2* rhs_in;
    double2* rhs_out;
};

// Dormand–Prince tableau, src/dopri.jl:1-28 (rational literals evaluated in double).
__constant__ double c_B[6][6] = {
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0},
    {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}};
__constant__ double c_nodes[6] = {1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1, 1};
__constant__ double c_b5[7] = {5179.0 / 57600, 0, 7571.0 / 16695, 393.0 / 640, -92097.0 / 339200, 187.0 / 2100, 1.0 / 40};
__constant__ double c_b4[7] = {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84, 0};
__constant__ double c_iC[4][7] = {
    {1, 0, 0, 0, 0, 0, 0},
    {-183.0 / 64, 0, 1500.0 / 371, -125.0 / 32, 9477.0 / 3392, -11.0 / 7, 3.0 / 2},
    {37.0 / 12, 0, -1000.0 / 159, 125.0 / 12, -729.0 / 106, 11.0 / 3, -4},
    {-145.0 / 128, 0, 1000.0 / 371, -375.0 / 64, 25515.0 / 6784, -55.0 / 28, 5.0 / 2}};

struct Smem {
    double2 buf[padded_len(8192)];
    double edge[LH_T];
    double wsum[5 * LH_NW];
    int flag;
};

// --------------------------------------------------------------------------
// PPT rate: exp(CSpline(|E|)) with the uniform index rule; |E| above the
// table raises the range flag (src/Ionisation.jl:256-269).
// --------------------------------------------------------------------------
__device__ __forceinline__ double ppt_rate(const DInst& in, double aE, int& bad) {
    if (aE < in.ppt_emin) return 0.0;
    if (!(aE <= in.ppt_emax)) {  // also catches NaN: no table index is formed
        bad = (aE > in.ppt_emax) ? LH_E_PPT_RANGE : LH_E_NONFINITE;
        return 0.0;
    }
    const double xmin = in.ppt_emin, xmax = in.ppt_emax;
    const int n = in.ppt_n;
    int i;
    if (aE <= xmin)
        i = 2;
    else if (aE >= xmax)
        i = n;
    else
        i = (int)ceil((aE - xmin) / (xmax - xmin) * (double)(n - 1)) + 1;
    i = min(max(i, 2), n);
    double4 pc = in.ppt[i - 1], pm = in.ppt[i - 2];
    if (aE == pc.x) return exp(pc.y);
    if (aE == pm.x) return exp(pm.y);
    double t = (aE - pm.x) / (pc.x - pm.x);
    double v = (pm.y + pm.z * t + (3 * (pc.y - pm.y) - 2 * pm.z - pc.z) * (t * t) +
                (2 * (pm.y - pc.y) + pm.z + pc.z) * (t * t * t));
    return exp(v);
}

// cumtrapz over the workgroup: thread owns Q consecutive samples y[q]
// (global index g0+q); on return y[q] holds the running integral.
template <int Q>
__device__ __forceinline__ void block_cumtrapz(double* y, double dx, int g0, Smem& sm) {
    const int tid = threadIdx.x;
    sm.edge[tid] = y[Q - 1];
    __syncthreads();
    double prev = (tid > 0) ? sm.edge[tid - 1] : 0.0;
    double acc = 0.0;
    double last = prev;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        double cur = y[q];
        double term = (g0 + q == 0) ? 0.0 : (0.5 * (last + cur)) * dx;
        last = cur;
        acc += term;
        y[q] = acc;
    }
    double off = block_exscan(acc, sm.wsum);
#pragma unroll
    for (int q = 0; q < Q; q++) y[q] = off + y[q];
}

// Pointwise field responses on the packed real series in LDS (in place):
// E = (c2r/Nto)/field_div; P = ρε0γ3 E³ + ρ·Plasma(E); P *= towin.
template <int MO>
__device__ __forceinline__ double field_sample(const Smem& sm, int tid, int q, double inv_nto, double fdiv) {
    constexpr int S = MO / LH_T;
    double2 v = sm.buf[pad(S * tid + (q >> 1))];
    return (inv_nto * ((q & 1) ? v.y : v.x)) / fdiv;
}

// The packed real series stays in LDS (this thread's Q samples are only
// touched by this thread); scan operands live in registers.
template <int MO>
__device__ __forceinline__ void pointwise_field(const DInst& in, Smem& sm) {
    constexpr int S = MO / LH_T;
    constexpr int Q = 2 * S;
    const int tid = threadIdx.x;
    const int g0 = Q * tid;
    const double inv_nto = in.inv_nto, fdiv = in.field_div;
    double a[Q];  // plasma polarisation (zero without plasma)
    if (in.plasma) {
        double jf[Q];
        int bad = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            double E = field_sample<MO>(sm, tid, q, inv_nto, fdiv);
            a[q] = ppt_rate(in, fabs(E), bad);
            jf[q] = a[q];  // rate, kept for the loss term
        }
        if (bad) sm.flag = bad;
        const double dt = in.dt_o;
        block_cumtrapz<Q>(a, dt, g0, sm);  // a = ∫rate
        const double pp1 = in.preion_p1, er = in.e_ratio, ip = in.ionpot;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            double E = field_sample<MO>(sm, tid, q, inv_nto, fdiv);
            double f = pp1 - exp(-a[q]);
            double rate = jf[q];
            a[q] = (f * er) * E;                                           // phase
            jf[q] = (E != 0.0) ? ((ip * rate) * (1 - f)) / E : 0.0;  // loss current
        }
        block_cumtrapz<Q>(a, dt, g0, sm);  // a = J (without loss term)
#pragma unroll
        for (int q = 0; q < Q; q++) a[q] = a[q] + jf[q];
        block_cumtrapz<Q>(a, dt, g0, sm);  // a = P
    }
    const bool kerr = (in.kerr == LH_KERR_FIELD), plasma = in.plasma;
    const double kfac = in.kerr_fac, rho = in.density;
    const double* tw = in.to_win;
#pragma unroll
    for (int s = 0; s < S; s++) {
        double2 v = sm.buf[pad(S * tid + s)];
        double E0 = (inv_nto * v.x) / fdiv, E1 = (inv_nto * v.y) / fdiv;
        double p0 = kerr ? kfac * (E0 * E0 * E0) : 0.0;
        double p1 = kerr ? kfac * (E1 * E1 * E1) : 0.0;
        if (plasma) {
            p0 = p0 + rho * a[2 * s];
            p1 = p1 + rho * a[2 * s + 1];
        }
        sm.buf[pad(S * tid + s)] = make_double2(p0 * tw[g0 + 2 * s], p1 * tw[g0 + 2 * s + 1]);
    }
}

// Field-grid mode-averaged transform on LDS data.  On entry buf[0..N) holds
// sc_t·Eω (zero to MO).  Calls epi(k, nl_k) for every output bin k < N.
template <int MO, class Epi>
__device__ __forceinline__ void rhs_field(const DInst& in, Smem& sm, const double2* __restrict__ tw, Epi epi) {
    __syncthreads();
    c2r_prepack<MO>(sm.buf, tw);
    __syncthreads();
    fft_lds<MO, true>(sm.buf, tw);
    pointwise_field<MO>(in, sm);
    __syncthreads();
    fft_lds<MO, false>(sm.buf, tw);
    for (int k = threadIdx.x; k < in.n_omega; k += LH_T) {
        double2 nl = cscal(in.sc_f, r2c_bin<MO>(sm.buf, tw, k));
        if (in.sidx[k]) {
            nl = cmul(nl, in.norm[k]);
            nl = cscal(in.omega_win[k], nl);
        }
        epi(k, nl);
    }
}

// Stage prologue: buf[k] = sc_t · (y + Σ_j Δz B_ij k_j)·exp(L hf); k1 first
// propagated by h0 (evaluate!, src/RK45.jl:209-223).
template <int MO>
__device__ __forceinline__ void stage_prologue(const DInst& in, Smem& sm, const double2* Y, double2* const* K,
                                               int stage, double dt, double h0, double hf) {
    const int N = in.n_omega;
    double c[6];
#pragma unroll
    for (int j = 0; j < 6; j++) c[j] = (j <= stage) ? dt * c_B[stage][j] : 0.0;
    for (int k = threadIdx.x; k < N; k += LH_T) {
        double2 L = in.linop[k];
        double2 yb = Y[k];
        if (stage == 0 && h0 != 0.0) {
            double2 k1 = cmul(K[0][k], cexp_lh(L, h0));
            K[0][k] = k1;
        }
#pragma unroll
        for (int j = 0; j < 6; j++)
            if (j <= stage && c_B[stage][j] != 0.0) yb = cadd(yb, cscal(c[j], K[j][k]));
        double2 yp = (hf != 0.0) ? cmul(yb, cexp_lh(L, hf)) : yb;
        sm.buf[pad(k)] = cscal(in.sc_t, yp);
    }
    for (int k = N + threadIdx.x; k <= MO; k += LH_T) sm.buf[pad(k)] = make_double2(0.0, 0.0);
}

// Accepted-step window filter (stepfun, src/Luna.jl:366-370) on the coarse
// grid: YN ← rfft(twin · irfft(ωwin · YN·exp(L h))).
template <int MC>
__device__ __forceinline__ void stepfun_filter(const DInst& in, Smem& sm, const double2* __restrict__ tw, double2* YN,
                                               double h) {
    const int N = in.n_omega;  // == MC + 1
    for (int k = threadIdx.x; k < N; k += LH_T) {
        double2 v = YN[k];
        if (h != 0.0) v = cmul(v, cexp_lh(in.linop[k], h));
        sm.buf[pad(k)] = cscal(in.omega_win[k], v);
    }
    __syncthreads();
    c2r_prepack<MC>(sm.buf, tw);
    __syncthreads();
    fft_lds<MC, true>(sm.buf, tw);
    for (int j = threadIdx.x; j < MC; j += LH_T) {
        double2 z = sm.buf[pad(j)];
        double x0 = (in.inv_nt * z.x) * in.t_win[2 * j];
        double x1 = (in.inv_nt * z.y) * in.t_win[2 * j + 1];
        sm.buf[pad(j)] = make_double2(x0, x1);
    }
    __syncthreads();
    fft_lds<MC, false>(sm.buf, tw);
    for (int k = threadIdx.x; k < N; k += LH_T) YN[k] = r2c_bin<MC>(sm.buf, tw, k);
    __syncthreads();
}

// --------------------------------------------------------------------------
// Persistent adaptive solve of one instance (Luna.run → RK45.solve_precon).
// --------------------------------------------------------------------------
template <int MO, int MC>
__device__ void solve_instance(const DInst& in, Smem& sm, const double2* __restrict__ tw) {
    const int N = in.n_omega;
    double2* K[7];
#pragma unroll
    for (int j = 0; j < 7; j++) K[j] = in.K[j];
    double2* const Y0 = in.Y;
    double2* Y = in.Y;
    double2* YN = in.YN;
    if (threadIdx.x == 0) sm.flag = LH_OK;

    long long steps = 0, accepted = 0, rejected = 0, rhs_evals = 0;
    int status = LH_OK, saved = 0, repeated = 0;
    double t = in.z0, tn = in.z0, dt = in.init_dz, dtn = in.init_dz, errlast = 0.0;

    // PreconStepper init: k1 = fbar(y0, t, t)
    stage_prologue<MO>(in, sm, Y, K, 0, 0.0, 0.0, 0.0);
    {
        double2* k0 = K[0];
        rhs_field<MO>(in, sm, tw, [&](int k, double2 nl) { k0[k] = nl; });
    }
    rhs_evals++;
    __syncthreads();
    if (sm.flag != LH_OK) status = sm.flag;

    while (status == LH_OK && tn <= in.zmax) {
        if (steps >= LH_MAX_STEPS) {
            status = LH_E_MAXSTEPS_INTERNAL;
            break;
        }
        // evaluate!
        double h0 = tn - t;
        dt = dtn;
        t = tn;
        for (int st = 0; st < 6 && status == LH_OK; st++) {
            double t2 = t + c_nodes[st] * dt;
            double hf = t2 - t, hb = t - t2;
            stage_prologue<MO>(in, sm, Y, K, st, dt, h0, hf);
            double2* kout = K[st + 1];
            const double2* Lp = in.linop;
            rhs_field<MO>(in, sm, tw, [&](int k, double2 nl) { kout[k] = cmul(nl, cexp_lh(Lp[k], hb)); });
            rhs_evals++;
            __syncthreads();
            if (sm.flag != LH_OK) status = sm.flag;
        }
        if (status != LH_OK) break;
        steps++;
        // yn = y + Δz Σ b5 k ; yerr = Σ (Δz k) e ; weaknorm (src/RK45.jl:170-181, 343-354)
        double sy = 0.0, syn = 0.0, se = 0.0;
        double cb[7], ce[7];
#pragma unroll
        for (int j = 0; j < 7; j++) {
            cb[j] = dt * c_b5[j];
            ce[j] = c_b5[j] - c_b4[j];
        }
        for (int k = threadIdx.x; k < N; k += LH_T) {
            double2 yv = Y[k];
            double2 yn = yv, ye = make_double2(0.0, 0.0);
#pragma unroll
            for (int j = 0; j < 7; j++) {
                if (c_b5[j] == 0.0) continue;
                double2 kv = K[j][k];
                yn = cadd(yn, cscal(cb[j], kv));
                ye = cadd(ye, cscal(ce[j], cscal(dt, kv)));
            }
            sy += yv.x * yv.x + yv.y * yv.y;
            syn += yn.x * yn.x + yn.y * yn.y;
            se += ye.x * ye.x + ye.y * ye.y;
            YN[k] = yn;
        }
        block_sum3(sy, syn, se, sm.wsum);
        double errwt = fmax(fmax(sqrt(sy), sqrt(syn)), in.atol);
        double err = sqrt(se) / in.rtol / errwt;
        bool ok = err <= 1.0;
        // stepcontrolPI!, src/RK45.jl:374-396, steplims! :399-406
        const double beta1 = 3.0 / 5 / 5, beta2 = -1.0 / 5 / 5, eps = 0.8;
        if (ok) {
            if (errlast == 0.0) errlast = err;
            double fac = (err == 0.0) ? 1.5 : in.safety * pow(eps / err, beta1) * pow(eps / errlast, beta2);
            dtn = fac * dt;
            errlast = err;
        } else {
            if (!isfinite(err))
                dtn = dt / 2;
            else
                dtn = dt * fmax(0.1, in.safety * pow(err, -1.0 / 5));
        }
        if (dtn > in.max_dz)
            dtn = in.max_dz;
        else if (dtn < in.min_dz) {
            dtn = in.min_dz;
            ok = true;
        }
        if (ok) {
            tn = t + dt;
            double2* tmp = K[0];
            K[0] = K[6];
            K[6] = tmp;
            stepfun_filter<MC>(in, sm, tw, YN, tn - t);
            // GridCondition saves via dense output (quirk: ks[1] holds k7)
            while (saved < in.save_n && in.zgrid[saved] <= tn) {
                double ti = in.zgrid[saved];
                if (in.saves) {
                    double2* out = in.saves + (size_t)saved * N;
                    if (ti == t) {
                        for (int k = threadIdx.x; k < N; k += LH_T) out[k] = Y[k];
                    } else if (ti == tn) {
                        for (int k = threadIdx.x; k < N; k += LH_T) out[k] = YN[k];
                    } else {
                        double sg = (ti - t) / dt;
                        double sp[4] = {sg, pow(sg, 2.0), pow(sg, 3.0), pow(sg, 4.0)};
                        double b[7];
#pragma unroll
                        for (int j = 0; j < 7; j++) {
                            double acc = sp[0] * c_iC[0][j];
                            for (int p = 1; p < 4; p++) acc = acc + sp[p] * c_iC[p][j];
                            b[j] = acc;
                        }
                        double h = ti - t;
                        for (int k = threadIdx.x; k < N; k += LH_T) {
                            double2 yi = make_double2(0.0, 0.0);
#pragma unroll
                            for (int j = 0; j < 7; j++) {
                                if (j == 1) continue;
                                const double2* kj = (j == 6) ? K[0] : K[j];
                                yi = cadd(yi, cscal(b[j], kj[k]));
                            }
                            double2 o = cadd(Y[k], cscal(dt, yi));
                            out[k] = cmul(o, cexp_lh(in.linop[k], h));
                        }
                    }
                }
                saved++;
            }
            double2* tmp2 = Y;
            Y = YN;
            YN = tmp2;
            accepted++;
            repeated = 0;
        } else {
            rejected++;
            repeated++;
            if (repeated > in.repeat_limit) status = LH_E_REPEAT_LIMIT;
        }
        __syncthreads();
    }
    // The final state s.yn lives in whichever buffer in.Y points to after the
    // swaps; copy it into the caller-visible buffer Y0.
    if (Y != Y0)
        for (int k = threadIdx.x; k < N; k += LH_T) Y0[k] = Y[k];
    if (threadIdx.x == 0 && in.stats) {
        lh_stats s;
        s.steps = steps;
        s.accepted = accepted;
        s.rejected = rejected;
        s.rhs_evals = rhs_evals;
        s.status = status;
        s.saved = saved;
        s.z = tn;
        s.dz = dtn;
        *in.stats = s;
    }
    __syncthreads();
}

// Dispatch on the instance's grid sizes.
__global__ __launch_bounds__(LH_T) void lh_solve_kernel(const DInst* insts, const double2* __restrict__ tw) {
    __shared__ Smem sm;
    const DInst& in = insts[blockIdx.x];
    if (in.grid_kind == LH_GRID_REAL && in.transform == LH_TRANS_MODEAVG) {
        if (in.n_to == 16384 && in.n_t == 8192)
            solve_instance<8192, 4096>(in, sm, tw);
        else if (in.n_to == 8192 && in.n_t == 4096)
            solve_instance<4096, 2048>(in, sm, tw);
        else if (in.n_to == 4096 && in.n_t == 2048)
            solve_instance<2048, 1024>(in, sm, tw);
        else
            goto unsupported;
        return;
    }
unsupported:
    if (threadIdx.x == 0 && in.stats) {
        lh_stats s;
        memset(&s, 0, sizeof(s));
        s.status = LH_E_ARG;
        *in.stats = s;
    }
}

__global__ __launch_bounds__(LH_T) void lh_rhs_kernel(const DInst* insts, const double2* __restrict__ tw,
                                                      int* status) {
    __shared__ Smem sm;
    const DInst& in = insts[blockIdx.x];
    if (threadIdx.x == 0) sm.flag = LH_OK;
    const int N = in.n_omega;
    for (int k = threadIdx.x; k < N; k += LH_T) sm.buf[pad(k)] = cscal(in.sc_t, in.rhs_in[k]);
    double2* out = in.rhs_out;
    if (in.grid_kind == LH_GRID_REAL && in.transform == LH_TRANS_MODEAVG) {
        if (in.n_to == 16384) {
            for (int k = N + threadIdx.x; k <= 8192; k += LH_T) sm.buf[pad(k)] = make_double2(0, 0);
            rhs_field<8192>(in, sm, tw, [&](int k, double2 nl) { out[k] = nl; });
        } else if (in.n_to == 8192) {
            for (int k = N + threadIdx.x; k <= 4096; k += LH_T) sm.buf[pad(k)] = make_double2(0, 0);
            rhs_field<4096>(in, sm, tw, [&](int k, double2 nl) { out[k] = nl; });
        } else if (in.n_to == 4096) {
            for (int k = N + threadIdx.x; k <= 2048; k += LH_T) sm.buf[pad(k)] = make_double2(0, 0);
            rhs_field<2048>(in, sm, tw, [&](int k, double2 nl) { out[k] = nl; });
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) status[blockIdx.x] = sm.flag;
}


// ==========================================================================
// Host side: contexts, constant upload, launches (C ABI, include/luna_hip.h)
// ==========================================================================
static const double EPS0 = 8.854187817620389e-12;  // CODATA 2014 ε0
static const double ECHARGE = 1.6021766208e-19;
static const double MELEC = 9.10938356e-31;

struct lh_ctx {
    int device;
    hipStream_t stream;
    lh_desc d;  // scalars only; host pointers cleared after upload
    int n_state;
    double ppt_emin, ppt_emax;
    double* omega_win;
    uint8_t* sidx;
    double* t_win;
    double* to_win;
    double2* linop;
    double2* norm;
    double4* ppt;
    // solver buffers (allocated on first use)
    double2* Y;
    double2* YN;
    double2* K[7];
    double2* saves;
    int saves_cap;
    double* zgrid;
    int zgrid_cap;
    lh_stats* dstats;
    DInst* dinst;
    double2* rhs_in;
    double2* rhs_out;
    int* dstatus;
    int last_status;
};

static std::mutex g_mu;
static double2* g_tw[64] = {nullptr};

#define HIPCK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            fprintf(stderr, "luna_hip: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return (e_ == hipErrorOutOfMemory) ? LH_E_NOMEM : LH_E_HIP;              \
        }                                                                            \
    } while (0)

// cos and sin of 2π m/N built from the first octant, so values on the axes are
// exact (0, ±1) and the table is symmetric to the last bit.
static void sincos_frac(long m, long N, double* c, double* s) {
    const long double tau = 6.283185307179586476925286766559005768L;
    m %= N;
    if (m < 0) m += N;
    long oct = (8 * m) / N;
    long f = m - oct * (N / 8);
    if (f == 0 && oct % 2 == 0) {
        const double ax[4][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
        *c = ax[oct / 2][0];
        *s = ax[oct / 2][1];
        return;
    }
    long double phi = tau * (long double)f / (long double)N;
    long double phip = tau * (long double)(N / 8 - f) / (long double)N;
    long double cp = cosl(phi), sp = sinl(phi), cq = cosl(phip), sq = sinl(phip);
    long double C, S;
    switch (oct) {
        case 0: C = cp; S = sp; break;
        case 1: C = sq; S = cq; break;
        case 2: C = -sp; S = cp; break;
        case 3: C = -cq; S = sq; break;
        case 4: C = -cp; S = -sp; break;
        case 5: C = -sq; S = -cq; break;
        case 6: C = sp; S = -cp; break;
        default: C = cq; S = -sq; break;
    }
    *c = (double)C;
    *s = (double)S;
}

// Global twiddle table tw[m] = exp(-2πi m/LH_TWN), one per device.
static int twiddles(int dev, const double2** out) {
    std::lock_guard<std::mutex> g(g_mu);
    if (dev < 0 || dev >= 64) return LH_E_ARG;
    if (!g_tw[dev]) {
        std::vector<double2> h(LH_TWN);
        for (int m = 0; m < LH_TWN; m++) {
            double c, s;
            sincos_frac(m, LH_TWN, &c, &s);
            h[m] = make_double2(c, -s);
        }
        double2* d = nullptr;
        HIPCK(hipMalloc(&d, sizeof(double2) * LH_TWN));
        HIPCK(hipMemcpy(d, h.data(), sizeof(double2) * LH_TWN, hipMemcpyHostToDevice));
        g_tw[dev] = d;
    }
    *out = g_tw[dev];
    return LH_OK;
}

template <class T>
static int upload(T** dst, const void* src, size_t count) {
    *dst = nullptr;
    if (!src || count == 0) return LH_OK;
    HIPCK(hipMalloc((void**)dst, sizeof(T) * count));
    HIPCK(hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
    return LH_OK;
}

static bool supported(const lh_desc* d) {
    if (d->grid_kind != LH_GRID_REAL || d->transform != LH_TRANS_MODEAVG || d->n_r != 1) return false;
    if (!(d->n_to == 16384 || d->n_to == 8192 || d->n_to == 4096)) return false;
    if (d->n_t * 2 != d->n_to) return false;
    if (d->n_omega != d->n_t / 2 + 1) return false;
    if (d->raman || d->kerr == LH_KERR_ENV) return false;
    return true;
}

static DInst base_inst(const lh_ctx* c) {
    DInst in;
    memset(&in, 0, sizeof(in));
    const lh_desc& d = c->d;
    in.grid_kind = d.grid_kind;
    in.transform = d.transform;
    in.n_omega = d.n_omega;
    in.n_t = d.n_t;
    in.n_to = d.n_to;
    in.n_r = d.n_r;
    in.kerr = d.kerr;
    in.plasma = d.plasma;
    in.raman = d.raman;
    in.ppt_n = d.ppt_n;
    in.field_div = d.field_div;
    in.density = d.density;
    in.kerr_fac = d.density * EPS0 * d.gamma3;  // (ρ ε0) γ3, Nonlinear.jl:30
    in.ionpot = d.ionpot;
    in.preion_p1 = d.preionfrac + 1;
    in.e_ratio = ECHARGE * ECHARGE / MELEC;
    in.dt_o = d.dt_o;
    in.inv_nto = 1.0 / (double)d.n_to;
    in.inv_nt = 1.0 / (double)d.n_t;
    if (d.grid_kind == LH_GRID_REAL) {
        int No = d.n_to / 2 + 1;
        in.sc_t = (double)(No - 1) / (double)(d.n_omega - 1);
        in.sc_f = (double)(d.n_omega - 1) / (double)(No - 1);
    } else {
        in.sc_t = (double)d.n_to / (double)d.n_omega;
        in.sc_f = (double)d.n_omega / (double)d.n_to;
    }
    in.ppt_emin = c->ppt_emin;
    in.ppt_emax = c->ppt_emax;
    in.omega_win = c->omega_win;
    in.sidx = c->sidx;
    in.t_win = c->t_win;
    in.to_win = c->to_win;
    in.linop = c->linop;
    in.norm = c->norm;
    in.ppt = c->ppt;
    return in;
}

// Per-context solver buffers sized for n_state and save_n.
static int ensure_solver_buffers(lh_ctx* c, int save_n, bool want_saves) {
    const size_t NS = (size_t)c->n_state;
    if (!c->Y) {
        HIPCK(hipMalloc(&c->Y, sizeof(double2) * NS));
        HIPCK(hipMalloc(&c->YN, sizeof(double2) * NS));
        for (int j = 0; j < 7; j++) HIPCK(hipMalloc(&c->K[j], sizeof(double2) * NS));
    }
    if (want_saves && save_n > c->saves_cap) {
        if (c->saves) HIPCK(hipFree(c->saves));
        c->saves = nullptr;
        HIPCK(hipMalloc(&c->saves, sizeof(double2) * NS * (size_t)save_n));
        c->saves_cap = save_n;
    }
    int zn = save_n > 0 ? save_n : 1;
    if (zn > c->zgrid_cap) {
        if (c->zgrid) HIPCK(hipFree(c->zgrid));
        c->zgrid = nullptr;
        HIPCK(hipMalloc(&c->zgrid, sizeof(double) * (size_t)zn));
        c->zgrid_cap = zn;
    }
    return LH_OK;
}

// range(z0, stop=zmax, length=n) as Julia evaluates it (correctly rounded lerp).
static void save_grid(const lh_rk_opts* o, std::vector<double>& g) {
    int n = o->save_n;
    g.assign(n > 0 ? n : 1, o->z0);
    long double a = o->z0, b = o->zmax;
    for (int i = 0; i < n; i++)
        g[i] = (n == 1) ? o->z0 : (double)((a * (long double)(n - 1 - i) + b * (long double)i) / (long double)(n - 1));
}

static int prepare_instance(lh_ctx* c, const double* Ew, const lh_rk_opts* o, bool want_saves, DInst* in,
                            lh_stats* dstats_slot) {
    if (o->save_n < 0 || o->repeat_limit < 0 || !(o->rtol > 0) || !(o->zmax >= o->z0)) return LH_E_ARG;
    int st = ensure_solver_buffers(c, o->save_n, want_saves);
    if (st) return st;
    std::vector<double> g;
    save_grid(o, g);
    HIPCK(hipMemcpyAsync(c->zgrid, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->Y, Ew, sizeof(double2) * (size_t)c->n_state, hipMemcpyHostToDevice, c->stream));
    *in = base_inst(c);
    in->Y = c->Y;
    in->YN = c->YN;
    for (int j = 0; j < 7; j++) in->K[j] = c->K[j];
    in->saves = want_saves ? c->saves : nullptr;
    in->zgrid = c->zgrid;
    in->stats = dstats_slot;
    in->z0 = o->z0;
    in->zmax = o->zmax;
    in->rtol = o->rtol;
    in->atol = o->atol;
    in->safety = o->safety;
    in->max_dz = o->max_dz;
    in->min_dz = o->min_dz;
    in->init_dz = o->init_dz;
    in->repeat_limit = o->repeat_limit;
    in->save_n = o->save_n;
    return LH_OK;
}

static int fetch_instance(lh_ctx* c, double* Ew, const lh_rk_opts* o, double* saves, double* zsaves,
                          hipStream_t s) {
    HIPCK(hipMemcpyAsync(Ew, c->Y, sizeof(double2) * (size_t)c->n_state, hipMemcpyDeviceToHost, s));
    if (saves && o->save_n > 0)
        HIPCK(hipMemcpyAsync(saves, c->saves, sizeof(double2) * (size_t)c->n_state * o->save_n,
                             hipMemcpyDeviceToHost, s));
    if (zsaves && o->save_n > 0)
        HIPCK(hipMemcpyAsync(zsaves, c->zgrid, sizeof(double) * o->save_n, hipMemcpyDeviceToHost, s));
    return LH_OK;
}

extern "C" {

const char* lh_strerror(int status) {
    switch (status) {
        case LH_OK: return "ok";
        case LH_E_ARG: return "invalid argument or unsupported configuration";
        case LH_E_HIP: return "HIP runtime error";
        case LH_E_PPT_RANGE: return "field strength exceeds maximum for PPT ionisation rate";
        case LH_E_REPEAT_LIMIT: return "reached limit for step repetition";
        case LH_E_NONFINITE: return "non-finite state";
        case LH_E_NOMEM: return "device allocation failed";
        case LH_E_NODEVICE: return "no HIP device";
        default: return "unknown status";
    }
}

int lh_device_count(int* count) {
    if (!count) return LH_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return LH_OK;
}

int lh_destroy(lh_ctx* c) {
    if (!c) return LH_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->omega_win, c->sidx, c->t_win, c->to_win, c->linop, c->norm, c->ppt, c->Y, c->YN,
                    c->saves, c->zgrid, c->dstats, c->dinst, c->rhs_in, c->rhs_out, c->dstatus};
    for (void* p : ptrs)
        if (p) hipFree(p);
    for (int j = 0; j < 7; j++)
        if (c->K[j]) hipFree(c->K[j]);
    if (c->stream) hipStreamDestroy(c->stream);
    free(c);
    return LH_OK;
}

int lh_create(lh_ctx** out, const lh_desc* d) {
    if (!out || !d) return LH_E_ARG;
    *out = nullptr;
    if (!supported(d)) return LH_E_ARG;
    if (!d->omega_win || !d->sidx || !d->t_win || !d->to_win || !d->norm) return LH_E_ARG;
    if (d->plasma && (!d->ppt_x || !d->ppt_y || !d->ppt_d || d->ppt_n < 4)) return LH_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LH_E_NODEVICE;
    lh_ctx* c = (lh_ctx*)calloc(1, sizeof(lh_ctx));
    if (!c) return LH_E_NOMEM;
    c->d = *d;
    c->n_state = d->n_omega * d->n_r;
    int st = LH_OK;
    if (hipGetDevice(&c->device) != hipSuccess) st = LH_E_HIP;
    if (!st && hipStreamCreate(&c->stream) != hipSuccess) st = LH_E_HIP;
    const size_t N = (size_t)d->n_omega, NS = (size_t)c->n_state;
    if (!st) st = upload(&c->omega_win, d->omega_win, N);
    if (!st) st = upload(&c->sidx, d->sidx, N);
    if (!st) st = upload(&c->t_win, d->t_win, (size_t)d->n_t);
    if (!st) st = upload(&c->to_win, d->to_win, (size_t)d->n_to);
    if (!st) st = upload(&c->norm, d->norm, NS);
    if (!st) {
        std::vector<double2> lin(NS, make_double2(0.0, 0.0));
        if (d->linop) memcpy(lin.data(), d->linop, sizeof(double2) * NS);
        st = upload(&c->linop, lin.data(), NS);
    }
    if (!st && d->plasma) {
        std::vector<double4> p((size_t)d->ppt_n);
        for (int i = 0; i < d->ppt_n; i++) p[i] = make_double4(d->ppt_x[i], d->ppt_y[i], d->ppt_d[i], 0.0);
        c->ppt_emin = d->ppt_x[0];
        c->ppt_emax = d->ppt_x[d->ppt_n - 1];
        st = upload(&c->ppt, p.data(), p.size());
    }
    // host pointers inside the copied descriptor belong to the caller
    c->d.omega = nullptr; c->d.sidx = nullptr; c->d.omega_win = nullptr; c->d.t_win = nullptr;
    c->d.to_win = nullptr; c->d.linop = nullptr; c->d.norm = nullptr; c->d.ppt_x = nullptr;
    c->d.ppt_y = nullptr; c->d.ppt_d = nullptr; c->d.raman_h = nullptr; c->d.qdht = nullptr;
    if (!st) {
        const double2* tw;
        st = twiddles(c->device, &tw);
    }
    if (!st && hipMalloc(&c->dinst, sizeof(DInst)) != hipSuccess) st = LH_E_NOMEM;
    if (!st && hipMalloc(&c->dstats, sizeof(lh_stats)) != hipSuccess) st = LH_E_NOMEM;
    if (!st && hipMalloc(&c->dstatus, sizeof(int)) != hipSuccess) st = LH_E_NOMEM;
    if (st) {
        lh_destroy(c);
        return st;
    }
    *out = c;
    return LH_OK;
}

int lh_ctx_status(lh_ctx* c) { return c ? c->last_status : LH_E_ARG; }

int lh_rhs_device(lh_ctx* c, void* nl_dev, const void* Ew_dev, double z, void* stream) {
    (void)z;  // constant density / Aeff / linear operator: no z dependence
    if (!c || !nl_dev || !Ew_dev) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const double2* tw;
    int st = twiddles(c->device, &tw);
    if (st) return st;
    DInst in = base_inst(c);
    in.rhs_in = (const double2*)Ew_dev;
    in.rhs_out = (double2*)nl_dev;
    HIPCK(hipMemcpyAsync(c->dinst, &in, sizeof(DInst), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(lh_rhs_kernel, dim3(1), dim3(LH_T), 0, s, c->dinst, tw, c->dstatus);
    HIPCK(hipGetLastError());
    return LH_OK;
}

int lh_rhs(lh_ctx* c, double* nl, const double* Ew, double z) {
    if (!c || !nl || !Ew) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    const size_t bytes = sizeof(double2) * (size_t)c->n_state;
    if (!c->rhs_in) {
        HIPCK(hipMalloc(&c->rhs_in, bytes));
        HIPCK(hipMalloc(&c->rhs_out, bytes));
    }
    HIPCK(hipMemcpyAsync(c->rhs_in, Ew, bytes, hipMemcpyHostToDevice, c->stream));
    int st = lh_rhs_device(c, c->rhs_out, c->rhs_in, z, c->stream);
    if (st) return st;
    int kst = LH_OK;
    HIPCK(hipMemcpyAsync(nl, c->rhs_out, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(&kst, c->dstatus, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    c->last_status = kst;
    return kst;
}

int lh_solve(lh_ctx* c, double* Ew, const lh_rk_opts* o, double* saves, double* zsaves, lh_stats* stats) {
    lh_ctx* cs[1] = {c};
    double* ews[1] = {Ew};
    double* svs[1] = {saves};
    double* zss[1] = {zsaves};
    lh_stats tmp;
    lh_stats* sp = stats ? stats : &tmp;
    int st = lh_batch_solve(cs, 1, ews, o, saves ? svs : nullptr, zss, sp);
    if (st) return st;
    return sp->status;
}

int lh_batch_solve(lh_ctx* const* ctxs, int n, double* const* Ew, const lh_rk_opts* o, double* const* saves,
                   double* const* zsaves, lh_stats* stats) {
    if (!ctxs || n <= 0 || !Ew || !o || !stats) return LH_E_ARG;
    const int dev = ctxs[0]->device;
    for (int i = 0; i < n; i++)
        if (!ctxs[i] || ctxs[i]->device != dev || !Ew[i]) return LH_E_ARG;
    for (int i = 0; i < n; i++)  // the same context twice would share buffers
        for (int j = i + 1; j < n; j++)
            if (ctxs[i] == ctxs[j]) return LH_E_ARG;
    HIPCK(hipSetDevice(dev));
    hipStream_t s = ctxs[0]->stream;
    const double2* tw;
    int st = twiddles(dev, &tw);
    if (st) return st;
    DInst* dinst = nullptr;
    lh_stats* dstats = nullptr;
    HIPCK(hipMalloc(&dinst, sizeof(DInst) * (size_t)n));
    HIPCK(hipMalloc(&dstats, sizeof(lh_stats) * (size_t)n));
    std::vector<DInst> h(n);
    for (int i = 0; i < n && !st; i++) {
        lh_ctx* c = ctxs[i];
        bool ws = saves && saves[i];
        // uploads of per-instance data are ordered on the launch stream
        hipStream_t keep = c->stream;
        c->stream = s;
        st = prepare_instance(c, Ew[i], o, ws, &h[i], dstats + i);
        c->stream = keep;
    }
    if (!st && hipMemcpyAsync(dinst, h.data(), sizeof(DInst) * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess)
        st = LH_E_HIP;
    if (!st) {
        hipLaunchKernelGGL(lh_solve_kernel, dim3(n), dim3(LH_T), 0, s, dinst, tw);
        if (hipGetLastError() != hipSuccess) st = LH_E_HIP;
    }
    if (!st && hipMemcpyAsync(stats, dstats, sizeof(lh_stats) * (size_t)n, hipMemcpyDeviceToHost, s) != hipSuccess)
        st = LH_E_HIP;
    for (int i = 0; i < n && !st; i++)
        st = fetch_instance(ctxs[i], Ew[i], o, saves ? saves[i] : nullptr, zsaves ? zsaves[i] : nullptr, s);
    if (hipStreamSynchronize(s) != hipSuccess && !st) st = LH_E_HIP;
    hipFree(dinst);
    hipFree(dstats);
    for (int i = 0; i < n; i++) ctxs[i]->last_status = stats[i].status;
    return st;
}

}  // extern "C"
