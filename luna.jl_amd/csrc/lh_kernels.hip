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
#include <string>
#include <time.h>
#include <vector>

#include "../../include/luna_hip.h"
#include "lh_device.h"

static const long long LH_MAX_STEPS = 500000;  // default cap on step attempts per instance

// --------------------------------------------------------------------------
// Device-side instance description
// --------------------------------------------------------------------------
struct DInst {
    int grid_kind, transform, n_omega, n_t, n_to, n_r;
    int kerr, plasma, raman, ppt_n;
    double density, kerr_fac, ionpot, preion_p1, e_ratio;
    double dt_o, e_scale, inv_nt, sc_t, ppt_emin, ppt_emax;
    const double* omega_win;
    const double* t_win;
    const double* to_win;
    const double2* linop;
    const double2* outfac;  // per-bin output factor: sc_f·norm·ωwin (sidx) or sc_f
    const double4* ppt;     // (x, y, D, 0) per spline node
    // envelope grid
    double kerr34;          // ¾·ρ·ε0·γ3 (Kerr_env)
    double inv_2n;          // 1/(2·Nto): inverse of the zero-padded Raman transform
    const double2* hwdt;    // rfft of h(t) zero-padded to 2·Nto, times δt: bins 0..Nto
    double2* escr;          // HBM stash of E(t) during the Raman convolution
    // solver state
    const double2* Yin;  // initial field (never written)
    double2* Y;          // receives the final state s.yn
    double2* YN;
    double2* K[7];        // one block: K[j] = K[0] + j·kstride
    long long kstride;
    double2* saves;       // [save_n][n_state] or null
    const double* zgrid;  // [save_n]
    lh_stats* stats;
    double z0, zmax, rtol, atol, safety, max_dz, min_dz, init_dz;
    int repeat_limit, save_n;
    int save_first;       // saves already written (a resumed run continues at this index)
    double zstop;         // > 0: stop after the first accepted step with tn > zstop
    long long max_steps;  // cap on step attempts
    // single-RHS I/O
    const double2* rhs_in;
    double2* rhs_out;
    // per-step statistics (steprec null: off)
    const double* st_omega;  // [Nω] ω
    const double* st_ew;     // [Nω] energy quadrature weights
    const double* st_t;      // [Nt] grid.t
    double st_aeff;          // Aeff
    double st_ediv;          // sqrt(ε0·c·Aeff/2)
    const double* st_win;    // [st_nwin][Nω] energy_w·window² of each energy window
    int st_nwin;
    double* steprec;         // [steprec_cap][LH_NSTAT]
    int steprec_cap;
    // z-dependent fill (pressure gradient); zg_np = 0: constant fill
    int zg_np, zg_two, zg_model, zg_loss, zd_n;
    double sc_f;             // to_freq! crop scale (Nω-1)/(No-1)
    double kerr_gamma3;      // γ3 of Kerr_field (0 without Kerr)
    double zsqrt_aeff, zfd_h;
    const double* zg_Z;      // [zg_np] profile positions
    const double* zg_P;      // [zg_np] profile pressures
    const double4* zd;       // [zd_n] density spline (P, ρ, D, 0)
    const double* zgam;      // [Nω] γ(λ)
    const double2* znwg;     // [Nω] neff_wg
    const double2* zpre;     // [Nω] norm prefactor
    const unsigned char* zsidx;  // [Nω]
    const double* zfd;       // [9][4] β1 stencil (ω, γ, nwg)
    int ztaper;              // 1: linear taper a(z), 2: piecewise linear; neff_wg from zU, zV
    double za0, za1, zaL, zsqrt_intg;
    int zt_n;                // taper = 2: nodes
    const double* zt_Z;      // [zt_n] positions
    const double* zt_A;      // [zt_n] radii
    const double* zU;        // [Nω] u_nm/k
    const double2* zV;       // [Nω] v_n/k
    const double* zfdt;      // [9][3] U, V at the β1 stencil points
    double2* zL;             // [Nω] linop at the last refresh (= linop)
    double2* zOF;            // [Nω] output factor at the last refresh (= outfac)
    // streaming output (lh_solve_stream; pub null: off): device-mapped,
    // page-locked host memory.  A step that writes saves publishes itself:
    // its filtered state and (z, dz, saved, accepted) go to slot k & 1 of
    // publication k, then pub is set to k + 1 with a system-scope release.
    unsigned long long* pub;
    double* pub_info;        // [2][4]
    double2* pub_y;          // [2][n_state]
    const int* pub_stop;     // set by the host: end with LH_E_STOPPED at the next publication
    // every-step mode (pub_ring > 0): every accepted step publishes, into
    // slot k % pub_ring of pub_ring slots (pub_info and pub_y hold that many);
    // the kernel never waits, the host checks that it has not been lapped
    int pub_ring;
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

// LDS of one workgroup: the padded transform buffer sized for the largest
// transform it runs (MO complex points + Nyquist), twiddles, scan scratch.
// plist: the rates of the samples that need one (the few hundred under the
// pulse), so that every thread takes one or two of them instead of the
// threads owning the pulse doing all of that work: a compacted list up to
// Nto = 8192, a window [lo, hi] of the round for Nto = 16384, whose 149 KB of
// other buffers leave 11 KB (1408 samples) for it.
template <int MO>
struct PList {
    // what the largest class's LDS leaves after the buffer, twiddles and the
    // per-wave reduction and scan slots
    static constexpr int CAP = (MO >= 8192) ? 1344 - 19 * (LH_NW - 8) : 4096;
};

template <int MO>
struct SmemT {
    double2 buf[padded_len(MO)];
    TwLds tw;
    double wsum[5 * LH_NW];      // block_sum3/block_sum4 and the statistics reductions
    double scan[2][7 * LH_NW];   // one-barrier scans and min/max, used in alternation
    double plist[PList<MO>::CAP];
    double zs[3];  // gradient fill: ρ(z), β1(z), taper radius a(z) from thread 0
    int flag;
};

// Number of state bins each thread owns for an oversampled half-length MO.
template <int MO>
struct Bins {
    static constexpr int N = MO / 2 + 1;  // Nω of a real grid with Nto = 2·MO, Nt = MO
    static constexpr int NPT = (N + LH_T - 1) / LH_T;
};

// --------------------------------------------------------------------------
// Maths.CSpline evaluation on uniform nodes (x, y, D) with the uniform index
// rule (src/Maths.jl:771-775, 840-848).  The PPT table (log rate) and the
// density-vs-pressure spline of a gradient fill use it.
// --------------------------------------------------------------------------
__device__ __forceinline__ double cspline_uniform(const double4* nodes, int n, double xmin, double xmax, double x) {
    int i;
    if (x <= xmin)
        i = 2;
    else if (x >= xmax)
        i = n;
    else
        i = (int)ceil((x - xmin) / (xmax - xmin) * (double)(n - 1)) + 1;
    i = min(max(i, 2), n);
    const double4 pc = nodes[i - 1], pm = nodes[i - 2];
    if (x == pc.x) return pc.y;
    if (x == pm.x) return pm.y;
    const double t = (x - pm.x) / (pc.x - pm.x);
    return (pm.y + pm.z * t + (3 * (pc.y - pm.y) - 2 * pm.z - pc.z) * (t * t) +
            (2 * (pm.y - pc.y) + pm.z + pc.z) * (t * t * t));
}

// The same for Q samples at once: every index first, then all table loads,
// then the splines, so the loads of a thread's samples are in flight together
// instead of forming one dependent load-and-wait chain per sample.  Inactive
// samples (below the table, or out of range) load the first interval and are
// discarded.
template <int Q>
__device__ __forceinline__ void ppt_rates(const DInst& in, const double* aE, double* rate, int& bad) {
    const double xmin = in.ppt_emin, xmax = in.ppt_emax;
    const int n = in.ppt_n;
    int idx[Q];
    bool act[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const double x = aE[q];
        const bool below = x < xmin;
        const bool out = !below && !(x <= xmax);  // above the table, or NaN
        if (out) bad = (x > xmax) ? LH_E_PPT_RANGE : LH_E_NONFINITE;
        act[q] = !below && !out;
        int i = 2;
        if (act[q]) {
            if (x <= xmin)
                i = 2;
            else if (x >= xmax)
                i = n;
            else
                i = (int)ceil((x - xmin) / (xmax - xmin) * (double)(n - 1)) + 1;
            i = min(max(i, 2), n);
        }
        idx[q] = i;
    }
    double4 pc[Q], pm[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        pc[q] = gl(in.ppt)[idx[q] - 1];
        pm[q] = gl(in.ppt)[idx[q] - 2];
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        double r = 0.0;
        if (act[q]) {
            const double x = aE[q];
            double v;
            if (x == pc[q].x) {
                v = pc[q].y;
            } else if (x == pm[q].x) {
                v = pm[q].y;
            } else {
                const double t = (x - pm[q].x) / (pc[q].x - pm[q].x);
                v = (pm[q].y + pm[q].z * t + (3 * (pc[q].y - pm[q].y) - 2 * pm[q].z - pc[q].z) * (t * t) +
                     (2 * (pm[q].y - pc[q].y) + pm[q].z + pc[q].z) * (t * t * t));
            }
            r = exp(v);
        }
        rate[q] = r;
    }
}

// cumtrapz over the workgroup for one round of QR samples per thread (global
// index g0+q), with a single barrier.  The trapezoid joining a thread's first
// sample to its predecessor's last comes from a lane move inside a
// wavefront; for lane 0 of waves ≥ 1 it is added after the barrier from the
// previous wave's last sample in LDS.  S: a scan slot of 3·LH_NW doubles (wave
// totals, first and last samples of each wave); consecutive calls alternate
// between the two slots, so a slot is rewritten only after a further barrier
// has passed all of its readers.  carry: running integral before this round;
// prev0: the integrand at the sample just before the round (thread 0 only).
// On return y[q] holds the running integral and carry the round's final value;
// returns the round's last integrand value.
template <int QR>
__device__ __forceinline__ double block_cumtrapz(double* y, double dx, int g0, double* S, double& carry,
                                                 double prev0) {
    const int tid = fresh_tid();
    const int lane = tid & 63, wid = tid >> 6;
    const double first = y[0], lastv = y[QR - 1];
    double prev = dpp_f64<DPP_WAVE_SHR1>(lastv);
    if (tid == 0) prev = prev0;
    const bool joined = (g0 != 0) && !(lane == 0 && wid > 0);
    double acc = 0.0;
    double last = prev;
#pragma unroll
    for (int q = 0; q < QR; q++) {
        const double cur = y[q];
        const double term = (q == 0) ? (joined ? (0.5 * (last + cur)) * dx : 0.0) : (0.5 * (last + cur)) * dx;
        last = cur;
        acc += term;
        y[q] = acc;
    }
    const double x = wave_incl_scan(acc);
    const double ex = dpp_f64<DPP_WAVE_SHR1>(x);  // exclusive (0 in lane 0)
    if (lane == 63) {
        S[wid] = x;
        S[2 * LH_NW + wid] = lastv;
    }
    if (lane == 0) S[LH_NW + wid] = first;
    __syncthreads();
    // lane w < LH_NW: wave w's joining trapezoid and total, scanned over those lanes
    double Ww = 0.0, c = 0.0;
    if (lane < LH_NW) {
        Ww = S[lane];
        c = (lane > 0) ? (0.5 * (S[2 * LH_NW + lane - 1] + S[LH_NW + lane])) * dx + Ww : Ww;
    }
    c = lanes_incl_scan(c);
    const double off = (carry + (lane_value(c, wid) - lane_value(Ww, wid))) + ex;
#pragma unroll
    for (int q = 0; q < QR; q++) y[q] = off + y[q];
    carry = carry + lane_value(c, LH_NW - 1);
    return S[3 * LH_NW - 1];
}

// Two chained cumtrapz over the workgroup with a single barrier: J = ∫a and
// P = ∫(J + l) for one round of QR samples per thread (global index g0+q).
// A run of samples acts on the integrals at the sample before it as
// J ← J + SJ, P ← P + n·dx·J + SP (n: its trapezoids; SJ, SP: its integrals
// from J = P = 0), and runs compose as
//   (n1, SJ1, SP1)·(n2, SJ2, SP2) = (n1 + n2, SJ1 + SJ2, SP1 + SP2 + n2·dx·SJ1),
// which the wave scan (DPP) and the cross-wave step apply in that order.  The
// trapezoid joining lane 0 of wave w ≥ 1 to wave w−1's last sample is a run
// of its own, formed after the barrier from the first and last samples in S
// (7·LH_NW doubles: wave totals n, SJ, SP; first a, l; last a, l).
// cJ, cP: J and P at the sample before the round; pa0, pl0: a and l there
// (thread 0).  On return a[q] = P at the sample, cJ, cP, pa0, pl0 those of the
// round's last sample.
struct Run2 {
    int n;
    double sj, sp;
};
__device__ __forceinline__ Run2 run2_cat(const Run2& x, const Run2& y, double dx) {
    return Run2{x.n + y.n, x.sj + y.sj, (x.sp + y.sp) + ((double)y.n * dx) * x.sj};
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ Run2 dpp_run2(const Run2& x) {
    return Run2{dpp_i32<CTRL, ROW_MASK>(x.n, 0), dpp_f64<CTRL, ROW_MASK>(x.sj), dpp_f64<CTRL, ROW_MASK>(x.sp)};
}

template <int QR>
__device__ __forceinline__ void block_cumtrapz2(double* a, const double* l, double dx, int g0, double* S, double& cJ,
                                                double& cP, double& pa0, double& pl0) {
    const int tid = fresh_tid();
    const int lane = tid & 63, wid = tid >> 6;
    const double afirst = a[0], lfirst = l[0], alast = a[QR - 1], llast = l[QR - 1];
    double aprev = dpp_f64<DPP_WAVE_SHR1>(alast), pprev = dpp_f64<DPP_WAVE_SHR1>(llast);
    if (tid == 0) {
        aprev = pa0;
        pprev = pl0;
    }
    const bool joined = (g0 != 0) && !(lane == 0 && wid > 0);
    // this thread's run; a[q] := its SP up to sample q
    double sj = 0.0, sp = 0.0;
#pragma unroll
    for (int q = 0; q < QR; q++) {
        const double cur = a[q];
        if (q > 0 || joined) {
            sj += (0.5 * (aprev + cur)) * dx;
            const double pc = sj + l[q];
            sp += (0.5 * (pprev + pc)) * dx;
            pprev = pc;
        } else {
            pprev = l[q];
        }
        aprev = cur;
        a[q] = sp;
    }
    const int nt = QR - (joined ? 0 : 1);
    // inclusive scan of the runs over the wave (DPP), then exclusive by one lane
    Run2 x{nt, sj, sp};
    x = run2_cat(dpp_run2<DPP_ROW_SHR1>(x), x, dx);
    x = run2_cat(dpp_run2<DPP_ROW_SHR2>(x), x, dx);
    x = run2_cat(dpp_run2<DPP_ROW_SHR4>(x), x, dx);
    x = run2_cat(dpp_run2<DPP_ROW_SHR8>(x), x, dx);
    x = run2_cat(dpp_run2<DPP_ROW_BCAST15, 0xa>(x), x, dx);
    x = run2_cat(dpp_run2<DPP_ROW_BCAST31, 0xc>(x), x, dx);
    const Run2 ex = dpp_run2<DPP_WAVE_SHR1>(x);
    if (lane == 63) {
        S[wid] = (double)x.n;
        S[LH_NW + wid] = x.sj;
        S[2 * LH_NW + wid] = x.sp;
        S[5 * LH_NW + wid] = alast;
        S[6 * LH_NW + wid] = llast;
    }
    if (lane == 0) {
        S[3 * LH_NW + wid] = afirst;
        S[4 * LH_NW + wid] = lfirst;
    }
    __syncthreads();
    // lane w < LH_NW: the join into wave w (w ≥ 1), then wave w's total
    Run2 e{0, 0.0, 0.0}, jn{0, 0.0, 0.0};
    if (lane < LH_NW) {
        if (lane > 0) {
            const double ta = (0.5 * (S[5 * LH_NW + lane - 1] + S[3 * LH_NW + lane])) * dx;
            const double sp1 = (0.5 * (S[6 * LH_NW + lane - 1] + (ta + S[4 * LH_NW + lane]))) * dx;
            jn = Run2{1, ta, sp1};
        }
        e = run2_cat(jn, Run2{(int)S[lane], S[LH_NW + lane], S[2 * LH_NW + lane]}, dx);
    }
    Run2 c = e;  // inclusive over lanes 0..LH_NW-1
    c = run2_cat(dpp_run2<DPP_ROW_SHR1>(c), c, dx);
    c = run2_cat(dpp_run2<DPP_ROW_SHR2>(c), c, dx);
    c = run2_cat(dpp_run2<DPP_ROW_SHR4>(c), c, dx);
    if constexpr (LH_NW > 8) c = run2_cat(dpp_run2<DPP_ROW_SHR8>(c), c, dx);
    const Run2 win = run2_cat(dpp_run2<DPP_ROW_SHR1>(c), jn, dx);  // runs before wave w's first lane
    // runs before this thread: the wave's, then the lanes' before it in the wave
    const Run2 w_in{__builtin_amdgcn_readlane(win.n, __builtin_amdgcn_readfirstlane(wid)), lane_value(win.sj, wid),
                    lane_value(win.sp, wid)};
    const Run2 t_in = run2_cat(w_in, ex, dx);
    const double J = cJ + t_in.sj;
    const double P = (cP + ((double)t_in.n * dx) * cJ) + t_in.sp;
#pragma unroll
    for (int q = 0; q < QR; q++) {
        const int nq = q + (joined ? 1 : 0);
        a[q] = (P + ((double)nq * dx) * J) + a[q];
    }
    const Run2 all{__builtin_amdgcn_readlane(c.n, LH_NW - 1), lane_value(c.sj, LH_NW - 1),
                   lane_value(c.sp, LH_NW - 1)};
    cP = (cP + ((double)all.n * dx) * cJ) + all.sp;
    cJ = cJ + all.sj;
    pa0 = S[6 * LH_NW - 1];
    pl0 = S[7 * LH_NW - 1];
}

// Exclusive prefix sum of one value per thread with a single barrier (scan
// slot S as for block_cumtrapz); also returns the workgroup total.
__device__ __forceinline__ double block_exscan1(double v, double* S, double* total) {
    const int tid = fresh_tid();
    const int lane = tid & 63, wid = tid >> 6;
    const double x = wave_incl_scan(v);
    const double ex = dpp_f64<DPP_WAVE_SHR1>(x);
    if (lane == 63) S[wid] = x;
    __syncthreads();
    const double Ww = (lane < LH_NW) ? S[lane] : 0.0;
    const double c = lanes_incl_scan(Ww);
    *total = lane_value(c, LH_NW - 1);
    return (lane_value(c, wid) - lane_value(Ww, wid)) + ex;
}

// Workgroup min of lo and max of hi with a single barrier (scan slot S).
__device__ __forceinline__ void block_minmax(int& lo, int& hi, double* S) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    wave_minmax(lo, hi);
    if (lane == 0) {
        S[2 * wid] = (double)lo;
        S[2 * wid + 1] = (double)hi;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < LH_NW; w++) {
        lo = min(lo, (int)S[2 * w]);
        hi = max(hi, (int)S[2 * w + 1]);
    }
}

template <class Smem>
__device__ __forceinline__ double field_at(const Smem& sm, int g, double e_scale) {
    double2 v = sm.buf[pad(g >> 1)];
    return ((g & 1) ? v.y : v.x) * e_scale;
}

// The integrals and scan slot pointwise_field carries from round to round
// (kept across the LDS windows of the series staged through HBM).
struct PwCarry {
    double cF, cJ, cP, lastRate, lastPhi, lastLoss, Fprev, fprev;
    int sl;
};

// Pointwise field responses on the packed real series in LDS (in place):
// E = c2r/Nto/field_div; P = ρε0γ3 E³ + ρ·Plasma(E); P *= towin.
// The series is processed in rounds of QR samples per thread (contiguous
// chunks), with the three cumtrapz integrals carried from round to round;
// each thread's samples are touched only by that thread.
// WIN: LDS holds the window of 2·LDS_N samples from sample `base` on (the
// series of the HBM-staged grid); rounds r0 .. r0 + rounds in the window run,
// with the carries taken from and left in *cw.
template <int MO, class Smem, bool RAMAN = false, bool WIN = false, int LDS_N = MO>
__device__ __forceinline__ void pointwise_field(const DInst& in, Smem& sm, double rho, double kfac,
                                                PwCarry* cw = nullptr, int r0 = 0, int base = 0) {
    constexpr int QT = 2 * MO / LH_T;          // samples per thread over all rounds
    constexpr int QR = (QT > 16) ? 16 : QT;    // samples per thread per round
    constexpr int NR = QT / QR;                // rounds
    constexpr int RW = WIN ? (2 * LDS_N) / (LH_T * QR) : NR;  // rounds of this call
    static_assert(!WIN || (2 * LDS_N) % (LH_T * QR) == 0, "whole rounds per LDS window");
    const int tid = fresh_tid();
    const int lb = WIN ? base : 0;             // sample index of LDS sample 0
    const double es = in.e_scale;
    const bool kerr = (in.kerr == LH_KERR_FIELD), plasma = in.plasma;
    const double2* praman = RAMAN ? gl(in.escr) + 3 * MO : nullptr;  // P(t) of raman_field, packed pairs
    const double* tw = gl(in.to_win);
    const double dt = in.dt_o, pp1 = in.preion_p1, er = in.e_ratio, ip = in.ionpot;
    double cF = 0.0, cJ = 0.0, cP = 0.0;           // carried integrals
    double lastRate = 0.0, lastPhi = 0.0, lastLoss = 0.0;
    double Fprev = 0.0, fprev = pp1 - 1.0;         // exp(-0) = 1
    int sl = 0;                                    // scan slot of the next workgroup scan
    if constexpr (WIN) {
        cF = cw->cF, cJ = cw->cJ, cP = cw->cP;
        lastRate = cw->lastRate, lastPhi = cw->lastPhi, lastLoss = cw->lastLoss;
        Fprev = cw->Fprev, fprev = cw->fprev;
        sl = cw->sl;
    }
#pragma unroll 1
    for (int r = r0; r < r0 + RW; r++) {
        const int g0 = r * (LH_T * QR) + QR * tid;
        double a[QR];
        if (plasma) {
            double jf[QR];
            int bad = 0;
            constexpr int QC = (QR < 4) ? QR : 4;
            constexpr int PCAP = PList<MO>::CAP;
#pragma unroll
            for (int q = 0; q < QR; q++) jf[q] = fabs(field_at(sm, g0 + q - lb, es));
            if constexpr (MO <= 4096) {
                // samples inside the table's range get a rate; above it (or NaN) raise the flag
                unsigned mask = 0;
                int cnt = 0;
#pragma unroll
                for (int q = 0; q < QR; q++) {
                    const double x = jf[q];
                    const bool below = x < in.ppt_emin;
                    const bool out = !below && !(x <= in.ppt_emax);
                    if (out) bad = (x > in.ppt_emax) ? LH_E_PPT_RANGE : LH_E_NONFINITE;
                    if (!below && !out) {
                        mask |= 1u << q;
                        cnt++;
                    }
                }
                double ntot;
                const int off = (int)block_exscan1((double)cnt, sm.scan[sl], &ntot);
                sl ^= 1;
                const int total = (int)ntot;
                if (total <= PCAP) {  // compacted: one or two samples per thread
                    // sample q of this thread sits at off + (set bits of mask below q)
#pragma unroll
                    for (int q = 0; q < QR; q++)
                        if ((mask >> q) & 1u) sm.plist[off + __popc(mask & ((1u << q) - 1u))] = jf[q];
                    __syncthreads();
                    int unused = 0;
                    for (int i = tid; i < total; i += LH_T) {
                        double r;
                        ppt_rates<1>(in, &sm.plist[i], &r, unused);
                        sm.plist[i] = r;
                    }
                    __syncthreads();
#pragma unroll
                    for (int q = 0; q < QR; q++)
                        a[q] = ((mask >> q) & 1u) ? sm.plist[off + __popc(mask & ((1u << q) - 1u))] : 0.0;
                } else {
#pragma unroll
                    for (int c = 0; c < QR; c += QC) ppt_rates<QC>(in, jf + c, a + c, bad);
                }
            } else {
                // The samples that need a rate (the few hundred under the
                // pulse) lie in a window [lo, hi] of the round; when it fits the
                // LDS list, all threads take its samples in turn (one or two
                // each) instead of the ~16 threads that own them doing all of it.
                int lo = 0x7fffffff, hi = -1;
#pragma unroll
                for (int q = 0; q < QR; q++)
                    if (!(jf[q] < in.ppt_emin)) {
                        lo = min(lo, g0 + q);
                        hi = max(hi, g0 + q);
                    }
                block_minmax(lo, hi, sm.scan[sl]);
                sl ^= 1;
                if (hi < lo) {
#pragma unroll
                    for (int q = 0; q < QR; q++) a[q] = 0.0;
                } else if (hi - lo < PCAP) {
                    for (int g = lo + tid; g <= hi; g += LH_T) {
                        double ae = fabs(field_at(sm, g - lb, es)), rr = 0.0;
                        if (!(ae < in.ppt_emin)) ppt_rates<1>(in, &ae, &rr, bad);
                        sm.plist[g - lo] = rr;
                    }
                    __syncthreads();
#pragma unroll
                    for (int q = 0; q < QR; q++) {
                        const int g = g0 + q;
                        a[q] = (g >= lo && g <= hi) ? sm.plist[g - lo] : 0.0;
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < QR; c += QC) ppt_rates<QC>(in, jf + c, a + c, bad);
                }
            }
#pragma unroll
            for (int q = 0; q < QR; q++) jf[q] = a[q];  // rate, kept for the loss term
            if (bad) sm.flag = bad;
            lastRate = block_cumtrapz<QR>(a, dt, g0, sm.scan[sl], cF, lastRate);  // a = ∫rate
            sl ^= 1;
#pragma unroll
            for (int q = 0; q < QR; q++) {
                double E = field_at(sm, g0 + q - lb, es);
                double F = a[q];
                double f = (F == Fprev) ? fprev : pp1 - exp(-F);  // F is flat outside the pulse
                Fprev = F;
                fprev = f;
                double rate = jf[q];
                a[q] = (f * er) * E;                                                     // phase
                jf[q] = (rate != 0.0 && E != 0.0) ? ((ip * rate) * (1 - f)) / E : 0.0;  // loss current
            }
            // a = P = ∫(J + loss), J = ∫(phase term)
            block_cumtrapz2<QR>(a, jf, dt, g0, sm.scan[sl], cJ, cP, lastPhi, lastLoss);
            sl ^= 1;
        }
#pragma unroll
        for (int s2 = 0; s2 < QR / 2; s2++) {
            const int g = g0 + 2 * s2;
            double2 v = sm.buf[pad((g - lb) >> 1)];
            double E0 = v.x * es, E1 = v.y * es;
            double p0 = kerr ? kfac * (E0 * E0 * E0) : 0.0;
            double p1 = kerr ? kfac * (E1 * E1 * E1) : 0.0;
            if (plasma) {
                p0 = p0 + rho * a[2 * s2];
                p1 = p1 + rho * a[2 * s2 + 1];
            }
            if constexpr (RAMAN) {  // RamanPolarField: Pout = ρ·E·P (src/Nonlinear.jl:323-325)
                const double2 pr = praman[g >> 1];
                p0 = p0 + (rho * E0) * pr.x;
                p1 = p1 + (rho * E1) * pr.y;
            }
            sm.buf[pad((g - lb) >> 1)] = make_double2(p0 * tw[g], p1 * tw[g + 1]);
        }
    }
    if constexpr (WIN) {
        cw->cF = cF, cw->cJ = cJ, cw->cP = cP;
        cw->lastRate = lastRate, cw->lastPhi = lastPhi, cw->lastLoss = lastLoss;
        cw->Fprev = Fprev, cw->fprev = fprev;
        cw->sl = sl;
    }
}

// Epilogues of a transform: called with v[i] = the RHS of bin
// k = (the thread's index) + (i0 + i)·LH_T for a chunk of at most 9 of this thread's
// bins (bins k ≥ n hold zero and are not written).
struct EpiStore {  // out[k] = v
    static constexpr bool FOLDS_OUTFAC = false;
    double2* out;
    template <int NP>
    __device__ __forceinline__ void operator()(const double2 (&v)[NP], int n, int i0) const {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const int k = fresh_tid() + (i0 + i) * LH_T;
            if (k < n) gl(out)[k] = v[i];
        }
    }
};
// out[k] = v·F[k], F stored by the stage prologue: on the envelope grids
// F = exp(L·(t − t2)) and v carries the output factor (FOLD false); on the
// field grid F = outfac·exp(L·(t − t2)) and v is the bare r2c bin (FOLD true).
template <bool FOLD>
struct EpiStageFactor {
    static constexpr bool FOLDS_OUTFAC = FOLD;
    double2* out;
    const double2* F;
    template <int NP>
    __device__ __forceinline__ void operator()(const double2 (&v)[NP], int n, int i0) const {
        static_assert(NP <= 9, "epilogue chunk");
        double2 f[NP];  // the chunk's loads before its first store (out may alias F as far as the compiler knows)
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const int k = fresh_tid() + (i0 + i) * LH_T;
            f[i] = (k < n) ? static_cast<const double2*>(gl(F))[k] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const int k = fresh_tid() + (i0 + i) * LH_T;
            if (k < n) gl(out)[k] = cmul(v[i], f[i]);
        }
    }
};
using EpiStageF = EpiStageFactor<false>;
using EpiStageFO = EpiStageFactor<true>;

// Field-grid mode-averaged transform on LDS data.  On entry buf[0..N) holds
// sc_t·Eω (zero to MO), or (PACKED) its c2r pre-pack already.  Hands the
// output bins k < N to epi.
template <int MO, class Smem, class Epi, bool PACKED, bool RAMAN>
__device__ __forceinline__ void rhs_field(const DInst& in, Smem& sm, Epi epi, double rho, double kfac) {
    __syncthreads();
    if constexpr (!PACKED) {
        c2r_prepack<MO>(sm.buf, sm.tw);
        __syncthreads();
    }
    fft_lds<MO, true>(sm.buf, sm.tw);
    if constexpr (RAMAN) raman_field<MO, Smem>(in, sm);
    pointwise_field<MO, Smem, RAMAN>(in, sm, rho, kfac);
    constexpr int N = Bins<MO>::N, NPT = Bins<MO>::NPT;
    static_assert(NPT <= 9, "one epilogue chunk");
    double2 res[NPT];
    if constexpr (Epi::FOLDS_OUTFAC) {
        __syncthreads();
        fft_lds<MO, false>(sm.buf, sm.tw);
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            int k = fresh_tid() + i * LH_T;
            res[i] = (k < N) ? r2c_bin<MO>(sm.buf, sm.tw, k) : make_double2(0.0, 0.0);
        }
    } else {
        double2 of[NPT];  // loads in flight during the forward FFT
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            int k = fresh_tid() + i * LH_T;
            of[i] = (k < N) ? static_cast<const double2*>(gl(in.outfac))[k] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        fft_lds<MO, false>(sm.buf, sm.tw);
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            int k = fresh_tid() + i * LH_T;
            res[i] = (k < N) ? cmul(r2c_bin<MO>(sm.buf, sm.tw, k), of[i]) : make_double2(0.0, 0.0);
        }
    }
    epi(res, N, 0);
}

// The seven stage vectors k1..k7, addressed arithmetically: a loop over j
// computes the address of k_j in registers (the index is uniform over the
// workgroup) instead of reading a pointer array from private memory before
// every vector's loads.
struct KSet {
    double2* base;     // the vectors lie in one block, `stride` elements apart; slot 7:
                       // the envelope stage's exp(L·(t − t2))
    long long stride;
    int s0, s6;        // slots of k1 and k7 (swapped by FSAL); k2..k6 keep slots 1..5
    __device__ __forceinline__ gptr<double2> operator[](int j) const {
        const int s = (j == 0) ? s0 : (j == 6) ? s6 : j;
        return gl(base + (long long)s * stride);
    }
    __device__ __forceinline__ void swap06() {  // FSAL: ks[1] .= ks[7] as a pointer swap
        const int t = s0;
        s0 = s6;
        s6 = t;
    }
};

// Stage prologue: buf[k] = sc_t · (y + Σ_j Δz B_ij k_j)·exp(L hf); k1 first
// propagated by h0 (evaluate!, src/RK45.jl:209-223).  Loops run over the
// k-vectors outside and this thread's bins inside, so each vector's loads
// for all bins are in flight together; with nine bins per thread and FOLD,
// chunks of three bins instead load every vector a chunk needs in one round
// (three rounds per stage instead of three to five: 125.3 → 122.2 ms for the
// Nto = 16384 class of the sweep, bit-identical results).
// FOLD (field grid): also stores the stage epilogue's factor
// outfac·exp(L·(t − t2)) = outfac·exp(−L·hf) into slot 7 of K, from the same
// exp and sincos as exp(L·hf) (the epilogue then needs neither a
// transcendental pass nor the output factor held across the forward FFT).
template <int MO, class Smem, class KS, bool PACK, bool FOLD = false>
__device__ __forceinline__ void stage_prologue_cols(const double2* __restrict__ linop, double sc_t, Smem& sm,
                                                    const double2* Y, const KS& K, int stage, double dt,
                                                    double h0, double hf, const double2* outfac = nullptr);

template <int MO, class Smem>
__device__ __forceinline__ void stage_prologue(const DInst& in, Smem& sm, const double2* Y, const KSet& K,
                                               int stage, double dt, double h0, double hf) {
    stage_prologue_cols<MO, Smem, KSet, true, true>(gl(in.linop), in.sc_t, sm, Y, K, stage, dt, h0, hf, gl(in.outfac));
}

// Same on explicit per-column pointers (one column of a radial state).
template <int MO, class Smem, class KS, bool PACK, bool FOLD>
__device__ __forceinline__ void stage_prologue_cols(const double2* __restrict__ linop, double sc_t, Smem& sm,
                                                    const double2* Y, const KS& K, int stage, double dt,
                                                    double h0, double hf, const double2* outfac) {
    constexpr int N = Bins<MO>::N, NPT = Bins<MO>::NPT;
    const int tid = fresh_tid();
    if constexpr (FOLD && NPT == 9) {
        // chunks of three bins per thread, each with every vector it needs
        // (y, L, the output factor and k1..k_{stage+1}) loaded in one round
        constexpr int CB = 3;
        double2* F = K[7];
#pragma unroll 1
        for (int c = 0; c < NPT; c += CB) {
            double2 yb[CB], L[CB], of[CB], kk[6][CB];
#pragma unroll
            for (int i = 0; i < CB; i++) {
                const int k = tid + (c + i) * LH_T;
                const bool v = k < N;
                yb[i] = v ? Y[k] : make_double2(0.0, 0.0);
                L[i] = v ? linop[k] : make_double2(0.0, 0.0);
                of[i] = v ? outfac[k] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const double2* Kj = K[j];
                const bool use = j <= stage && c_B[stage][j] != 0.0;
#pragma unroll
                for (int i = 0; i < CB; i++) {
                    const int k = tid + (c + i) * LH_T;
                    kk[j][i] = (use && k < N) ? Kj[k] : make_double2(0.0, 0.0);
                }
            }
            if (stage == 0 && h0 != 0.0) {
                double2* K0 = K[0];
#pragma unroll
                for (int i = 0; i < CB; i++) {
                    const int k = tid + (c + i) * LH_T;
                    kk[0][i] = cmul(kk[0][i], cexp_lh(L[i], h0));
                    if (k < N) K0[k] = kk[0][i];
                }
            }
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const double cj = dt * c_B[stage][j];
                if (j <= stage && cj != 0.0) {
#pragma unroll
                    for (int i = 0; i < CB; i++) yb[i] = cadd(yb[i], cscal(cj, kk[j][i]));
                }
            }
#pragma unroll
            for (int i = 0; i < CB; i++) {
                const int k = tid + (c + i) * LH_T;
                if (k < N) {
                    double2 yp = yb[i], inv = make_double2(1.0, 0.0);
                    if (hf != 0.0) yp = cmul(yb[i], cexp_lh_inv(L[i], hf, inv));
                    c2r_pack_bin<MO>(sm.buf, sm.tw, k, cscal(sc_t, yp));
                    F[k] = cmul(of[i], inv);
                }
            }
        }
        return;
    }
    double2 yb[NPT], L[NPT];
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        int k = tid + i * LH_T;
        bool v = k < N;
        yb[i] = v ? Y[k] : make_double2(0.0, 0.0);
        L[i] = v ? linop[k] : make_double2(0.0, 0.0);
    }
    if (stage == 0 && h0 != 0.0) {
        double2* K0 = K[0];
        constexpr int CH = (NPT > 9) ? 8 : NPT;  // a chunk's loads before its first store
        static_assert(NPT % CH == 0, "prologue chunks");
#pragma unroll
        for (int c = 0; c < NPT; c += CH) {
            double2 kv[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                int k = tid + (c + i) * LH_T;
                kv[i] = (k < N) ? K0[k] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int i = 0; i < CH; i++) {
                int k = tid + (c + i) * LH_T;
                if (k < N) K0[k] = cmul(kv[i], cexp_lh(L[c + i], h0));
            }
        }
    }
    // two k-vectors per iteration: both vectors' loads are in flight together
    // (the sum still runs over j in order)
#pragma unroll 1
    for (int j = 0; j <= stage; j += 2) {
        const double c0 = dt * c_B[stage][j];
        const double c1 = (j + 1 <= stage) ? dt * c_B[stage][j + 1] : 0.0;
        const double2* K0j = K[j];
        const double2* K1j = K[j + 1];
        double2 a[NPT], b[NPT];
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            int k = tid + i * LH_T;
            a[i] = (k < N && c0 != 0.0) ? K0j[k] : make_double2(0.0, 0.0);
            b[i] = (k < N && c1 != 0.0) ? K1j[k] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            if (c0 != 0.0) yb[i] = cadd(yb[i], cscal(c0, a[i]));
            if (c1 != 0.0) yb[i] = cadd(yb[i], cscal(c1, b[i]));
        }
    }
    if constexpr (FOLD) {
        double2* F = K[7];
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            int k = tid + i * LH_T;
            if (k < N) {
                const double2 of = outfac[k];
                double2 yp = yb[i], inv = make_double2(1.0, 0.0);
                if (hf != 0.0) yp = cmul(yb[i], cexp_lh_inv(L[i], hf, inv));
                c2r_pack_bin<MO>(sm.buf, sm.tw, k, cscal(sc_t, yp));
                F[k] = cmul(of, inv);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            int k = tid + i * LH_T;
            if (k < N) {
                double2 yp = (hf != 0.0) ? cmul(yb[i], cexp_lh(L[i], hf)) : yb[i];
                if constexpr (PACK)  // the c2r pre-pack of the RHS, bin by bin (N − 1 = MO/2)
                    c2r_pack_bin<MO>(sm.buf, sm.tw, k, cscal(sc_t, yp));
                else
                    sm.buf[pad(k)] = cscal(sc_t, yp);
            }
        }
    }
    if constexpr (!PACK)
        for (int k = N + tid; k <= MO; k += LH_T) sm.buf[pad(k)] = make_double2(0.0, 0.0);
}

// Statistics of the spectrum held in LDS (defined with the statistics below).
// Partial energies Σ_k window_w[w][k]·|Eω_k|² of a thread's bins for the
// energy windows (energy_window: energyfun_ω(Eω .* window), src/Stats.jl:69-79).
struct WinSums {
    double s[LH_STAT_MAX_WINDOWS];
};
__device__ __forceinline__ void win_add(const DInst& in, WinSums& w, int k, double a) {
#pragma unroll
    for (int q = 0; q < LH_STAT_MAX_WINDOWS; q++)
        if (q < in.st_nwin) w.s[q] += gl(in.st_win)[(size_t)q * in.n_omega + k] * a;
}

template <int KIND, int NT, class Smem>
__device__ __forceinline__ void stats_time(const DInst& in, Smem& sm, double s0, double s1, double se,
                                           const WinSums& ws, double z, double dz, double rho, double* rec);

// Accepted-step window filter (stepfun, src/Luna.jl:366-370) on the coarse
// grid: YN ← rfft(twin · irfft(ωwin · YN·exp(L h))).  With STATS and a
// record pointer, the statistics of the filtered YN follow from the values
// still in registers (no read-back of YN).
template <int MC, class Smem, bool STATS = false>
__device__ __forceinline__ void stepfun_filter_cols(const DInst& in, const double2* __restrict__ linop, Smem& sm,
                                                    double2* YN, double h, double* rec = nullptr, double z = 0.0,
                                                    double dz = 0.0, double rho = 0.0);

template <int MC, class Smem>
__device__ __forceinline__ void stepfun_filter(const DInst& in, Smem& sm, double2* YN, double h, double* rec,
                                               double z, double dz, double rho) {
    stepfun_filter_cols<MC, Smem, true>(in, gl(in.linop), sm, YN, h, rec, z, dz, rho);
}

template <int MC, class Smem, bool STATS>
__device__ __forceinline__ void stepfun_filter_cols(const DInst& in, const double2* __restrict__ linop, Smem& sm,
                                                    double2* YN, double h, double* rec, double z, double dz,
                                                    double rho) {
    constexpr int N = MC + 1;
    constexpr int NPT = (N + LH_T - 1) / LH_T;
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        int k = fresh_tid() + i * LH_T;
        if (k < N) {
            double2 v = YN[k];
            if (h != 0.0) v = cmul(v, cexp_lh(linop[k], h));
            sm.buf[pad(k)] = cscal(gl(in.omega_win)[k], v);
        }
    }
    __syncthreads();
    c2r_prepack<MC>(sm.buf, sm.tw);
    __syncthreads();
    fft_lds<MC, true>(sm.buf, sm.tw);
    const double inv_nt = in.inv_nt;
    for (int j = fresh_tid(); j < MC; j += LH_T) {
        double2 z = sm.buf[pad(j)];
        double x0 = (inv_nt * z.x) * gl(in.t_win)[2 * j];
        double x1 = (inv_nt * z.y) * gl(in.t_win)[2 * j + 1];
        sm.buf[pad(j)] = make_double2(x0, x1);
    }
    __syncthreads();
    fft_lds<MC, false>(sm.buf, sm.tw);
    if constexpr (STATS) {
        if (rec) {
            // analytic spectrum (copyto_fft!, src/Stats.jl:441-448): bins
            // 0..Nω-2 doubled, the +fs/2 bin dropped, zero up to Nt = 2·MC
            double2 yv[NPT];
            double wv[NPT], ev[NPT];
#pragma unroll
            for (int i = 0; i < NPT; i++) {
                int k = fresh_tid() + i * LH_T;
                wv[i] = (k < N) ? gl(in.st_omega)[k] : 0.0;
                ev[i] = (k < N) ? gl(in.st_ew)[k] : 0.0;
                yv[i] = make_double2(0.0, 0.0);
                if (k < N) {
                    yv[i] = r2c_bin<MC>(sm.buf, sm.tw, k);
                    YN[k] = yv[i];
                }
            }
            __syncthreads();  // every bin read before the buffer is reused
            double s0 = 0.0, s1 = 0.0, se = 0.0;
            WinSums ws = {{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
            for (int i = 0; i < NPT; i++) {
                int k = fresh_tid() + i * LH_T;
                double a = yv[i].x * yv[i].x + yv[i].y * yv[i].y;
                s0 += a;
                s1 += wv[i] * a;
                se += ev[i] * a;
                if (k < N) win_add(in, ws, k, a);
                if (k < N - 1) sm.buf[pad(k)] = make_double2(2.0 * yv[i].x, 2.0 * yv[i].y);
            }
            for (int k = N - 1 + fresh_tid(); k < 2 * MC; k += LH_T) sm.buf[pad(k)] = make_double2(0.0, 0.0);
            stats_time<0, 2 * MC, Smem>(in, sm, s0, s1, se, ws, z, dz, rho, rec);
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        int k = fresh_tid() + i * LH_T;
        if (k < N) YN[k] = r2c_bin<MC>(sm.buf, sm.tw, k);
    }
    __syncthreads();
}

// ==========================================================================
// Envelope grid (EnvGrid, GNLSE): complex fields, Kerr_env and RamanPolarEnv
// ==========================================================================
template <int NO, int NN>
struct EnvBins {
    static constexpr int N = NN;
    static constexpr int NPT = (NN + LH_T - 1) / LH_T;
};

// Index of state bin k in the oversampled FFT array (copy_scale_both!,
// src/NonlinearRHS.jl:89-95): first N/2 bins at the start, last N/2 at the end.
template <int NO, int NN>
__device__ __forceinline__ int env_slot(int k) {
    return (k < NN / 2) ? k : k + (NO - NN);
}

template <int NO, int NN, class Smem>
__device__ __forceinline__ void stage_prologue_env(const DInst& in, Smem& sm, const double2* Y, const KSet& K,
                                                   int stage, double dt, double h0, double hf) {
    constexpr int N = NN, NPT = EnvBins<NO, NN>::NPT;
    const int tid = fresh_tid();
    double2 yb[NPT], L[NPT];
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        int k = tid + i * LH_T;
        bool v = k < N;
        yb[i] = v ? Y[k] : make_double2(0.0, 0.0);
        L[i] = v ? static_cast<const double2*>(gl(in.linop))[k] : make_double2(0.0, 0.0);
    }
    if (stage == 0 && h0 != 0.0) {
        double2* K0 = K[0];
        constexpr int CH = (NPT > 9) ? 8 : NPT;  // a chunk's loads before its first store
        static_assert(NPT % CH == 0, "prologue chunks");
#pragma unroll
        for (int c = 0; c < NPT; c += CH) {
            double2 kv[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                int k = tid + (c + i) * LH_T;
                kv[i] = (k < N) ? K0[k] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int i = 0; i < CH; i++) {
                int k = tid + (c + i) * LH_T;
                if (k < N) K0[k] = cmul(kv[i], cexp_lh(L[c + i], h0));
            }
        }
    }
#pragma unroll 1
    for (int j = 0; j <= stage; j++) {
        const double cj = dt * c_B[stage][j];
        if (cj == 0.0) continue;
        const double2* Kj = K[j];
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            int k = tid + i * LH_T;
            if (k < N) yb[i] = cadd(yb[i], cscal(cj, Kj[k]));
        }
    }
    const double sc_t = in.sc_t;
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        int k = tid + i * LH_T;
        if (k < N) {
            // exp(L·hf), and for the stage epilogue exp(L·(t − t2)) = exp(−L·hf)
            // from the same exp and sincos into slot 7 (saves the envelope
            // epilogue, with 16 bins per thread, its own transcendental pass)
            double2 yp = yb[i], inv = make_double2(1.0, 0.0);
            if (hf != 0.0) yp = cmul(yb[i], cexp_lh_inv(L[i], hf, inv));
            sm.buf[pad(env_slot<NO, NN>(k))] = cscal(sc_t, yp);
            K[7][k] = inv;
        }
    }
    if constexpr (NO > NN)
        for (int k = NN / 2 + tid; k < NO - NN / 2; k += LH_T) sm.buf[pad(k)] = make_double2(0.0, 0.0);
}

// TransModeAvg on an envelope grid with Kerr_env and RamanPolarEnv
// (src/NonlinearRHS.jl:25-32, 47-53, 362-373, 389-399; src/Nonlinear.jl:47-49,
// 63-73, 282-333).  The Raman convolution of ½|E|² with h(t) over the
// zero-padded 2·Nto grid is done as a real transform: r2c of 2·Nto reals (a
// length-Nto complex FFT), ×(hω δt), c2r — three LDS FFTs, with E(t) kept in
// HBM meanwhile.  Calls epi(k, nl_k) for every state bin k < Nω.
template <int NO, int NN, class Smem, class Epi>
__device__ __forceinline__ void rhs_env(const DInst& in, Smem& sm, Epi epi) {
    constexpr int PER = NO / LH_T;  // complex samples per thread (contiguous chunk)
    const int tid = fresh_tid();
    const int i0 = PER * tid;
    __syncthreads();
    fft_lds<NO, true>(sm.buf, sm.tw);
    const double es = in.e_scale;
    if (in.raman) {
        double2 E[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            double2 z = sm.buf[pad(i0 + q)];
            E[q] = make_double2(z.x * es, z.y * es);
            gl(in.escr)[i0 + q] = E[q];
        }
        __syncthreads();  // every thread has read its E before the packing overwrites
#pragma unroll
        for (int q = 0; q < PER / 2; q++) {
            double a0 = 0.5 * (E[2 * q].x * E[2 * q].x + E[2 * q].y * E[2 * q].y);
            double a1 = 0.5 * (E[2 * q + 1].x * E[2 * q + 1].x + E[2 * q + 1].y * E[2 * q + 1].y);
            sm.buf[pad(i0 / 2 + q)] = make_double2(a0, a1);        // ½|E|², reals packed in pairs
            sm.buf[pad(NO / 2 + i0 / 2 + q)] = make_double2(0.0, 0.0);  // zero padding to 2·Nto
        }
        __syncthreads();
        fft_lds<NO, false>(sm.buf, sm.tw);
        // r2c post-step of the 2·Nto-point transform, ×hω·δt, and the c2r
        // pre-pack of the product, pairwise (k, Nto-k) in place
        constexpr int TS = LH_TWN / (2 * NO);
        for (int k = tid; k <= NO / 2; k += LH_T) {
            int m = NO - k;
            if (k == 0) {
                double2 Y0 = cmul(gl(in.hwdt)[0], r2c_bin<NO>(sm.buf, sm.tw, 0));
                double2 YM = cmul(gl(in.hwdt)[NO], r2c_bin<NO>(sm.buf, sm.tw, NO));
                Y0.y = 0.0;
                YM.y = 0.0;
                double2 A = cadd(Y0, cconj(YM)), B = csub(Y0, cconj(YM));
                sm.buf[pad(0)] = cadd(A, cmul_pi(B));
            } else if (k == NO / 2) {
                double2 Yk = cmul(gl(in.hwdt)[k], r2c_bin<NO>(sm.buf, sm.tw, k));
                double2 A = cadd(Yk, cconj(Yk));
                double2 B = cmul(csub(Yk, cconj(Yk)), cconj(tw_get(sm.tw, k * TS)));
                sm.buf[pad(k)] = cadd(A, cmul_pi(B));
            } else {
                double2 Yk = cmul(gl(in.hwdt)[k], r2c_bin<NO>(sm.buf, sm.tw, k));
                double2 Ym = cmul(gl(in.hwdt)[m], r2c_bin<NO>(sm.buf, sm.tw, m));
                double2 wk = cconj(tw_get(sm.tw, k * TS)), wm = cconj(tw_get(sm.tw, m * TS));
                double2 Ak = cadd(Yk, cconj(Ym)), Bk = cmul(csub(Yk, cconj(Ym)), wk);
                double2 Am = cadd(Ym, cconj(Yk)), Bm = cmul(csub(Ym, cconj(Yk)), wm);
                sm.buf[pad(k)] = cadd(Ak, cmul_pi(Bk));
                sm.buf[pad(m)] = cadd(Am, cmul_pi(Bm));
            }
        }
        __syncthreads();
        fft_lds<NO, true>(sm.buf, sm.tw);  // P(t) on the 2·Nto grid, packed in pairs
        double P[PER];
#pragma unroll
        for (int q = 0; q < PER / 2; q++) {
            double2 z = sm.buf[pad(i0 / 2 + q)];
            P[2 * q] = z.x * in.inv_2n;
            P[2 * q + 1] = z.y * in.inv_2n;
        }
        __syncthreads();
        const double rho = in.density, k34 = in.kerr34;
        const bool kerr = (in.kerr == LH_KERR_ENV);
        const double* tw = gl(in.to_win);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            double2 e = gl(in.escr)[i0 + q];
            double2 p = make_double2(0.0, 0.0);
            if (kerr) {
                double s = k34 * (e.x * e.x + e.y * e.y);
                p = make_double2(s * e.x, s * e.y);
            }
            double2 re = make_double2(rho * e.x, rho * e.y);
            p = cadd(p, make_double2(re.x * P[q], re.y * P[q]));
            sm.buf[pad(i0 + q)] = make_double2(p.x * tw[i0 + q], p.y * tw[i0 + q]);
        }
    } else {
        const double k34 = in.kerr34;
        const bool kerr = (in.kerr == LH_KERR_ENV);
        const double* tw = gl(in.to_win);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            double2 z = sm.buf[pad(i0 + q)];
            double2 e = make_double2(z.x * es, z.y * es);
            double2 p = make_double2(0.0, 0.0);
            if (kerr) {
                double s = k34 * (e.x * e.x + e.y * e.y);
                p = make_double2(s * e.x, s * e.y);
            }
            sm.buf[pad(i0 + q)] = make_double2(p.x * tw[i0 + q], p.y * tw[i0 + q]);
        }
    }
    constexpr int N = NN, NPT = EnvBins<NO, NN>::NPT;
    double2 of[NPT];
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        int k = tid + i * LH_T;
        of[i] = (k < N) ? static_cast<const double2*>(gl(in.outfac))[k] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<NO, false>(sm.buf, sm.tw);
    constexpr int CH = (NPT > 9) ? 8 : NPT;
    static_assert(NPT % CH == 0, "epilogue chunks");
#pragma unroll
    for (int c = 0; c < NPT; c += CH) {
        double2 res[CH];
#pragma unroll
        for (int i = 0; i < CH; i++) {
            int k = tid + (c + i) * LH_T;
            res[i] = (k < N) ? cmul(sm.buf[pad(env_slot<NO, NN>(k))], of[c + i]) : make_double2(0.0, 0.0);
        }
        epi(res, N, c);
    }
}

// stepfun on the envelope grid (src/Luna.jl:366-370 with a c2c FT):
// YN ← fft(twin · ifft(ωwin · YN·exp(L h))).
template <int NT, class Smem>
__device__ __forceinline__ void stepfun_env(const DInst& in, Smem& sm, double2* YN, double h, double* rec, double z,
                                            double dz, double rho) {
    constexpr int NPT = (NT + LH_T - 1) / LH_T;
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        int k = fresh_tid() + i * LH_T;
        if (k < NT) {
            double2 v = YN[k];
            if (h != 0.0) v = cmul(v, cexp_lh(gl(in.linop)[k], h));
            sm.buf[pad(k)] = cscal(gl(in.omega_win)[k], v);
        }
    }
    __syncthreads();
    fft_lds<NT, true>(sm.buf, sm.tw);
    const double inv_nt = in.inv_nt;
    for (int j = fresh_tid(); j < NT; j += LH_T) {
        double2 z = sm.buf[pad(j)];
        double w = gl(in.t_win)[j];
        sm.buf[pad(j)] = make_double2((inv_nt * z.x) * w, (inv_nt * z.y) * w);
    }
    __syncthreads();
    fft_lds<NT, false>(sm.buf, sm.tw);
    double s0 = 0.0, s1 = 0.0, se = 0.0;
    WinSums ws = {{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        int k = fresh_tid() + i * LH_T;
        if (k < NT) {
            double2 v = sm.buf[pad(k)];
            YN[k] = v;
            if (rec) {  // the envelope's analytic field is ifft(Eω): the spectrum stays in place
                double a = v.x * v.x + v.y * v.y;
                s0 += a;
                s1 += gl(in.st_omega)[k] * a;
                se += gl(in.st_ew)[k] * a;
                win_add(in, ws, k, a);
            }
        }
    }
    if (rec)
        stats_time<1, NT, Smem>(in, sm, s0, s1, se, ws, z, dz, rho, rec);
    else
        __syncthreads();
}

// ==========================================================================
// Envelope grid beyond one workgroup's LDS (Nt = Nto = 16384, KIND 2).  Each
// transform of NO = 2·M points runs as two M-point LDS transforms (even and
// odd samples) joined by one radix-2 step in HBM; between transforms the
// series live in the instance's HBM work area escr = [E(t) | W | T]
// (NO + NO + M complex): E the scaled field, W the working series, T the
// half-length stash.  Bins per thread exceed the register budget of the
// LDS-resident classes, so bin loops run in chunks of 8.
// ==========================================================================
// exp(−2πi k/L) for the zero-padded Raman transforms: L up to LH_TWN from
// the table; L = 2·LH_TWN gives even k from the table and odd k one more
// factor exp(−2πi/L).
template <int L>
__device__ __forceinline__ double2 tw_frac(const TwLds& tw, int k) {
    if constexpr (L <= LH_TWN) {
        return tw_get(tw, k * (LH_TWN / L));
    } else {
        static_assert(L == 2 * LH_TWN, "tw_frac: L up to 2·LH_TWN");
        double2 w = tw_get(tw, k >> 1);
        if (k & 1) {
            double sn, cs;
            sincos(-6.283185307179586476925286766559 / (double)L, &sn, &cs);
            w = cmul(w, make_double2(cs, sn));
        }
        return w;
    }
}

// Unnormalised complex FFT of NO = 2·M points held in HBM (natural order, in
// place in g); tmp: M points of HBM.  X[k] = E[k] + w^k O[k], X[k+M] =
// E[k] − w^k O[k], with E, O the transforms of the even and odd samples and
// w = exp(∓2πi/NO).  Starts and ends with a workgroup barrier.
template <int M, bool INV, class Smem>
__device__ __forceinline__ void fft_big(double2* g, double2* tmp, Smem& sm) {
    const int tid = fresh_tid();
    __syncthreads();
    for (int j = tid; j < M; j += LH_T) sm.buf[pad(j)] = g[2 * j];
    __syncthreads();
    fft_lds<M, INV>(sm.buf, sm.tw);
    for (int j = tid; j < M; j += LH_T) tmp[j] = sm.buf[pad(j)];
    __syncthreads();
    for (int j = tid; j < M; j += LH_T) sm.buf[pad(j)] = g[2 * j + 1];
    __syncthreads();
    fft_lds<M, INV>(sm.buf, sm.tw);
    constexpr int TS = LH_TWN / (2 * M);
    for (int k = tid; k < M; k += LH_T) {
        double2 w = tw_get(sm.tw, k * TS);
        if (INV) w.y = -w.y;
        const double2 o = cmul(w, sm.buf[pad(k)]), e = tmp[k];
        g[k] = cadd(e, o);
        g[k + M] = csub(e, o);
    }
    __syncthreads();
}

// r2c post-step of a 2·NO-point real transform whose packed NO-point
// transform Z is in HBM: X[k] = E + w^k O (as r2c_bin).
template <int NO, class Smem>
__device__ __forceinline__ double2 r2c_bin_big(const double2* Z, const Smem& sm, int k) {
    double2 zk = Z[k == NO ? 0 : k];
    double2 zc = cconj(Z[k == 0 ? 0 : NO - k]);
    double2 E = cscal(0.5, cadd(zk, zc));
    double2 d = csub(zk, zc);
    double2 O = make_double2(0.5 * d.y, -(0.5 * d.x));
    double2 w = (k == NO) ? make_double2(-1.0, 0.0) : tw_frac<2 * NO>(sm.tw, k);
    return cadd(E, cmul(w, O));
}

// The convolution step of the zero-padded Raman transform (2·NO reals packed
// into the NO-point transform Z in HBM): r2c post-step, ×hω·δt, and the c2r
// pre-pack of the product, pairwise (k, NO − k) in place.
template <int NO, class Smem>
__device__ __forceinline__ void raman_spectrum_big(const DInst& in, const Smem& sm, double2* W) {
    const int tid = fresh_tid();
    for (int k = tid; k <= NO / 2; k += LH_T) {
        const int m = NO - k;
        if (k == 0) {
            double2 Y0 = cmul(gl(in.hwdt)[0], r2c_bin_big<NO>(W, sm, 0));
            double2 YM = cmul(gl(in.hwdt)[NO], r2c_bin_big<NO>(W, sm, NO));
            Y0.y = 0.0;
            YM.y = 0.0;
            const double2 A = cadd(Y0, cconj(YM)), B = csub(Y0, cconj(YM));
            W[0] = cadd(A, cmul_pi(B));
        } else if (k == NO / 2) {
            const double2 Yk = cmul(gl(in.hwdt)[k], r2c_bin_big<NO>(W, sm, k));
            const double2 A = cadd(Yk, cconj(Yk));
            const double2 B = cmul(csub(Yk, cconj(Yk)), cconj(tw_frac<2 * NO>(sm.tw, k)));
            W[k] = cadd(A, cmul_pi(B));
        } else {
            const double2 Yk = cmul(gl(in.hwdt)[k], r2c_bin_big<NO>(W, sm, k));
            const double2 Ym = cmul(gl(in.hwdt)[m], r2c_bin_big<NO>(W, sm, m));
            const double2 wk = cconj(tw_frac<2 * NO>(sm.tw, k)), wm = cconj(tw_frac<2 * NO>(sm.tw, m));
            const double2 Ak = cadd(Yk, cconj(Ym)), Bk = cmul(csub(Yk, cconj(Ym)), wk);
            const double2 Am = cadd(Ym, cconj(Yk)), Bm = cmul(csub(Ym, cconj(Yk)), wm);
            W[k] = cadd(Ak, cmul_pi(Bk));
            W[m] = cadd(Am, cmul_pi(Bm));
        }
    }
}

// RamanPolarField with third-harmonic terms (src/Nonlinear.jl:183-247,
// 284-333, thg = true) on the field grid, for a series of Nto = 2·MO reals
// held packed in LDS (z[j] = x[2j] + i x[2j+1], before the e_scale factor).
// E² over the zero-padded 2·Nto reals is one 2·MO-point complex transform,
// run through HBM as two MO-point LDS transforms (fft_big); ×hω·δt; back;
// P(t) = irfft/(2·Nto) lands in escr[3·MO, 4·MO) as packed pairs, which the
// pointwise phase adds as ρ·E·P.  The LDS series is restored unchanged from
// its HBM copy escr[0, MO); escr[MO, 3·MO) is the transform's series.
template <int MO, class Smem>
__device__ __forceinline__ void raman_field(const DInst& in, Smem& sm) {
    constexpr int NR = 2 * MO;
    double2* S = gl(in.escr);
    double2* W = gl(in.escr) + MO;
    double2* T = gl(in.escr) + 3 * MO;
    const int tid = fresh_tid();
    const double es = in.e_scale;
    for (int j = tid; j < MO; j += LH_T) {
        const double2 z = sm.buf[pad(j)];
        const double e0 = z.x * es, e1 = z.y * es;
        S[j] = z;
        W[j] = make_double2(e0 * e0, e1 * e1);  // E² (thg: E.^2), reals packed in pairs
        W[j + MO] = make_double2(0.0, 0.0);     // zero padding to 2·Nto reals
    }
    fft_big<MO, false>(W, T, sm);
    raman_spectrum_big<NR>(in, sm, W);
    fft_big<MO, true>(W, T, sm);
    const double inv2n = in.inv_2n;
    for (int j = tid; j < MO; j += LH_T) {
        const double2 p = W[j];
        T[j] = make_double2(p.x * inv2n, p.y * inv2n);
        sm.buf[pad(j)] = S[j];
    }
    __syncthreads();
}

// Stage prologue on the big envelope grid: W = sc_t·(y + Δz Σ B k)·exp(L hf),
// exp(−L hf) into slot 7 (as stage_prologue_env, chunked).
template <int NN, class Smem>
__device__ __forceinline__ void stage_prologue_big(const DInst& in, Smem& sm, const double2* Y, const KSet& K,
                                                   int stage, double dt, double h0, double hf) {
    constexpr int NPT = NN / LH_T, CH = 8;
    static_assert(NN % (LH_T * CH) == 0, "big-grid chunks");
    double2* W = gl(in.escr) + NN;
    const int tid = fresh_tid();
    const double sc_t = in.sc_t;
    double2* K7 = K[7];
#pragma unroll 1
    for (int c = 0; c < NPT; c += CH) {
        double2 yb[CH], L[CH];
#pragma unroll
        for (int i = 0; i < CH; i++) {
            const int k = tid + (c + i) * LH_T;
            yb[i] = Y[k];
            L[i] = gl(in.linop)[k];
        }
        if (stage == 0 && h0 != 0.0) {
            double2* K0 = K[0];
            double2 kv[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) kv[i] = K0[tid + (c + i) * LH_T];
#pragma unroll
            for (int i = 0; i < CH; i++) K0[tid + (c + i) * LH_T] = cmul(kv[i], cexp_lh(L[i], h0));
        }
#pragma unroll 1
        for (int j = 0; j <= stage; j++) {
            const double cj = dt * c_B[stage][j];
            if (cj == 0.0) continue;
            const double2* Kj = K[j];
#pragma unroll
            for (int i = 0; i < CH; i++) yb[i] = cadd(yb[i], cscal(cj, Kj[tid + (c + i) * LH_T]));
        }
#pragma unroll
        for (int i = 0; i < CH; i++) {
            const int k = tid + (c + i) * LH_T;
            double2 yp = yb[i], inv = make_double2(1.0, 0.0);
            if (hf != 0.0) yp = cmul(yb[i], cexp_lh_inv(L[i], hf, inv));
            W[k] = cscal(sc_t, yp);
            K7[k] = inv;
        }
    }
}

// TransModeAvg on the big envelope grid (the operations of rhs_env with the
// transforms in HBM): W holds sc_t·Eω on entry.
template <int NO, class Smem, class Epi>
__device__ __forceinline__ void rhs_big(const DInst& in, Smem& sm, Epi epi) {
    constexpr int M = NO / 2;
    double2* E = gl(in.escr);
    double2* W = gl(in.escr) + NO;
    double2* T = gl(in.escr) + 2 * NO;
    const int tid = fresh_tid();
    fft_big<M, true>(W, T, sm);
    const double es = in.e_scale, k34 = in.kerr34, rho = in.density;
    const bool kerr = (in.kerr == LH_KERR_ENV);
    const double* twin = gl(in.to_win);
    if (in.raman) {
        for (int i = tid; i < NO; i += LH_T) {
            const double2 z = W[i];
            E[i] = make_double2(z.x * es, z.y * es);
        }
        __syncthreads();
        for (int j = tid; j < NO; j += LH_T) {
            if (j < M) {
                const double2 a = E[2 * j], b = E[2 * j + 1];
                W[j] = make_double2(0.5 * (a.x * a.x + a.y * a.y), 0.5 * (b.x * b.x + b.y * b.y));
            } else {
                W[j] = make_double2(0.0, 0.0);  // zero padding to 2·NO reals
            }
        }
        fft_big<M, false>(W, T, sm);
        raman_spectrum_big<NO>(in, sm, W);
        fft_big<M, true>(W, T, sm);  // P(t) on the 2·NO grid, packed in pairs: the first NO values in W[0..M)
        for (int j = tid; j < M; j += LH_T) T[j] = W[j];
        __syncthreads();
        const double inv2n = in.inv_2n;
        for (int i = tid; i < NO; i += LH_T) {
            const double2 pp = T[i >> 1];
            const double P = ((i & 1) ? pp.y : pp.x) * inv2n;
            const double2 e = E[i];
            double2 p = make_double2(0.0, 0.0);
            if (kerr) {
                const double s = k34 * (e.x * e.x + e.y * e.y);
                p = make_double2(s * e.x, s * e.y);
            }
            p = cadd(p, make_double2((rho * e.x) * P, (rho * e.y) * P));
            W[i] = make_double2(p.x * twin[i], p.y * twin[i]);
        }
    } else {
        for (int i = tid; i < NO; i += LH_T) {
            const double2 z = W[i];
            const double2 e = make_double2(z.x * es, z.y * es);
            double2 p = make_double2(0.0, 0.0);
            if (kerr) {
                const double s = k34 * (e.x * e.x + e.y * e.y);
                p = make_double2(s * e.x, s * e.y);
            }
            W[i] = make_double2(p.x * twin[i], p.y * twin[i]);
        }
    }
    fft_big<M, false>(W, T, sm);
    constexpr int NPT = NO / LH_T, CH = 8;
#pragma unroll 1
    for (int c = 0; c < NPT; c += CH) {
        double2 res[CH];
#pragma unroll
        for (int i = 0; i < CH; i++) {
            const int k = tid + (c + i) * LH_T;
            res[i] = cmul(W[k], gl(in.outfac)[k]);
        }
        epi(res, NO, c);
    }
}

// stepfun on the big envelope grid: YN ← fft(twin · ifft(ωwin · YN·exp(L h))).
template <int NT, class Smem>
__device__ __forceinline__ void stepfun_big(const DInst& in, Smem& sm, double2* YN, double h) {
    double2* W = gl(in.escr) + NT;
    double2* T = gl(in.escr) + 2 * NT;
    const int tid = fresh_tid();
    for (int k = tid; k < NT; k += LH_T) {
        double2 v = YN[k];
        if (h != 0.0) v = cmul(v, cexp_lh(gl(in.linop)[k], h));
        W[k] = cscal(gl(in.omega_win)[k], v);
    }
    fft_big<NT / 2, true>(W, T, sm);
    const double inv_nt = in.inv_nt;
    for (int j = tid; j < NT; j += LH_T) {
        const double2 z = W[j];
        const double w = gl(in.t_win)[j];
        W[j] = make_double2((inv_nt * z.x) * w, (inv_nt * z.y) * w);
    }
    fft_big<NT / 2, false>(W, T, sm);
    for (int k = tid; k < NT; k += LH_T) YN[k] = W[k];
    __syncthreads();
}

// ==========================================================================
// Field grid of Nto = 4·LDS points (KIND 3): the c2r/r2c transforms of the
// Nto reals are MO = Nto/2-point complex transforms staged through HBM
// (fft_big), the pointwise responses run over the series in two LDS windows.
// HBM work area escr = [W | T], MO + MO/2 complex: W the packed series,
// T the half-length stash of fft_big.
// ==========================================================================
// c2r pre-pack of bin k into Z in HBM for a spectrum zero above M/2 (as
// c2r_pack_bin, with the twiddles of 2M points from tw_frac).
template <int M>
__device__ __forceinline__ void c2r_pack_bin_big(double2* Z, const TwLds& tw, int k, double2 Xk) {
    const double2 Xm = make_double2(0.0, 0.0);
    if (k == 0) {
        Xk.y = 0.0;
        double2 A = cadd(Xk, cconj(Xm));
        double2 B = csub(Xk, cconj(Xm));
        Z[0] = cadd(A, cmul_pi(B));
    } else if (k == M / 2) {
        double2 A = cadd(Xk, cconj(Xk));
        double2 B = cmul(csub(Xk, cconj(Xk)), cconj(tw_frac<2 * M>(tw, k)));
        Z[k] = cadd(A, cmul_pi(B));
    } else {
        const int m = M - k;
        double2 wk = cconj(tw_frac<2 * M>(tw, k)), wm = cconj(tw_frac<2 * M>(tw, m));
        double2 Ak = cadd(Xk, cconj(Xm));
        double2 Bk = cmul(csub(Xk, cconj(Xm)), wk);
        double2 Am = cadd(Xm, cconj(Xk));
        double2 Bm = cmul(csub(Xm, cconj(Xk)), wm);
        Z[k] = cadd(Ak, cmul_pi(Bk));
        Z[m] = cadd(Am, cmul_pi(Bm));
    }
}

// Stage prologue of the big field grid (stage_prologue with FOLD, in chunks of
// 8 bins per thread): the c2r pre-pack of sc_t·(y + Δz Σ B k)·exp(L hf) into
// W, outfac·exp(−L hf) into slot 7; k1 first propagated by h0.
template <int MO, class Smem>
__device__ __forceinline__ void stage_prologue_bigfield(const DInst& in, Smem& sm, const double2* Y, const KSet& K,
                                                        int stage, double dt, double h0, double hf) {
    constexpr int N = MO / 2 + 1, CH = 8, NC = (N + LH_T * CH - 1) / (LH_T * CH);
    double2* W = gl(in.escr);
    const double2* linop = gl(in.linop);
    const double2* outfac = gl(in.outfac);
    const int tid = fresh_tid();
    const double sc_t = in.sc_t;
    double2* K7 = K[7];
#pragma unroll 1
    for (int c = 0; c < NC; c++) {
        double2 yb[CH], L[CH];
#pragma unroll
        for (int i = 0; i < CH; i++) {
            const int k = tid + (c * CH + i) * LH_T;
            yb[i] = (k < N) ? Y[k] : make_double2(0.0, 0.0);
            L[i] = (k < N) ? linop[k] : make_double2(0.0, 0.0);
        }
        if (stage == 0 && h0 != 0.0) {
            double2* K0 = K[0];
            double2 kv[CH];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int k = tid + (c * CH + i) * LH_T;
                kv[i] = (k < N) ? K0[k] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int k = tid + (c * CH + i) * LH_T;
                if (k < N) K0[k] = cmul(kv[i], cexp_lh(L[i], h0));
            }
        }
#pragma unroll 1
        for (int j = 0; j <= stage; j++) {
            const double cj = dt * c_B[stage][j];
            if (cj == 0.0) continue;
            const double2* Kj = K[j];
#pragma unroll
            for (int i = 0; i < CH; i++) {
                const int k = tid + (c * CH + i) * LH_T;
                if (k < N) yb[i] = cadd(yb[i], cscal(cj, Kj[k]));
            }
        }
        double2 ofv[CH];  // every load before the first store to slot 7
#pragma unroll
        for (int i = 0; i < CH; i++) {
            const int k = tid + (c * CH + i) * LH_T;
            ofv[i] = (k < N) ? outfac[k] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < CH; i++) {
            const int k = tid + (c * CH + i) * LH_T;
            if (k < N) {
                double2 yp = yb[i], inv = make_double2(1.0, 0.0);
                if (hf != 0.0) yp = cmul(yb[i], cexp_lh_inv(L[i], hf, inv));
                c2r_pack_bin_big<MO>(W, sm.tw, k, cscal(sc_t, yp));
                K7[k] = cmul(ofv[i], inv);
            }
        }
    }
}

// TransModeAvg on the big field grid: W holds the c2r pre-pack of sc_t·Eω on
// entry.  Inverse transform, the pointwise responses over the series in two
// LDS windows (carrying the plasma integrals), forward transform, r2c
// post-step; the epilogue gets bin k < Nω (times outfac unless it folds it).
template <int MO, class Smem, class Epi>
__device__ __forceinline__ void rhs_bigfield(const DInst& in, Smem& sm, Epi epi, double rho, double kfac) {
    constexpr int H = MO / 2;  // complex pairs per LDS window
    double2* W = gl(in.escr);
    double2* T = gl(in.escr) + MO;
    const int tid = fresh_tid();
    fft_big<H, true>(W, T, sm);
    PwCarry cr;
    cr.cF = cr.cJ = cr.cP = 0.0;
    cr.lastRate = cr.lastPhi = cr.lastLoss = 0.0;
    cr.Fprev = 0.0;
    cr.fprev = in.preion_p1 - 1.0;
    cr.sl = 0;
#pragma unroll 1
    for (int h = 0; h < 2; h++) {
        __syncthreads();
        for (int j = tid; j < H; j += LH_T) sm.buf[pad(j)] = W[h * H + j];
        __syncthreads();
        pointwise_field<MO, Smem, false, true, H>(in, sm, rho, kfac, &cr, 2 * h, 2 * H * h);
        __syncthreads();
        for (int j = tid; j < H; j += LH_T) W[h * H + j] = sm.buf[pad(j)];
    }
    fft_big<H, false>(W, T, sm);
    constexpr int N = MO / 2 + 1, CH = 8, NC = (N + LH_T * CH - 1) / (LH_T * CH);
    const double2* outfac = gl(in.outfac);
#pragma unroll 1
    for (int c = 0; c < NC; c++) {
        double2 res[CH];
#pragma unroll
        for (int i = 0; i < CH; i++) {
            const int k = tid + (c * CH + i) * LH_T;
            res[i] = make_double2(0.0, 0.0);
            if (k < N) {
                res[i] = r2c_bin_big<MO>(W, sm, k);
                if constexpr (!Epi::FOLDS_OUTFAC) res[i] = cmul(res[i], outfac[k]);
            }
        }
        epi(res, N, c * CH);
    }
    __syncthreads();
}

// hω·δt on the big grid: h zero-padded to 2·NO reals, packed into NO complex
// points in W, transformed, post-stepped into hwdt[0..NO].  One workgroup.
template <int NO>
__global__ __launch_bounds__(LH_T) void lh_raman_hw_big_kernel(const double* __restrict__ h, double dt,
                                                               const double2* __restrict__ twg, double2* work,
                                                               double2* hwdt) {
    __shared__ SmemT<NO / 2> sm;
    tw_load(sm.tw, twg);
    double2* W = work;
    double2* T = work + NO;
    for (int j = threadIdx.x; j < NO; j += LH_T)
        W[j] = (j < NO / 2) ? make_double2(h[2 * j], h[2 * j + 1]) : make_double2(0.0, 0.0);
    fft_big<NO / 2, false>(W, T, sm);
    for (int k = threadIdx.x; k <= NO; k += LH_T) {
        const double2 X = r2c_bin_big<NO>(W, sm, k);
        hwdt[k] = make_double2(X.x * dt, X.y * dt);
    }
}

// hω·δt for the Raman convolution: rfft of h(t) zero-padded to 2·Nto reals
// (the reference's FT * h on the doubled grid, src/Nonlinear.jl:307-312).
template <int NO>
__global__ __launch_bounds__(LH_T) void lh_raman_hw_kernel(const double* __restrict__ h, double dt,
                                                           const double2* __restrict__ twg, double2* hwdt) {
    __shared__ SmemT<NO> sm;
    tw_load(sm.tw, twg);
    for (int j = threadIdx.x; j < NO; j += LH_T)
        sm.buf[pad(j)] = (j < NO / 2) ? make_double2(h[2 * j], h[2 * j + 1]) : make_double2(0.0, 0.0);
    __syncthreads();
    fft_lds<NO, false>(sm.buf, sm.tw);
    for (int k = threadIdx.x; k <= NO; k += LH_T) {
        double2 X = r2c_bin<NO>(sm.buf, sm.tw, k);
        hwdt[k] = make_double2(X.x * dt, X.y * dt);
    }
}

// --------------------------------------------------------------------------
// Persistent adaptive solve of one instance (Luna.run → RK45.solve_precon).
// --------------------------------------------------------------------------
// KIND 0: field grid, MO = Nto/2 complex points, MC = Nt/2.
// KIND 1: envelope grid, MO = Nto complex points, MC = Nt = Nω.
// KIND 2: envelope grid with the transforms staged through HBM, MO = MC = Nt.
template <int KIND, int MO, int MC>
struct Sizes {
    static constexpr int N = (KIND == 0 || KIND == 3) ? Bins<MO>::N : MC;
    static constexpr int NPT = (N + LH_T - 1) / LH_T;
    static constexpr int LDS = (KIND == 2 || KIND == 3) ? MO / 2 : MO;  // LDS transform length (SmemT)
};

// ==========================================================================
// z-dependent gas fill (Capillary.gradient + LinearOps.make_linop)
// ==========================================================================
#define LH_EPS0 8.854187817620389e-12  // CODATA 2014 ε0
#define LH_CLIGHT 299792458.0

// 8th-order central first-derivative weights at offsets -4..4 (times 1/h)
__constant__ double c_fd8[9] = {1.0 / 280, -4.0 / 105, 1.0 / 5, -4.0 / 5, 0.0, 4.0 / 5, -1.0 / 5, 4.0 / 105, -1.0 / 280};

// Pressure profile p(z) of Capillary.gradient, src/Capillary.jl:306-340.
__device__ __forceinline__ double zg_pressure(const DInst& in, double z) {
    const double* Z = gl(in.zg_Z);
    const double* P = gl(in.zg_P);
    const int n = in.zg_np;
    if (in.zg_two) {
        const double L = Z[1], p0 = P[0], p1 = P[1];
        if (z > L) return p1;
        if (z <= 0) return p0;
        return sqrt(p0 * p0 + z / L * (p1 * p1 - p0 * p0));
    }
    if (z <= Z[0]) return P[0];
    if (z >= Z[n - 1]) return P[n - 1];
    int i = 0;
    for (int j = 0; j < n; j++)
        if (Z[j] < z) i = j;  // findlast(x -> x < z, Z)
    return sqrt(P[i] * P[i] + (z - Z[i]) / (Z[i + 1] - Z[i]) * (P[i + 1] * P[i + 1] - P[i] * P[i]));
}

// Principal square root of x + iy.
__device__ __forceinline__ double2 csqrt_lh(double x, double y) {
    if (x == 0.0 && y == 0.0) return make_double2(0.0, y);
    const double t = sqrt((fabs(x) + hypot(x, y)) * 0.5);
    if (x >= 0.0) return make_double2(t, y / (2.0 * t));
    return make_double2(fabs(y) / (2.0 * t), copysign(t, y));
}

// neff(mode, εco, nwg) of a Marcatili mode, src/Capillary.jl:216-234.
__device__ __forceinline__ double2 zg_neff(const DInst& in, double eco, double2 nwg) {
    double2 n;
    if (in.zg_model == 0)
        n = csqrt_lh(eco - nwg.x, -nwg.y);
    else
        n = make_double2(1 + (eco - 1) / 2 - nwg.x, -nwg.y);
    if (!in.zg_loss) n.y = 0.0;
    return n;
}

// Re-evaluate the z-dependent constants at z (make_prop!(linop!) evaluates
// linop!(out, t2) whenever t2 changes): ρ(z) and β1(z) on thread 0, then per
// bin the operator -i(ω/c·conj_clamp(neff) - ω β1) (src/LinearOps.jl:337-358,
// 254) and the output factor sc_f·pre/β·√Aeff·ωwin (β = ω/c·Re neff),
// written to gl(in.zL) / gl(in.zOF).  Returns ρ(z) and the Kerr factor ρ ε0 γ3.
// Radius of the taper at z: linear a0 + (a1 - a0)·clamp(z/L, 0, 1), or the
// piecewise-linear profile through (zt_Z, zt_A), constant outside.
__device__ __forceinline__ double zg_radius(const DInst& in, double z) {
    if (in.ztaper == 2) {
        const double *Z = gl(in.zt_Z), *A = gl(in.zt_A);
        const int n = in.zt_n;
        if (z <= Z[0]) return A[0];
        if (z >= Z[n - 1]) return A[n - 1];
        // the segment [Z[i], Z[i+1]] with Z[i] < z <= Z[i+1] (nodes strictly
        // increasing), by bisection: Z[lo] < z <= Z[hi] throughout
        int lo = 0, hi = n - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (Z[mid] < z)
                lo = mid;
            else
                hi = mid;
        }
        const int i = lo;
        return A[i] + (A[i + 1] - A[i]) * ((z - Z[i]) / (Z[i + 1] - Z[i]));
    }
    double f = z / in.zaL;
    f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
    return in.za0 + (in.za1 - in.za0) * f;
}

// neff_wg(mode, ω; z) at radius a from U = u_nm/k and V = v_n/k
// (src/Capillary.jl:188-214): (U/a)²(1 - iV/a)² (:full); (U/a)²/2, plus
// i(U/a)²V/a with loss (:reduced).
__device__ __forceinline__ double2 zg_nwg_taper(const DInst& in, double U, double2 V, double a) {
    const double q = U / a, q2 = q * q;
    if (in.zg_model == 0) {
        const double tr = 1.0 + V.y / a, ti = -(V.x / a);  // 1 - iV/a
        return make_double2(q2 * (tr * tr - ti * ti), q2 * (2.0 * tr * ti));
    }
    double2 r = make_double2(q2 / 2, 0.0);
    if (in.zg_loss) {
        r.x += -(q2 * (V.y / a));
        r.y += q2 * (V.x / a);
    }
    return r;
}

template <class Smem>
__device__ __forceinline__ void zg_refresh(const DInst& in, Smem& sm, double z, double& rho, double& kfac) {
    const int tid = fresh_tid();
    if (tid == 0) {
        const double p = zg_pressure(in, z);
        const double r = cspline_uniform(gl(in.zd), in.zd_n, gl(in.zd)[0].x, gl(in.zd)[in.zd_n - 1].x, p);
        // β1 = Modes.dispersion(mode, 1, ω0, z): β(ω) = ω/c·Re neff(mode, ω; z)
        // with neff's own clamp (src/Capillary.jl:162-186)
        const double a = in.ztaper ? zg_radius(in, z) : 0.0;
        double acc = 0.0;
        for (int j = 0; j < 9; j++) {
            const double* f = gl(in.zfd) + 4 * j;
            const double sc = sqrt(1 + f[1] * r);
            double2 nwj = make_double2(f[2], f[3]);
            if (in.ztaper) {
                const double* t = gl(in.zfdt) + 3 * j;
                nwj = zg_nwg_taper(in, t[0], make_double2(t[1], t[2]), a);
            }
            double nr = zg_neff(in, sc * sc, nwj).x;
            if (in.zg_model == 0 && nr < 1e-3) nr = 1e-3;
            acc += c_fd8[j] * (f[0] / LH_CLIGHT * nr);
        }
        sm.zs[0] = r;
        sm.zs[1] = acc / in.zfd_h;
        sm.zs[2] = a;
    }
    __syncthreads();
    rho = sm.zs[0];
    const double b1 = sm.zs[1];
    const double a = sm.zs[2];
    const double sqrt_aeff = in.ztaper ? a * in.zsqrt_intg : in.zsqrt_aeff;  // Aeff(z) = a(z)²·aeff_intg
    kfac = rho * LH_EPS0 * in.kerr_gamma3;
    for (int k = tid; k < in.n_omega; k += LH_T) {
        double2 L = make_double2(0.0, 0.0), of = make_double2(in.sc_f, 0.0);
        if (gl(in.zsidx)[k]) {
            const double w = gl(in.st_omega)[k];
            const double sc = sqrt(1 + gl(in.zgam)[k] * rho);  // coren(ω; z)^2
            const double2 nw = in.ztaper ? zg_nwg_taper(in, gl(in.zU)[k], gl(in.zV)[k], a) : static_cast<const double2*>(gl(in.znwg))[k];
            const double2 n = zg_neff(in, sc * sc, nw);
            const double ncr = (n.x < 1e-3) ? 1e-3 : n.x;  // conj_clamp
            const double lim = 3000 * LH_CLIGHT / w;
            const double nci = -((n.y < 0.0) ? 0.0 : (n.y > lim ? lim : n.y));
            const double wc = w / LH_CLIGHT;
            const double zr = wc * ncr - w * b1, zi = wc * nci;
            L = make_double2(zi, -zr);
            const double beta = wc * n.x;
            const double2 pre = gl(in.zpre)[k];
            const double nfx = pre.x / beta * sqrt_aeff, nfy = pre.y / beta * sqrt_aeff;
            of = make_double2((in.sc_f * nfx) * gl(in.omega_win)[k], (in.sc_f * nfy) * gl(in.omega_win)[k]);
        }
        gl(in.zL)[k] = L;
        gl(in.zOF)[k] = of;
    }
    __syncthreads();
}

// ==========================================================================
// Per-step statistics (Stats.default + collect_stats, src/Stats.jl:454-532)
// ==========================================================================
// Maths.linterpx, src/Maths.jl:248-252.
__device__ __forceinline__ double linterpx(double x1, double x2, double y1, double y2, double val) {
    double slope = (y2 - y1) / (x2 - x1);
    double icpt = y1 - slope * x1;
    return (val - icpt) / slope;
}

// One statistics record of the state Y (complex [Nω], HBM) at (z, dz).
//   frequency domain: ω0 = Σω|E|²/Σ|E|² (Maths.moment), energy = Σ w|E|²
//   analytic field (plan_analytic, src/Stats.jl:421-448): field grid
//     Et = ifft([2E_0 .. 2E_{Nω-2}, 0 ..]) over Nt = 2(Nω-1) samples;
//     envelope grid Et = ifft(E).  One inverse FFT of length Nt = MO in LDS.
//   peakpower = max|Et|², peakintensity = peakpower/Aeff,
//   fwhm_t via Maths.level_xings (:linear, src/Maths.jl:162-214),
//   electrondensity (field grid with plasma): rate = PPT(|Re Et|/√(ε0cAeff/2)),
//     ρ·(1 - exp(-cumtrapz(rate, δt)[end])), peak rate = max rate.
// Writes rec[LH_NSTAT] from thread 0; a PPT range/NaN error sets sm.flag.
template <int KIND, int MO, int MC, class Smem>
__device__ __forceinline__ void step_stats(const DInst& in, Smem& sm, const double2* Y, double z, double dz,
                                           double rho, double* rec) {
    constexpr int N = Sizes<KIND, MO, MC>::N;
    constexpr int NT = MO;  // analytic-signal samples = length of grid.t
    const int tid = fresh_tid();
    double s0 = 0.0, s1 = 0.0, se = 0.0;
    WinSums ws = {{0.0, 0.0, 0.0, 0.0}};
    for (int k = tid; k < NT; k += LH_T) {
        double2 v = (k < N) ? Y[k] : make_double2(0.0, 0.0);
        if (k < N) {
            double a = v.x * v.x + v.y * v.y;
            s0 += a;
            s1 += gl(in.st_omega)[k] * a;
            se += gl(in.st_ew)[k] * a;
            win_add(in, ws, k, a);
        }
        if constexpr (KIND == 0)  // copyto_fft!: bins 0..Nω-2 doubled, the +fs/2 bin dropped
            sm.buf[pad(k)] = (k < N - 1) ? make_double2(2.0 * v.x, 2.0 * v.y) : make_double2(0.0, 0.0);
        else
            sm.buf[pad(k)] = v;
    }
    stats_time<KIND, NT, Smem>(in, sm, s0, s1, se, ws, z, dz, rho, rec);
}

// The analytic spectrum is in LDS (not yet synchronised) and each thread holds
// its partial sums Σ|E|², Σω|E|², Σw|E|²: the record of the state.
template <int KIND, int NT, class Smem>
__device__ __forceinline__ void stats_time(const DInst& in, Smem& sm, double s0, double s1, double se,
                                           const WinSums& ws, double z, double dz, double rho, double* rec) {
    const int tid = fresh_tid();
    const int lane = tid & 63, wid = tid >> 6;
    block_sum3(s0, s1, se, sm.wsum);  // ends synchronised: the spectrum is complete
    // Aeff at this record's z: a taper's a(z)²·aeff_intg from the last refresh
    // (at z: the step's final stage, or lh_stats_eval's own), else the constant
    double st_aeff = in.st_aeff, st_ediv = in.st_ediv;
    if (KIND == 0 && in.zg_np > 0 && in.ztaper) {
        const double sa = sm.zs[2] * in.zsqrt_intg;
        st_aeff = sa * sa;
        st_ediv = sqrt(LH_EPS0 * LH_CLIGHT * st_aeff / 2);
    }
    {
        // energy windows: reduced and written here, so the sums do not stay live
        const double nan = __longlong_as_double(0x7ff8000000000000LL);
        double e[LH_STAT_MAX_WINDOWS] = {nan, nan, nan, nan};
        if (in.st_nwin > 0) {
            double w0 = ws.s[0], w1 = ws.s[1], w2 = ws.s[2], w3 = ws.s[3];
            block_sum4(w0, w1, w2, w3, sm.wsum);
            e[0] = w0;
            e[1] = (in.st_nwin > 1) ? w1 : nan;
            e[2] = (in.st_nwin > 2) ? w2 : nan;
            e[3] = (in.st_nwin > 3) ? w3 : nan;
        }
        if (tid == 0) {
#pragma unroll
            for (int q = 0; q < LH_STAT_MAX_WINDOWS; q++) rec[LH_STAT_WINDOW0 + q] = e[q];
        }
    }
    fft_lds<NT, true>(sm.buf, sm.tw);
    // pass 1: P = |Et|² (kept in .x), rate (kept in .y); argmax, max/sum of
    // rate.  Sample j = tid + m·LH_T; rates in batches of QC (spline-table
    // loads issued together).
    constexpr int MPT = NT / LH_T, QC = (MPT < 4) ? MPT : 4;
    static_assert(NT % LH_T == 0 && MPT % QC == 0, "stats sample split");
    const double inv = 1.0 / (double)NT;  // FFTW's ifft normalisation
    const bool ion = (KIND == 0) && in.plasma;
    double pmax = -1.0, rmax = 0.0, rsum = 0.0;
    int imax = NT, bad = LH_OK;
#pragma unroll 1
    for (int c = 0; c < MPT; c += QC) {
        double p[QC], aE[QC], r[QC];
#pragma unroll
        for (int q = 0; q < QC; q++) {
            const double2 e = sm.buf[pad(tid + (c + q) * LH_T)];
            const double re = e.x * inv, im = e.y * inv;
            p[q] = re * re + im * im;
            aE[q] = fabs(re / st_ediv);
            r[q] = 0.0;
        }
        bool need = false;  // any sample at or above the table start (or NaN)
#pragma unroll
        for (int q = 0; q < QC; q++) need |= !(aE[q] < in.ppt_emin);
        if (ion && need) {  // most of the window is below it: those waves skip the table
            ppt_rates<QC>(in, aE, r, bad);
#pragma unroll
            for (int q = 0; q < QC; q++) {
                rmax = fmax(rmax, r[q]);
                rsum += r[q];
            }
        }
#pragma unroll
        for (int q = 0; q < QC; q++) {
            if (p[q] > pmax) {  // j increases within a thread: the first maximum is kept
                pmax = p[q];
                imax = tid + (c + q) * LH_T;
            }
            sm.buf[pad(tid + (c + q) * LH_T)] = make_double2(p[q], r[q]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double p2 = __shfl_xor(pmax, off, 64);
        int i2 = __shfl_xor(imax, off, 64);
        if (p2 > pmax || (p2 == pmax && i2 < imax)) {
            pmax = p2;
            imax = i2;
        }
        rmax = fmax(rmax, __shfl_xor(rmax, off, 64));
        rsum += __shfl_xor(rsum, off, 64);
    }
    if (lane == 0) {
        sm.wsum[4 * wid + 0] = pmax;
        sm.wsum[4 * wid + 1] = (double)imax;
        sm.wsum[4 * wid + 2] = rmax;
        sm.wsum[4 * wid + 3] = rsum;
    }
    if (bad != LH_OK) sm.flag = bad;
    __syncthreads();
    pmax = sm.wsum[0];
    imax = (int)sm.wsum[1];
    rmax = sm.wsum[2];
    rsum = sm.wsum[3];
    for (int w = 1; w < LH_NW; w++) {
        double p2 = sm.wsum[4 * w];
        int i2 = (int)sm.wsum[4 * w + 1];
        if (p2 > pmax || (p2 == pmax && i2 < imax)) {
            pmax = p2;
            imax = i2;
        }
        rmax = fmax(rmax, sm.wsum[4 * w + 2]);
        rsum += sm.wsum[4 * w + 3];
    }
    __syncthreads();
    // pass 2: level crossings around the maximum at val = 0.5·max (t increasing,
    // so x < xmax ⇔ j < imax):  :min takes the last j < imax with P < val and the
    // first j > imax with P < val; :max the first j < imax with P > val and the
    // last j > imax with P > val.
    const double val = 0.5 * pmax;
    int aL = -1, aR = NT, bL = NT, bR = -1;
#pragma unroll
    for (int m = 0; m < MPT; m++) {
        const int j = tid + m * LH_T;
        const double p = sm.buf[pad(j)].x;
        if (j < imax) {
            if (p < val) aL = max(aL, j);
            if (p > val) bL = min(bL, j);
        } else if (j > imax) {
            if (p < val) aR = min(aR, j);
            if (p > val) bR = max(bR, j);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        aL = max(aL, __shfl_xor(aL, off, 64));
        aR = min(aR, __shfl_xor(aR, off, 64));
        bL = min(bL, __shfl_xor(bL, off, 64));
        bR = max(bR, __shfl_xor(bR, off, 64));
    }
    if (lane == 0) {
        sm.wsum[4 * wid + 0] = (double)aL;
        sm.wsum[4 * wid + 1] = (double)aR;
        sm.wsum[4 * wid + 2] = (double)bL;
        sm.wsum[4 * wid + 3] = (double)bR;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 0; w < LH_NW; w++) {
            aL = max(aL, (int)sm.wsum[4 * w + 0]);
            aR = min(aR, (int)sm.wsum[4 * w + 1]);
            bL = min(bL, (int)sm.wsum[4 * w + 2]);
            bR = max(bR, (int)sm.wsum[4 * w + 3]);
        }
        const double* t = gl(in.st_t);
        const double nan = __longlong_as_double(0x7ff8000000000000LL);
        // a missing crossing or a neighbour off the grid is the reference's
        // caught exception: NaN
        double fmin_ = nan, fmax_ = nan;
        if (imax < NT && aL >= 0 && aR < NT) {
            double left = linterpx(t[aL], t[aL + 1], sm.buf[pad(aL)].x, sm.buf[pad(aL + 1)].x, val);
            double right = linterpx(t[aR - 1], t[aR], sm.buf[pad(aR - 1)].x, sm.buf[pad(aR)].x, val);
            fmin_ = fabs(right - left);
        }
        if (imax < NT && bL >= 1 && bL < NT && bR >= 0 && bR + 1 < NT) {
            double left = linterpx(t[bL - 1], t[bL], sm.buf[pad(bL - 1)].x, sm.buf[pad(bL)].x, val);
            double right = linterpx(t[bR], t[bR + 1], sm.buf[pad(bR)].x, sm.buf[pad(bR + 1)].x, val);
            fmax_ = fabs(right - left);
        }
        double ed = nan, pr = nan;
        if (ion) {
            // cumtrapz!(out, δt) end value: Σ ½(r_{j-1} + r_j)·δt
            const double dt = t[1] - t[0];
            double integral = (rsum - 0.5 * (sm.buf[pad(0)].y + sm.buf[pad(NT - 1)].y)) * dt;
            ed = (1.0 - exp(-integral)) * rho;
            pr = rmax;
        }
        rec[LH_STAT_Z] = z;
        rec[LH_STAT_DZ] = dz;
        rec[LH_STAT_OMEGA0] = s1 / s0;
        rec[LH_STAT_ENERGY] = se;
        rec[LH_STAT_PEAKPOWER] = pmax;
        rec[LH_STAT_PEAKINTENSITY] = pmax / st_aeff;
        rec[LH_STAT_FWHM_T_MIN] = fmin_;
        rec[LH_STAT_FWHM_T_MAX] = fmax_;
        rec[LH_STAT_ELECTRONDENSITY] = ed;
        rec[LH_STAT_PEAK_IONISATION_RATE] = pr;
        rec[LH_STAT_DENSITY] = rho;
    }
    __syncthreads();
}

template <int KIND, int MO, int MC, class Smem>
__device__ __forceinline__ void do_prologue(const DInst& in, Smem& sm, const double2* Y, const KSet& K, int stage,
                                            double dt, double h0, double hf) {
    if constexpr (KIND == 0)
        stage_prologue<MO, Smem>(in, sm, Y, K, stage, dt, h0, hf);
    else if constexpr (KIND == 1)
        stage_prologue_env<MO, MC, Smem>(in, sm, Y, K, stage, dt, h0, hf);
    else if constexpr (KIND == 2)
        stage_prologue_big<MC, Smem>(in, sm, Y, K, stage, dt, h0, hf);
    else
        stage_prologue_bigfield<MO, Smem>(in, sm, Y, K, stage, dt, h0, hf);
}

// PACKED: the field-grid buffer already holds the c2r pre-pack (the stage
// prologue writes it); the single-RHS entry fills plain bins instead.
// RAMAN (field grid): compiled with RamanPolarField (the launch picks it for a
// class holding a Raman instance; the rest run without its registers).
template <int KIND, int MO, int MC, class Smem, class Epi, bool PACKED = true, bool RAMAN = false>
__device__ __forceinline__ void do_rhs(const DInst& in, Smem& sm, Epi epi, double rho, double kfac) {
    if constexpr (KIND == 0)
        rhs_field<MO, Smem, Epi, PACKED, RAMAN>(in, sm, epi, rho, kfac);
    else if constexpr (KIND == 1)
        rhs_env<MO, MC, Smem>(in, sm, epi);
    else if constexpr (KIND == 2)
        rhs_big<MO, Smem>(in, sm, epi);
    else
        rhs_bigfield<MO, Smem>(in, sm, epi, rho, kfac);  // W holds the pre-pack in both entries
}

template <int KIND, int MO, int MC, class Smem>
__device__ __forceinline__ void do_stepfun(const DInst& in, Smem& sm, double2* YN, double h, double* rec, double z,
                                           double dz, double rho) {
    if constexpr (KIND == 0)
        stepfun_filter<MC, Smem>(in, sm, YN, h, rec, z, dz, rho);
    else if constexpr (KIND == 1)
        stepfun_env<MC, Smem>(in, sm, YN, h, rec, z, dz, rho);
    else if constexpr (KIND == 2)
        stepfun_big<MC, Smem>(in, sm, YN, h);  // no statistics on this grid
    else
        stepfun_filter_cols<MC, Smem, false>(in, gl(in.linop), sm, YN, h);  // coarse grid in LDS; no statistics
}

// ZDEP: compiled with the z-dependent fill (the launch picks it only for a
// class holding a gradient or taper; the rest run without its registers).
template <int KIND, int MO, int MC, class Smem, bool ZDEP, bool STREAM, bool RAMAN>
__device__ __forceinline__ void solve_instance(const DInst& in, Smem& sm) {
    constexpr int N = Sizes<KIND, MO, MC>::N, NPT = Sizes<KIND, MO, MC>::NPT;
    KSet K{gl(in.K[0]), in.kstride, 0, 6};
    double2* const Y0 = gl(in.Y);
    double2* Y = gl(in.Y);
    double2* YN = gl(in.YN);
    for (int k = fresh_tid(); k < N; k += LH_T) Y[k] = gl(in.Yin)[k];
    if (threadIdx.x == 0) sm.flag = LH_OK;
    __syncthreads();

    long long steps = 0, accepted = 0, rejected = 0, rhs_evals = 0, npub = 0;
    int status = LH_OK, saved = in.save_first, repeated = 0;
    double t = in.z0, tn = in.z0, dt = in.init_dz, dtn = in.init_dz, errlast = 0.0;
    // density and Kerr factor of the current z (constant unless a gradient fill
    // re-evaluates them, with the operator, at every new z)
    double rho = in.density, kfac = in.kerr_fac;
    const bool zdep = ZDEP && (KIND == 0) && in.zg_np > 0;
    if (zdep) zg_refresh(in, sm, in.z0, rho, kfac);

    // PreconStepper init: k1 = fbar(y0, t, t)
    do_prologue<KIND, MO, MC, Smem>(in, sm, Y, K, 0, 0.0, 0.0, 0.0);
    {
        do_rhs<KIND, MO, MC, Smem, EpiStore, true, RAMAN>(in, sm, EpiStore{K[0]}, rho, kfac);
    }
    rhs_evals++;
    __syncthreads();
    if (sm.flag != LH_OK) status = sm.flag;

    while (status == LH_OK && tn <= in.zmax) {
        if (steps >= in.max_steps) {
            status = LH_E_MAXSTEPS;
            break;
        }
        // evaluate!
        double h0 = tn - t;
        dt = dtn;
        t = tn;
        if (zdep) {
            // prop!(ks[1], t, tn) with the operator at tn; each stage then
            // re-evaluates it at its own t2
            zg_refresh(in, sm, t, rho, kfac);
            if (h0 != 0.0) {
                double2* K0 = K[0];
                for (int k = fresh_tid(); k < N; k += LH_T) K0[k] = cmul(K0[k], cexp_lh(gl(in.linop)[k], h0));
                __syncthreads();
            }
            h0 = 0.0;
        }
#pragma unroll 1
        for (int st = 0; st < 6 && status == LH_OK; st++) {
            double t2 = t + c_nodes[st] * dt;
            double hf = t2 - t;
            if (zdep && st != 5) zg_refresh(in, sm, t2, rho, kfac);  // stages 5 and 6 share t + dt
            do_prologue<KIND, MO, MC, Smem>(in, sm, Y, K, st, dt, h0, hf);
            if constexpr (KIND == 0 || KIND == 3)
                do_rhs<KIND, MO, MC, Smem, EpiStageFO, true, RAMAN>(in, sm, EpiStageFO{K[st + 1], K[7]}, rho, kfac);
            else
                do_rhs<KIND, MO, MC, Smem, EpiStageF, true, RAMAN>(in, sm, EpiStageF{K[st + 1], K[7]}, rho, kfac);
            rhs_evals++;
            __syncthreads();
            if (sm.flag != LH_OK) status = sm.flag;
        }
        if (status != LH_OK) break;
        steps++;
        // yn = y + Δz Σ b5 k ; yerr = Σ (Δz k) e ; weaknorm (src/RK45.jl:170-181, 343-354)
        double sy = 0.0, syn = 0.0, se = 0.0;
        if constexpr (NPT > 16) {
            // the big grids: the same sums over chunks of 8 bins
            constexpr int CH = 8;
#pragma unroll 1
            for (int c = 0; c < NPT; c += CH) {
                double2 yv[CH], yn[CH], ye[CH];
#pragma unroll
                for (int i = 0; i < CH; i++) {
                    const int k = fresh_tid() + (c + i) * LH_T;
                    yv[i] = (k < N) ? Y[k] : make_double2(0.0, 0.0);
                    yn[i] = yv[i];
                    ye[i] = make_double2(0.0, 0.0);
                }
#pragma unroll 1
                for (int j = 0; j < 7; j++) {
                    if (j == 1) continue;
                    const double cb = dt * c_b5[j], ce = c_b5[j] - c_b4[j];
                    const double2* Kj = K[j];
#pragma unroll
                    for (int i = 0; i < CH; i++) {
                        const int k = fresh_tid() + (c + i) * LH_T;
                        const double2 kv = (k < N) ? Kj[k] : make_double2(0.0, 0.0);
                        yn[i] = cadd(yn[i], cscal(cb, kv));
                        ye[i] = cadd(ye[i], cscal(ce, cscal(dt, kv)));
                    }
                }
#pragma unroll
                for (int i = 0; i < CH; i++) {
                    const int k = fresh_tid() + (c + i) * LH_T;
                    sy += yv[i].x * yv[i].x + yv[i].y * yv[i].y;
                    syn += yn[i].x * yn[i].x + yn[i].y * yn[i].y;
                    se += ye[i].x * ye[i].x + ye[i].y * ye[i].y;
                    if (k < N) YN[k] = yn[i];
                }
            }
        } else if constexpr (NPT == 9) {
            // chunks of three bins, y and the six k-vectors of each loaded in
            // one round (three rounds instead of seven: 123.9 → 121.4 ms for
            // the Nto = 16384 class; results within 1e-13, steps unchanged)
            constexpr int CB = 3;
#pragma unroll 1
            for (int c = 0; c < NPT; c += CB) {
                double2 yv[CB], kk[7][CB];
#pragma unroll
                for (int i = 0; i < CB; i++) {
                    const int k = fresh_tid() + (c + i) * LH_T;
                    yv[i] = (k < N) ? Y[k] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int j = 0; j < 7; j++) {
                    if (j == 1) continue;
                    const double2* Kj = K[j];
#pragma unroll
                    for (int i = 0; i < CB; i++) {
                        const int k = fresh_tid() + (c + i) * LH_T;
                        kk[j][i] = (k < N) ? Kj[k] : make_double2(0.0, 0.0);
                    }
                }
                double2 yn[CB], ye[CB];
#pragma unroll
                for (int i = 0; i < CB; i++) {
                    yn[i] = yv[i];
                    ye[i] = make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int j = 0; j < 7; j++) {
                    if (j == 1) continue;  // b5[2] = b4[2] = 0
                    const double cb = dt * c_b5[j], ce = c_b5[j] - c_b4[j];
#pragma unroll
                    for (int i = 0; i < CB; i++) {
                        const int k = fresh_tid() + (c + i) * LH_T;
                        if (k < N) {
                            yn[i] = cadd(yn[i], cscal(cb, kk[j][i]));
                            ye[i] = cadd(ye[i], cscal(ce, cscal(dt, kk[j][i])));
                        }
                    }
                }
#pragma unroll
                for (int i = 0; i < CB; i++) {
                    const int k = fresh_tid() + (c + i) * LH_T;
                    if (k < N) {
                        sy += yv[i].x * yv[i].x + yv[i].y * yv[i].y;
                        syn += yn[i].x * yn[i].x + yn[i].y * yn[i].y;
                        se += ye[i].x * ye[i].x + ye[i].y * ye[i].y;
                        YN[k] = yn[i];
                    }
                }
            }
        } else {
            double2 yv[NPT], yn[NPT], ye[NPT];
#pragma unroll
            for (int i = 0; i < NPT; i++) {
                int k = fresh_tid() + i * LH_T;
                yv[i] = (k < N) ? Y[k] : make_double2(0.0, 0.0);
                yn[i] = yv[i];
                ye[i] = make_double2(0.0, 0.0);
            }
#pragma unroll 1
            for (int j = 0; j < 7; j++) {
                if (j == 1) continue;  // b5[2] = b4[2] = 0
                const double cb = dt * c_b5[j], ce = c_b5[j] - c_b4[j];
                const double2* Kj = K[j];
#pragma unroll
                for (int i = 0; i < NPT; i++) {
                    int k = fresh_tid() + i * LH_T;
                    if (k < N) {
                        double2 kv = Kj[k];
                        yn[i] = cadd(yn[i], cscal(cb, kv));
                        ye[i] = cadd(ye[i], cscal(ce, cscal(dt, kv)));
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < NPT; i++) {
                int k = fresh_tid() + i * LH_T;
                if (k < N) {
                    sy += yv[i].x * yv[i].x + yv[i].y * yv[i].y;
                    syn += yn[i].x * yn[i].x + yn[i].y * yn[i].y;
                    se += ye[i].x * ye[i].x + ye[i].y * ye[i].y;
                    YN[k] = yn[i];
                }
            }
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
            K.swap06();  // FSAL: ks[1] .= ks[7] (src/RK45.jl:184-186)
            // stepfun, then MemoryOutput's statistics of the filtered state
            // (src/Output.jl:65-67) from the same LDS pass
            double* rec = (gl(in.steprec) && accepted < in.steprec_cap) ? gl(in.steprec) + (size_t)accepted * LH_NSTAT
                                                                      : nullptr;
            do_stepfun<KIND, MO, MC, Smem>(in, sm, YN, tn - t, rec, tn, dtn, rho);
            if (rec && sm.flag != LH_OK) status = sm.flag;
            // GridCondition saves via dense output; ks[1] and ks[7] both hold k7
            const int saved_before = saved;
            while (saved < in.save_n && gl(in.zgrid)[saved] <= tn) {
                double ti = gl(in.zgrid)[saved];
                if (zdep && gl(in.saves) && ti != t && ti != tn) zg_refresh(in, sm, ti, rho, kfac);  // prop!(out, t, ti)
                if (gl(in.saves)) {
                    double2* out = gl(in.saves) + (size_t)saved * N;
                    if (ti == t) {
                        for (int k = fresh_tid(); k < N; k += LH_T) out[k] = Y[k];
                    } else if (ti == tn) {
                        for (int k = fresh_tid(); k < N; k += LH_T) out[k] = YN[k];
                    } else {
                        double sg = (ti - t) / dt;
                        const double sp1 = sg, sp2 = pow(sg, 2.0), sp3 = pow(sg, 3.0), sp4 = pow(sg, 4.0);
                        double h = ti - t;
                        // this thread's bins in registers (chunks of at most 9),
                        // one k-vector at a time, every load of a chunk before
                        // its first store to out
                        constexpr int CH = (NPT > 9) ? 8 : NPT;  // every access below checks k < N
                        if constexpr (NPT == 9) {
                            // chunks of three bins, y, L and the six k-vectors of
                            // each loaded in one round (three rounds per save
                            // instead of seven: 121.9 → 120.1 ms for the Nto =
                            // 16384 class, bit-identical results)
                            constexpr int CB = 3;
#pragma unroll 1
                            for (int c = 0; c < NPT; c += CB) {
                                double2 kk[7][CB], yv[CB], Lv[CB], yi[CB];
#pragma unroll
                                for (int i = 0; i < CB; i++) {
                                    const int k = fresh_tid() + (c + i) * LH_T;
                                    yv[i] = (k < N) ? Y[k] : make_double2(0.0, 0.0);
                                    Lv[i] = (k < N) ? static_cast<const double2*>(gl(in.linop))[k] : make_double2(0.0, 0.0);
                                    yi[i] = make_double2(0.0, 0.0);
                                }
#pragma unroll
                                for (int j = 0; j < 7; j++) {
                                    if (j == 1) continue;
                                    const double2* kj = (j == 6) ? K[0] : K[j];
#pragma unroll
                                    for (int i = 0; i < CB; i++) {
                                        const int k = fresh_tid() + (c + i) * LH_T;
                                        kk[j][i] = (k < N) ? kj[k] : make_double2(0.0, 0.0);
                                    }
                                }
#pragma unroll
                                for (int j = 0; j < 7; j++) {
                                    if (j == 1) continue;
                                    const double bj =
                                        ((sp1 * c_iC[0][j] + sp2 * c_iC[1][j]) + sp3 * c_iC[2][j]) + sp4 * c_iC[3][j];
#pragma unroll
                                    for (int i = 0; i < CB; i++) yi[i] = cadd(yi[i], cscal(bj, kk[j][i]));
                                }
#pragma unroll
                                for (int i = 0; i < CB; i++) {
                                    const int k = fresh_tid() + (c + i) * LH_T;
                                    if (k < N) out[k] = cmul(cadd(yv[i], cscal(dt, yi[i])), cexp_lh(Lv[i], h));
                                }
                            }
                        } else
#pragma unroll 1
                        for (int c = 0; c < NPT; c += CH) {
                            double2 yi[CH], yv[CH], Lv[CH];
#pragma unroll
                            for (int i = 0; i < CH; i++) yi[i] = make_double2(0.0, 0.0);
#pragma unroll 1
                            for (int j = 0; j < 7; j++) {
                                if (j == 1) continue;
                                const double2* kj = (j == 6) ? K[0] : K[j];
                                // weight b_j(σ) = Σ_p σ^p C[p][j] of the dense output
                                const double bj =
                                    ((sp1 * c_iC[0][j] + sp2 * c_iC[1][j]) + sp3 * c_iC[2][j]) + sp4 * c_iC[3][j];
#pragma unroll
                                for (int i = 0; i < CH; i++) {
                                    const int k = fresh_tid() + (c + i) * LH_T;
                                    if (k < N) yi[i] = cadd(yi[i], cscal(bj, kj[k]));
                                }
                            }
#pragma unroll
                            for (int i = 0; i < CH; i++) {
                                const int k = fresh_tid() + (c + i) * LH_T;
                                yv[i] = (k < N) ? Y[k] : make_double2(0.0, 0.0);
                                Lv[i] = (k < N) ? static_cast<const double2*>(gl(in.linop))[k] : make_double2(0.0, 0.0);
                            }
#pragma unroll
                            for (int i = 0; i < CH; i++) {
                                const int k = fresh_tid() + (c + i) * LH_T;
                                if (k < N) out[k] = cmul(cadd(yv[i], cscal(dt, yi[i])), cexp_lh(Lv[i], h));
                            }
                        }
                    }
                }
                saved++;
            }
            if (STREAM && (saved > saved_before || in.pub_ring > 0)) {
                // the output call of this step (src/Output.jl:65-81, 262-302):
                // its saves are in place; hand over the state, then publish
                const int slot = (in.pub_ring > 0) ? (int)(npub % in.pub_ring) : (int)(npub & 1);
                double2* py = in.pub_y + (size_t)slot * N;
                for (int k = fresh_tid(); k < N; k += LH_T) py[k] = YN[k];
                if (threadIdx.x == 0) {
                    double* info = in.pub_info + 4 * slot;
                    info[0] = tn;
                    info[1] = dtn;
                    info[2] = (double)saved;
                    info[3] = (double)(accepted + 1);
                }
                __threadfence_system();
                __syncthreads();
                if (threadIdx.x == 0) {
                    __hip_atomic_store(in.pub, (unsigned long long)(npub + 1), __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                    if (__hip_atomic_load(in.pub_stop, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM))
                        sm.flag = LH_E_STOPPED;
                }
                __syncthreads();
                if (sm.flag == LH_E_STOPPED && status == LH_OK && tn <= in.zmax) status = LH_E_STOPPED;
                npub++;
            }
            double2* tmp2 = Y;
            Y = YN;
            YN = tmp2;
            accepted++;
            repeated = 0;
            if (in.zstop > 0.0 && tn > in.zstop && tn <= in.zmax) status = LH_E_STOPPED;
        } else {
            rejected++;
            repeated++;
            if (repeated > in.repeat_limit) status = LH_E_REPEAT_LIMIT;
        }
        __syncthreads();
    }
    // The final state s.yn lives in whichever buffer Y points to after the
    // swaps; copy it into the caller-visible buffer Y0.
    if (Y != Y0)
        for (int k = fresh_tid(); k < N; k += LH_T) Y0[k] = Y[k];
    if (threadIdx.x == 0 && gl(in.stats)) {
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

// One persistent kernel per grid size: one workgroup per instance.  STREAM
// builds in the per-step publication of lh_solve_stream (a separate kernel,
// so the others keep their register allocation).
template <int KIND, int MO, int MC, bool ZDEP, bool STREAM = false, bool RAMAN = false>
__global__ __launch_bounds__(LH_T) void lh_solve_kernel(const DInst* insts, const double2* __restrict__ twg) {
    using Smem = SmemT<Sizes<KIND, MO, MC>::LDS>;
    __shared__ Smem sm;
    tw_load(sm.tw, twg);
    __syncthreads();
    solve_instance<KIND, MO, MC, Smem, ZDEP, STREAM, RAMAN>(insts[blockIdx.x], sm);
}

template <int KIND, int MO, int MC, class Smem>
__device__ __forceinline__ void rhs_once(const DInst& in, Smem& sm) {
    constexpr int N = Sizes<KIND, MO, MC>::N;
    double rho = in.density, kfac = in.kerr_fac;
    if (KIND == 0 && in.zg_np > 0) zg_refresh(in, sm, in.z0, rho, kfac);  // the call's z
    for (int k = threadIdx.x; k < N; k += LH_T) {
        double2 v = cscal(in.sc_t, gl(in.rhs_in)[k]);
        if constexpr (KIND == 0)
            sm.buf[pad(k)] = v;
        else if constexpr (KIND == 1)
            sm.buf[pad(env_slot<MO, MC>(k))] = v;
        else if constexpr (KIND == 2)
            gl(in.escr)[MO + k] = v;  // the big grid's working series W
        else
            c2r_pack_bin_big<MO>(gl(in.escr), sm.tw, k, v);  // the big field grid's pre-pack in W
    }
    if constexpr (KIND == 0) {
        for (int k = N + threadIdx.x; k <= MO; k += LH_T) sm.buf[pad(k)] = make_double2(0, 0);
    } else if constexpr (KIND == 1 && MO > MC) {
        for (int k = MC / 2 + threadIdx.x; k < MO - MC / 2; k += LH_T) sm.buf[pad(k)] = make_double2(0, 0);
    }
    double2* out = gl(in.rhs_out);
    if (KIND == 0 && in.raman)
        do_rhs<KIND, MO, MC, Smem, EpiStore, false, true>(in, sm, EpiStore{out}, rho, kfac);
    else
        do_rhs<KIND, MO, MC, Smem, EpiStore, false, false>(in, sm, EpiStore{out}, rho, kfac);
}

template <int KIND, int MO, int MC>
__global__ __launch_bounds__(LH_T) void lh_rhs_kernel(const DInst* insts, const double2* __restrict__ twg,
                                                      int* status) {
    using Smem = SmemT<Sizes<KIND, MO, MC>::LDS>;
    __shared__ Smem sm;
    tw_load(sm.tw, twg);
    if (threadIdx.x == 0) sm.flag = LH_OK;
    __syncthreads();
    rhs_once<KIND, MO, MC, Smem>(insts[blockIdx.x], sm);
    __syncthreads();
    if (threadIdx.x == 0) status[blockIdx.x] = sm.flag;
}

// Statistics of one given state (rhs_in) into steprec[0..LH_NSTAT).
template <int KIND, int MO, int MC>
__global__ __launch_bounds__(LH_T) void lh_stats_kernel(const DInst* insts, const double2* __restrict__ twg,
                                                        double z, double dz, int* status) {
    __shared__ SmemT<MO> sm;
    tw_load(sm.tw, twg);
    if (threadIdx.x == 0) sm.flag = LH_OK;
    __syncthreads();
    const DInst& in = insts[blockIdx.x];
    double rho = in.density, kfac = in.kerr_fac;
    if (KIND == 0 && in.zg_np > 0) zg_refresh(in, sm, z, rho, kfac);
    step_stats<KIND, MO, MC, SmemT<MO>>(in, sm, gl(in.rhs_in), z, dz, rho, gl(in.steprec));
    if (threadIdx.x == 0) status[blockIdx.x] = sm.flag;
}

// Size classes with a compiled kernel.
enum {
    LH_CLASS_16384 = 0,    // field grid, Nto = 16384, Nt = 8192
    LH_CLASS_8192 = 1,     // field grid, Nto = 8192, Nt = 4096
    LH_CLASS_4096 = 2,     // field grid, Nto = 4096, Nt = 2048
    LH_CLASS_ENV8192 = 3,  // envelope grid, Nto = Nt = 8192
    LH_CLASS_ENV4096 = 4,  // envelope grid, Nto = Nt = 4096
    LH_CLASS_RADIAL = 5,   // radial transform (solved context by context)
    LH_CLASS_ENV16384 = 6, // envelope grid, Nto = Nt = 16384, transforms staged through HBM
    LH_CLASS_32768 = 7,    // field grid, Nto = 32768, Nt = 16384, transforms staged through HBM
    LH_NCLASS = 8
};

static int size_class_of(int grid_kind, int n_to, int n_t) {
    if (grid_kind == LH_GRID_REAL) {
        if (n_to == 32768 && n_t == 16384) return LH_CLASS_32768;
        if (n_to == 16384 && n_t == 8192) return LH_CLASS_16384;
        if (n_to == 8192 && n_t == 4096) return LH_CLASS_8192;
        if (n_to == 4096 && n_t == 2048) return LH_CLASS_4096;
    } else if (grid_kind == LH_GRID_ENV) {
        if (n_to == 16384 && n_t == 16384) return LH_CLASS_ENV16384;
        if (n_to == 8192 && n_t == 8192) return LH_CLASS_ENV8192;
        if (n_to == 4096 && n_t == 4096) return LH_CLASS_ENV4096;
    }
    return -1;
}

template <int KIND, int MO, int MC>
static void launch_solve_k(bool zdep, bool raman, int n, const DInst* dinst, const double2* tw, hipStream_t s) {
    if constexpr (KIND == 0) {
        if (zdep) {  // a z-dependent fill never carries Raman (lh_set_gradient refuses it)
            hipLaunchKernelGGL((lh_solve_kernel<KIND, MO, MC, true>), dim3(n), dim3(LH_T), 0, s, dinst, tw);
            return;
        }
        if (raman) {
            hipLaunchKernelGGL((lh_solve_kernel<KIND, MO, MC, false, false, true>), dim3(n), dim3(LH_T), 0, s, dinst,
                               tw);
            return;
        }
    }
    hipLaunchKernelGGL((lh_solve_kernel<KIND, MO, MC, false>), dim3(n), dim3(LH_T), 0, s, dinst, tw);
}

static void launch_solve_stream(int cls, bool raman, const DInst* dinst, const double2* tw, hipStream_t s) {
    if (raman) {
        switch (cls) {
            case LH_CLASS_16384: hipLaunchKernelGGL((lh_solve_kernel<0, 8192, 4096, false, true, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); return;
            case LH_CLASS_8192: hipLaunchKernelGGL((lh_solve_kernel<0, 4096, 2048, false, true, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); return;
            case LH_CLASS_4096: hipLaunchKernelGGL((lh_solve_kernel<0, 2048, 1024, false, true, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); return;
            default: break;  // the envelope classes handle Raman at run time
        }
    }
    switch (cls) {
        case LH_CLASS_16384: hipLaunchKernelGGL((lh_solve_kernel<0, 8192, 4096, false, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); break;
        case LH_CLASS_8192: hipLaunchKernelGGL((lh_solve_kernel<0, 4096, 2048, false, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); break;
        case LH_CLASS_4096: hipLaunchKernelGGL((lh_solve_kernel<0, 2048, 1024, false, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); break;
        case LH_CLASS_ENV8192: hipLaunchKernelGGL((lh_solve_kernel<1, 8192, 8192, false, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); break;
        case LH_CLASS_ENV16384: hipLaunchKernelGGL((lh_solve_kernel<2, 16384, 16384, false, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); break;
        case LH_CLASS_32768: hipLaunchKernelGGL((lh_solve_kernel<3, 16384, 8192, false, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); break;
        default: hipLaunchKernelGGL((lh_solve_kernel<1, 4096, 4096, false, true>), dim3(1), dim3(LH_T), 0, s, dinst, tw); break;
    }
}

// zdep: some instance of the class has a z-dependent fill; raman: some
// instance carries a field-grid Raman response (field grid only, each).
static void launch_solve(int cls, int n, const DInst* dinst, const double2* tw, hipStream_t s, bool zdep,
                         bool raman = false) {
    switch (cls) {
        case LH_CLASS_16384: launch_solve_k<0, 8192, 4096>(zdep, raman, n, dinst, tw, s); break;
        case LH_CLASS_8192: launch_solve_k<0, 4096, 2048>(zdep, raman, n, dinst, tw, s); break;
        case LH_CLASS_4096: launch_solve_k<0, 2048, 1024>(zdep, raman, n, dinst, tw, s); break;
        case LH_CLASS_ENV8192: launch_solve_k<1, 8192, 8192>(false, false, n, dinst, tw, s); break;
        case LH_CLASS_ENV16384: launch_solve_k<2, 16384, 16384>(false, false, n, dinst, tw, s); break;
        case LH_CLASS_32768: launch_solve_k<3, 16384, 8192>(false, false, n, dinst, tw, s); break;
        default: launch_solve_k<1, 4096, 4096>(false, false, n, dinst, tw, s); break;
    }
}

static void launch_rhs(int cls, const DInst* dinst, const double2* tw, int* status, hipStream_t s) {
    switch (cls) {
        case LH_CLASS_16384: hipLaunchKernelGGL((lh_rhs_kernel<0, 8192, 4096>), dim3(1), dim3(LH_T), 0, s, dinst, tw, status); break;
        case LH_CLASS_8192: hipLaunchKernelGGL((lh_rhs_kernel<0, 4096, 2048>), dim3(1), dim3(LH_T), 0, s, dinst, tw, status); break;
        case LH_CLASS_4096: hipLaunchKernelGGL((lh_rhs_kernel<0, 2048, 1024>), dim3(1), dim3(LH_T), 0, s, dinst, tw, status); break;
        case LH_CLASS_ENV8192: hipLaunchKernelGGL((lh_rhs_kernel<1, 8192, 8192>), dim3(1), dim3(LH_T), 0, s, dinst, tw, status); break;
        case LH_CLASS_ENV16384: hipLaunchKernelGGL((lh_rhs_kernel<2, 16384, 16384>), dim3(1), dim3(LH_T), 0, s, dinst, tw, status); break;
        case LH_CLASS_32768: hipLaunchKernelGGL((lh_rhs_kernel<3, 16384, 8192>), dim3(1), dim3(LH_T), 0, s, dinst, tw, status); break;
        default: hipLaunchKernelGGL((lh_rhs_kernel<1, 4096, 4096>), dim3(1), dim3(LH_T), 0, s, dinst, tw, status); break;
    }
}

static void launch_stats(int cls, const DInst* dinst, const double2* tw, double z, double dz, int* status,
                         hipStream_t s) {
    switch (cls) {
        case LH_CLASS_16384: hipLaunchKernelGGL((lh_stats_kernel<0, 8192, 4096>), dim3(1), dim3(LH_T), 0, s, dinst, tw, z, dz, status); break;
        case LH_CLASS_8192: hipLaunchKernelGGL((lh_stats_kernel<0, 4096, 2048>), dim3(1), dim3(LH_T), 0, s, dinst, tw, z, dz, status); break;
        case LH_CLASS_4096: hipLaunchKernelGGL((lh_stats_kernel<0, 2048, 1024>), dim3(1), dim3(LH_T), 0, s, dinst, tw, z, dz, status); break;
        case LH_CLASS_ENV8192: hipLaunchKernelGGL((lh_stats_kernel<1, 8192, 8192>), dim3(1), dim3(LH_T), 0, s, dinst, tw, z, dz, status); break;
        default: hipLaunchKernelGGL((lh_stats_kernel<1, 4096, 4096>), dim3(1), dim3(LH_T), 0, s, dinst, tw, z, dz, status); break;
    }
}

static void launch_raman_hw(int n_to, const double* h, double dt, const double2* tw, double2* hwdt, double2* work,
                            hipStream_t s) {
    if (n_to == 16384) {
        hipLaunchKernelGGL((lh_raman_hw_big_kernel<16384>), dim3(1), dim3(LH_T), 0, s, h, dt, tw, work, hwdt);
        return;
    }
    if (n_to == 8192)
        hipLaunchKernelGGL((lh_raman_hw_kernel<8192>), dim3(1), dim3(LH_T), 0, s, h, dt, tw, hwdt);
    else
        hipLaunchKernelGGL((lh_raman_hw_kernel<4096>), dim3(1), dim3(LH_T), 0, s, h, dt, tw, hwdt);
}

// Radially symmetric transform: per-column kernels and FP64 MFMA contraction.
#include "lh_radial.hip"

// ==========================================================================
// Host side: contexts, constant upload, launches (C ABI, include/luna_hip.h)
// ==========================================================================
static const double EPS0 = 8.854187817620389e-12;  // CODATA 2014 ε0
static const double ECHARGE = 1.6021766208e-19;
static const double MELEC = 9.10938356e-31;
static const double CLIGHT = 299792458.0;

struct lh_ctx {
    int device;
    hipStream_t stream;
    lh_desc d;  // scalars only; host pointers cleared after upload
    int n_state;
    double ppt_emin, ppt_emax;
    double* omega_win;
    double* t_win;
    double* to_win;
    double2* linop;
    double2* outfac;
    double4* ppt;
    double2* hwdt;  // envelope Raman: rfft(h, 2·Nto)·δt
    double2* escr;  // envelope Raman: E(t) stash; big envelope grid: its HBM work area
    size_t escr_n;  // points of escr
    // solver buffers (allocated on first use)
    double2* Yin;
    int have_state;
    int have_result;
    int result_save_n;
    double2* Y;
    double2* YN;
    double2* K[7];  // one allocation: K[j] = K[0] + j·kstride
    long long kstride;
    double2* saves;
    int saves_cap;
    double* zgrid;
    int zgrid_cap;
    double* zgrid_h;  // host copy of the save grid in zgrid (an unchanged grid is not uploaded again)
    int zgrid_hn;
    lh_stats* dstats;
    DInst* dinst;
    double2* rhs_in;
    double2* rhs_out;
    int* dstatus;
    int last_status;
    int rhs_pending;
    hipStream_t rhs_stream;
    // radial transform
    double* Mt;          // transposed QDHT matrix
    double* Ek;          // [Nr][Nto] time series per k column
    double* Er;          // [Nr][Nto] time series per r column
    double* partial;     // [Nr][3] weaknorm partial sums
    DInst* pw;           // pointwise parameters (e_scale = 1)
    RadArgs* rargs;      // [0]: solver, [1]: single RHS
    RadCtl* rctl;        // [0]: solver, [1]: single RHS
    hipGraphExec_t rgraph;
    RadPlan rplan;       // variant of the RHS launch sequence (rgraph is captured with it)
    // per-step statistics
    double* omega;       // [Nω] ω (shared)
    double* st_ew;       // [Nω] energy weights (shared)
    double* st_t;        // [Nt] grid.t (shared)
    double* st_win;      // [st_nwin][Nω] window weights (shared)
    int st_nwin;
    double st_aeff, st_ediv;
    int st_on, st_cap;
    double* steprec;     // [st_cap][LH_NSTAT]
    double* st_rec1;     // [LH_NSTAT], lh_stats_eval
    int result_nrec;     // records kept by the last solve
    int have_records;    // the last solve (lh_batch_solve* or lh_solve_noalloc) ran
    // gradient fill
    int zg_on, zg_np, zg_two, zg_model, zg_loss, zd_n;
    double zsqrt_aeff, zfd_h;
    double *zg_Z, *zg_P, *zfd, *zgam;  // shared
    double4* zd;                       // shared
    double2 *znwg, *zpre;              // shared
    unsigned char* zsidx;              // shared
    double2 *zL, *zOF;                 // per context
    int zg_taper;                      // 1 linear, 2 piecewise-linear taper
    double zg_a0, zg_a1, zg_aL, zg_sqrt_intg;
    int zt_n;
    double *zt_Z, *zt_A;               // shared
    double *zU, *zfdt;                 // shared
    double2* zV;                       // shared
};

static std::mutex g_mu;
static double2* g_tw[64] = {nullptr};

// Working memory of lh_solve_noalloc: one fixed array per GPU, held until the
// process ends.  Its name is a host-side stand-in, not a device address; host
// code uses only the per-device address that hipGetSymbolAddress returns
// (g_scratch, set once per device by lh_create with that device current).
static __device__ __align__(256) unsigned char scratch[LH_NOALLOC_SCRATCH_BYTES];
static unsigned char* g_scratch[64] = {nullptr};
// HIP-event durations of the last resident batch, per size class
static float g_last_ms[8] = {0};
static int g_last_n[8] = {0};

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

// Device address of `scratch` on dev (the current device), looked up once.
static int scratch_init(int dev) {
    std::lock_guard<std::mutex> g(g_mu);
    if (dev < 0 || dev >= 64) return LH_E_ARG;
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIPCK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return LH_E_HIP;
        g_scratch[dev] = (unsigned char*)p;
    }
    return LH_OK;
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

// Constants that many contexts share (a scan's PPT table, the windows of one
// grid) are uploaded once per device and reference-counted: identical content
// maps to one device buffer, which then stays resident in the L2/MALL caches
// for every instance of a batch instead of one private copy each.
// The twiddle table and scratch address of dev as lh_create left them; null
// if no context was ever created there (nothing is allocated here).
static const double2* twiddles_existing(int dev) {
    std::lock_guard<std::mutex> g(g_mu);
    return (dev >= 0 && dev < 64) ? g_tw[dev] : nullptr;
}
static unsigned char* scratch_existing(int dev) {
    std::lock_guard<std::mutex> g(g_mu);
    return (dev >= 0 && dev < 64) ? g_scratch[dev] : nullptr;
}

struct SharedBuf {
    int device;
    uint64_t hash;
    std::vector<unsigned char> host;  // content, for an exact match
    void* dev;
    int refs;
};
static std::vector<SharedBuf*> g_shared;

static uint64_t fnv1a(const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    uint64_t h = 1469598103934665603ULL;
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ULL;
    return h;
}

template <class T>
static int shared_upload(int device, T** dst, const void* src, size_t count) {
    *dst = nullptr;
    if (!src || count == 0) return LH_OK;
    const size_t bytes = sizeof(T) * count;
    const uint64_t h = fnv1a(src, bytes);
    std::lock_guard<std::mutex> g(g_mu);
    for (SharedBuf* e : g_shared)
        if (e->device == device && e->hash == h && e->host.size() == bytes && !memcmp(e->host.data(), src, bytes)) {
            e->refs++;
            *dst = (T*)e->dev;
            return LH_OK;
        }
    void* d = nullptr;
    HIPCK(hipMalloc(&d, bytes));
    if (hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return LH_E_HIP;
    }
    SharedBuf* e = new SharedBuf{device, h, std::vector<unsigned char>((const unsigned char*)src, (const unsigned char*)src + bytes), d, 1};
    g_shared.push_back(e);
    *dst = (T*)d;
    return LH_OK;
}

static void shared_release(const void* dev) {
    if (!dev) return;
    std::lock_guard<std::mutex> g(g_mu);
    for (size_t i = 0; i < g_shared.size(); i++)
        if (g_shared[i]->dev == dev) {
            if (--g_shared[i]->refs == 0) {
                (void)hipFree(g_shared[i]->dev);
                delete g_shared[i];
                g_shared.erase(g_shared.begin() + i);
            }
            return;
        }
}

template <class T>
static int upload(T** dst, const void* src, size_t count) {
    *dst = nullptr;
    if (!src || count == 0) return LH_OK;
    HIPCK(hipMalloc((void**)dst, sizeof(T) * count));
    HIPCK(hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
    return LH_OK;
}


static int ctx_class(const lh_ctx* c);

static bool supported(const lh_desc* d) {
    if (d->transform == LH_TRANS_RADIAL) {
        // Nr a multiple of the 64-wide contraction tile; field grid with Kerr and
        // PPT plasma, or envelope grid (Nt = Nto) with Kerr_env
        if (d->n_r < 64 || d->n_r > 4096 || d->n_r % 64 != 0 || d->raman || d->qdht == nullptr) return false;
        if (d->grid_kind == LH_GRID_ENV)
            return rad_env_size_ok(d->n_to) && d->n_t == d->n_to && d->n_omega == d->n_t && !d->plasma &&
                   d->kerr != LH_KERR_FIELD;
        return d->grid_kind == LH_GRID_REAL && rad_size_ok(d->n_to) && d->n_t == d->n_to / 2 &&
               d->n_omega == d->n_t / 2 + 1 && d->kerr != LH_KERR_ENV;
    }
    if (d->transform != LH_TRANS_MODEAVG || d->n_r != 1) return false;
    if (size_class_of(d->grid_kind, d->n_to, d->n_t) < 0) return false;
    if (d->grid_kind == LH_GRID_REAL) {
        if (d->n_omega != d->n_t / 2 + 1) return false;
        if ((d->raman && !d->raman_h) || d->kerr == LH_KERR_ENV) return false;
    } else {
        if (d->n_omega != d->n_t) return false;
        if (d->plasma || d->kerr == LH_KERR_FIELD) return false;
        if (d->raman && !d->raman_h) return false;
    }
    return true;
}

// to_time!/to_freq! length-ratio scales, src/NonlinearRHS.jl:16-53.
static double scale_to_time(const lh_desc& d) {
    if (d.grid_kind == LH_GRID_REAL) return (double)(d.n_to / 2) / (double)(d.n_omega - 1);
    return (double)d.n_to / (double)d.n_omega;
}
static double scale_to_freq(const lh_desc& d) {
    if (d.grid_kind == LH_GRID_REAL) return (double)(d.n_omega - 1) / (double)(d.n_to / 2);
    return (double)d.n_omega / (double)d.n_to;
}

static int ctx_class(const lh_ctx* c) {
    if (c->d.transform == LH_TRANS_RADIAL) return LH_CLASS_RADIAL;
    return size_class_of(c->d.grid_kind, c->d.n_to, c->d.n_t);
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
    in.density = d.density;
    in.kerr_fac = d.density * EPS0 * d.gamma3;  // (ρ ε0) γ3, Nonlinear.jl:30
    in.kerr34 = 0.75 * in.kerr_fac;             // 3/4·fac, Nonlinear.jl:48
    in.inv_2n = 1.0 / (2.0 * (double)d.n_to);
    in.hwdt = c->hwdt;
    in.escr = c->escr;
    in.ionpot = d.ionpot;
    in.preion_p1 = d.preionfrac + 1;
    in.e_ratio = ECHARGE * ECHARGE / MELEC;
    in.dt_o = d.dt_o;
    in.e_scale = (1.0 / (double)d.n_to) / d.field_div;  // E = c2r/Nto/(nlscale·√Aeff)
    in.inv_nt = 1.0 / (double)d.n_t;
    in.sc_t = scale_to_time(d);
    in.ppt_emin = c->ppt_emin;
    in.ppt_emax = c->ppt_emax;
    in.omega_win = c->omega_win;
    in.t_win = c->t_win;
    in.to_win = c->to_win;
    in.linop = c->linop;
    in.outfac = c->outfac;
    in.ppt = c->ppt;
    in.st_omega = c->omega;
    in.st_ew = c->st_ew;
    in.st_t = c->st_t;
    in.st_aeff = c->st_aeff;
    in.st_ediv = c->st_ediv;
    in.st_win = c->st_win;
    in.st_nwin = c->st_nwin;
    in.steprec = c->st_on ? c->steprec : nullptr;
    in.steprec_cap = c->st_on ? c->st_cap : 0;
    in.sc_f = scale_to_freq(d);
    in.kerr_gamma3 = (d.kerr == LH_KERR_FIELD) ? d.gamma3 : 0.0;
    if (c->zg_on) {
        in.zg_np = c->zg_np;
        in.zg_two = c->zg_two;
        in.zg_model = c->zg_model;
        in.zg_loss = c->zg_loss;
        in.zd_n = c->zd_n;
        in.zsqrt_aeff = c->zsqrt_aeff;
        in.zfd_h = c->zfd_h;
        in.zg_Z = c->zg_Z;
        in.zg_P = c->zg_P;
        in.zd = c->zd;
        in.zgam = c->zgam;
        in.znwg = c->znwg;
        in.zpre = c->zpre;
        in.zsidx = c->zsidx;
        in.zfd = c->zfd;
        in.ztaper = c->zg_taper;
        in.za0 = c->zg_a0;
        in.za1 = c->zg_a1;
        in.zaL = c->zg_aL;
        in.zsqrt_intg = c->zg_sqrt_intg;
        in.zt_n = c->zt_n;
        in.zt_Z = c->zt_Z;
        in.zt_A = c->zt_A;
        in.zU = c->zU;
        in.zV = c->zV;
        in.zfdt = c->zfdt;
        in.zL = c->zL;
        in.zOF = c->zOF;
        in.linop = c->zL;  // refreshed in place at every new z
        in.outfac = c->zOF;
    }
    return in;
}

// Per-context solver buffers sized for n_state and save_n.
static int ensure_solver_buffers(lh_ctx* c, int save_n, bool want_saves) {
    const size_t NS = (size_t)c->n_state;
    if (!c->Y) {
        HIPCK(hipMalloc(&c->Yin, sizeof(double2) * NS));
        HIPCK(hipMalloc(&c->Y, sizeof(double2) * NS));
        HIPCK(hipMalloc(&c->YN, sizeof(double2) * NS));
        // the seven stage vectors in one block, each starting on a 256-byte
        // boundary, and slot 7 for the envelope stage factor exp(L·(t − t2))
        const size_t NSp = (NS + 15) & ~(size_t)15;
        HIPCK(hipMalloc(&c->K[0], sizeof(double2) * NSp * 8));
        for (int j = 1; j < 7; j++) c->K[j] = c->K[0] + j * NSp;
        c->kstride = (long long)NSp;
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
        c->zgrid_hn = 0;
        HIPCK(hipMalloc(&c->zgrid, sizeof(double) * (size_t)zn));
        c->zgrid_cap = zn;
    }
    return LH_OK;
}

// range(save_z0, stop=zmax, length=n) as Julia evaluates it (correctly rounded
// lerp): GridCondition(tmin, tmax, saveN), src/Output.jl:490-496.
static double save_grid_at(const lh_rk_opts* o, int i) {
    const int n = o->save_n;
    const long double a = o->save_z0, b = o->zmax;
    if (n <= 1) return o->save_z0;
    return (double)((a * (long double)(n - 1 - i) + b * (long double)i) / (long double)(n - 1));
}

static void save_grid_into(const lh_rk_opts* o, double* g) {
    const int n = o->save_n > 0 ? o->save_n : 1;
    for (int i = 0; i < n; i++) g[i] = save_grid_at(o, i);
}

static void save_grid(const lh_rk_opts* o, std::vector<double>& g) {
    g.assign(o->save_n > 0 ? o->save_n : 1, o->save_z0);
    save_grid_into(o, g.data());
}

static bool opts_ok(const lh_rk_opts* o) {
    return o->save_n >= 0 && o->repeat_limit >= 0 && o->rtol > 0 && o->zmax >= o->z0 && o->save_first >= 0 &&
           o->save_first <= o->save_n && o->max_steps >= 0 && o->zstop >= 0 && (o->every_step == 0 || o->every_step == 1);
}

static void inst_options(DInst* in, const lh_rk_opts* o) {
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
    in->save_first = o->save_first;
    in->zstop = o->zstop;
    in->max_steps = o->max_steps > 0 ? o->max_steps : LH_MAX_STEPS;
}

static int prepare_instance(lh_ctx* c, const lh_rk_opts* o, bool want_saves, DInst* in, lh_stats* dstats_slot,
                            hipStream_t s) {
    if (!opts_ok(o)) return LH_E_ARG;
    if (!c->have_state) return LH_E_ARG;
    int st = ensure_solver_buffers(c, o->save_n, want_saves);
    if (st) return st;
    std::vector<double> g;
    save_grid(o, g);
    // the kernels only read the grid: a sweep repeated with the same options
    // skips the upload (and its synchronisation) for every instance
    const bool same = c->zgrid_h && c->zgrid_hn == (int)g.size() &&
                      memcmp(c->zgrid_h, g.data(), sizeof(double) * g.size()) == 0;
    if (!same) {
        HIPCK(hipMemcpyAsync(c->zgrid, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, s));
        HIPCK(hipStreamSynchronize(s));  // g is a local buffer
        double* h = (double*)realloc(c->zgrid_h, sizeof(double) * g.size());
        if (h) {
            memcpy(h, g.data(), sizeof(double) * g.size());
            c->zgrid_h = h;
            c->zgrid_hn = (int)g.size();
        } else {
            c->zgrid_hn = 0;
        }
    }
    *in = base_inst(c);
    in->Yin = c->Yin;
    in->Y = c->Y;
    in->YN = c->YN;
    for (int j = 0; j < 7; j++) in->K[j] = c->K[j];
    in->kstride = c->kstride;
    in->saves = want_saves ? c->saves : nullptr;
    in->zgrid = c->zgrid;
    in->stats = dstats_slot;
    inst_options(in, o);
    c->have_result = 0;
    c->have_records = 0;
    c->result_save_n = want_saves ? o->save_n : 0;
    return LH_OK;
}

// --------------------------------------------------------------------------
// Plan wisdom: measured radial plans keyed by (n_to, n_r)
// --------------------------------------------------------------------------
// The role of FFTW's wisdom in the reference (src/Utils.jl:84-108 imports it
// before planning and exports it after): a context created for a shape with
// wisdom takes the recorded plan without measuring; any other takes
// RAD_DEFAULT_PLAN.  Process-wide, guarded by g_wisdom_mu.
struct WisdomEntry {
    int n_to, n_r;
    RadPlan plan;
    double us;
};
static std::mutex g_wisdom_mu;
static std::vector<WisdomEntry> g_wisdom;

static RadPlan wisdom_lookup(int n_to, int n_r) {
    std::lock_guard<std::mutex> lk(g_wisdom_mu);
    for (const WisdomEntry& w : g_wisdom)
        if (w.n_to == n_to && w.n_r == n_r) return w.plan;
    return RAD_DEFAULT_PLAN;
}

static void wisdom_record(int n_to, int n_r, RadPlan p, double us) {
    std::lock_guard<std::mutex> lk(g_wisdom_mu);
    for (WisdomEntry& w : g_wisdom)
        if (w.n_to == n_to && w.n_r == n_r) {
            w.plan = p;
            w.us = us;
            return;
        }
    g_wisdom.push_back({n_to, n_r, p, us});
}

// --------------------------------------------------------------------------
// Radial transform: set-up, one RHS, solve
// --------------------------------------------------------------------------
static int rad_setup(lh_ctx* c, const lh_desc* d) {
    const size_t nr = (size_t)d->n_r, nto = (size_t)d->n_to;
    std::vector<double> mt(nr * nr);
    for (size_t k = 0; k < nr; k++)
        for (size_t n = 0; n < nr; n++) mt[k * nr + n] = d->qdht[n * nr + k];
    int st = upload(&c->Mt, mt.data(), mt.size());
    if (st) return st;
    const size_t col = (d->grid_kind == LH_GRID_ENV) ? 2 * nto : nto;  // complex samples on the envelope grid
    HIPCK(hipMalloc(&c->Ek, sizeof(double) * col * nr));
    HIPCK(hipMalloc(&c->Er, sizeof(double) * col * nr));
    HIPCK(hipMalloc(&c->partial, sizeof(double) * 3 * nr));
    HIPCK(hipMalloc(&c->rargs, sizeof(RadArgs) * 2));
    HIPCK(hipMalloc(&c->rctl, sizeof(RadCtl) * 2));
    c->rplan = wisdom_lookup(d->n_to, d->n_r);
    DInst pw = base_inst(c);
    pw.e_scale = 1.0;  // the 1/Nto of the inverse transform is folded into the k→r contraction
    return upload(&c->pw, &pw, 1);
}

static void rad_ctl_init(RadCtl* h, const lh_rk_opts* o) {
    memset(h, 0, sizeof(*h));
    if (o) {
        h->t = h->tn = o->z0;
        h->dt = h->dtn = o->init_dz;
        h->saved = o->save_first;
    }
    for (int j = 0; j < 8; j++) h->perm[j] = j;
    h->status = LH_OK;
}

static RadArgs rad_args(const lh_ctx* c, int slot) {
    RadArgs a;
    memset(&a, 0, sizeof(a));
    const lh_desc& d = c->d;
    a.n_omega = d.n_omega;
    a.n_t = d.n_t;
    a.n_to = d.n_to;
    a.n_r = d.n_r;
    a.col = (d.grid_kind == LH_GRID_ENV) ? 2 * d.n_to : d.n_to;
    a.sc_t = scale_to_time(d);
    a.gemm_kr = d.qdht_inv / (double)d.n_to;  // ldiv!(Eto, QDHT, Eto) with irfft's 1/Nto
    a.gemm_rk = d.qdht_fwd;                   // mul!(Pto, QDHT, Pto)
    a.linop = c->linop;
    a.outfac = c->outfac;
    a.Mt = c->Mt;
    a.pw = c->pw;
    a.Ek = c->Ek;
    a.Er = c->Er;
    a.partial = c->partial;
    a.ctl = c->rctl + slot;
    return a;
}

// Per-step output of a radial run (lh_solve_stream): the steps are replayed
// one at a time and each accepted step that saves (every accepted step with
// every) is handed to cb with the saves since the previous call and the state,
// copied out of device memory into page-locked host buffers.
struct RadHook {
    lh_output_cb cb;
    void* user;
    bool every;
    double2* y_h;          // [n_state]
    double2* sv_h;         // [save_n][n_state]
    const double* zgrid;   // [save_n] host save grid
};

// Full propagation of one radial context on stream s (buffers and save grid
// prepared by prepare_instance).  The per-step launch sequence is captured
// once per context as a graph and replayed until the control block says done.
static int rad_solve(lh_ctx* c, const lh_rk_opts* o, bool want_saves, hipStream_t s, lh_stats* out,
                     const RadHook* hook = nullptr) {
    const double2* tw;
    int st = twiddles(c->device, &tw);
    if (st) return st;
    const int n_to = c->d.n_to, n_r = c->d.n_r;
    RadArgs a = rad_args(c, 0);
    a.Y[0] = c->Y;
    a.Y[1] = c->YN;
    for (int j = 0; j < 7; j++) a.K[j] = c->K[j];
    a.kstride = c->kstride;
    a.saves = want_saves ? c->saves : nullptr;
    a.zgrid = c->zgrid;
    a.zmax = o->zmax;
    a.rtol = o->rtol;
    a.atol = o->atol;
    a.safety = o->safety;
    a.max_dz = o->max_dz;
    a.min_dz = o->min_dz;
    a.repeat_limit = o->repeat_limit;
    a.save_n = o->save_n;
    a.zstop = o->zstop;
    a.max_steps = o->max_steps > 0 ? o->max_steps : LH_MAX_STEPS;
    RadCtl h;
    rad_ctl_init(&h, o);
    h.rhs_evals = 1;
    HIPCK(hipStreamSynchronize(s));
    HIPCK(hipMemcpy(c->rargs, &a, sizeof(a), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(c->rctl, &h, sizeof(h), hipMemcpyHostToDevice));
    HIPCK(hipMemcpyAsync(c->Y, c->Yin, sizeof(double2) * (size_t)c->n_state, hipMemcpyDeviceToDevice, s));
    rad_rhs_dispatch(c->rplan, c->d.grid_kind == LH_GRID_ENV, n_to, n_r, c->rargs, tw, -1, s);  // PreconStepper init: k1 = fbar(y0)
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(&h, c->rctl, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    if (h.status != LH_OK) {
        h.done = 1;
        HIPCK(hipMemcpy(c->rctl, &h, sizeof(h), hipMemcpyHostToDevice));
    }
    if (!h.done && !c->rgraph) {
        hipGraph_t g = nullptr;
        HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rad_step_dispatch(c->rplan, c->d.grid_kind == LH_GRID_ENV, n_to, n_r, c->rargs, tw, s);
        hipError_t e1 = hipGetLastError();
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e1 != hipSuccess || e2 != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            return LH_E_HIP;
        }
        hipError_t e3 = hipGraphInstantiate(&c->rgraph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e3 != hipSuccess) {
            c->rgraph = nullptr;
            return LH_E_HIP;
        }
    }
    // replay in chunks (one step at a time with a hook); steps after
    // completion return at once
    int chunk = hook ? 1 : 4;
    long long acc_done = 0;
    int saves_done = o->save_first;
    bool stop_req = false;
    const size_t NS = (size_t)c->n_state;
    while (!h.done) {
        for (int i = 0; i < chunk; i++) HIPCK(hipGraphLaunch(c->rgraph, s));
        HIPCK(hipMemcpyAsync(&h, c->rctl, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        if (!hook) {
            if (chunk < 16) chunk *= 2;
            continue;
        }
        if (h.accepted == acc_done) continue;
        acc_done = h.accepted;
        if (!stop_req && (hook->every || h.saved > saves_done)) {
            const int nsv = h.saved - saves_done;
            if (nsv > 0)
                HIPCK(hipMemcpyAsync(hook->sv_h + (size_t)saves_done * NS, c->saves + (size_t)saves_done * NS,
                                     sizeof(double2) * NS * (size_t)nsv, hipMemcpyDeviceToHost, s));
            HIPCK(hipMemcpyAsync(hook->y_h, h.ycur ? c->YN : c->Y, sizeof(double2) * NS, hipMemcpyDeviceToHost, s));
            HIPCK(hipStreamSynchronize(s));
            lh_step_out so;
            memset(&so, 0, sizeof(so));
            so.first_save = saves_done;
            so.n_saves = nsv;
            so.saves = (const double*)(hook->sv_h + (size_t)saves_done * NS);
            so.zsaves = hook->zgrid + saves_done;
            so.z = h.tn;
            so.dz = h.dtn;
            so.y = (const double*)hook->y_h;
            so.accepted = h.accepted;
            if (hook->cb(hook->user, &so) != 0) stop_req = true;
        }
        saves_done = h.saved;
        if (stop_req && !h.done) {
            // nothing more is handed over; the run ends here
            h.status = LH_E_STOPPED;
            h.done = 1;
            HIPCK(hipMemcpy(c->rctl, &h, sizeof(h), hipMemcpyHostToDevice));
        }
    }
    if (h.ycur == 1)  // the final state s.yn lives in YN
        HIPCK(hipMemcpyAsync(c->Y, c->YN, sizeof(double2) * (size_t)c->n_state, hipMemcpyDeviceToDevice, s));
    HIPCK(hipStreamSynchronize(s));
    out->steps = h.steps;
    out->accepted = h.accepted;
    out->rejected = h.rejected;
    out->rhs_evals = h.rhs_evals;
    out->status = h.status;
    out->saved = h.saved;
    out->z = h.tn;
    out->dz = h.dtn;
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
        case LH_E_MAXSTEPS: return "reached the cap on step attempts";
        case LH_E_STOPPED: return "stopped before the end of the propagation (zstop)";
        case LH_E_TOO_LARGE: return "problem exceeds the fixed working memory of lh_solve_noalloc";
        case LH_E_HOST_WAIT: return "the host fell behind the per-step output by more than its ring of slots";
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
    for (const void* p : {(const void*)c->omega_win, (const void*)c->t_win, (const void*)c->to_win, (const void*)c->ppt,
                          (const void*)c->omega, (const void*)c->st_ew, (const void*)c->st_t, (const void*)c->st_win,
                          (const void*)c->zg_Z,
                          (const void*)c->zg_P, (const void*)c->zfd, (const void*)c->zgam, (const void*)c->zd,
                          (const void*)c->znwg, (const void*)c->zpre, (const void*)c->zsidx})
        shared_release(p);
    void* ptrs[] = {c->linop, c->outfac, c->hwdt, c->escr, c->Yin, c->Y, c->YN,
                    c->saves, c->zgrid, c->dstats, c->dinst, c->rhs_in, c->rhs_out, c->dstatus,
                    c->steprec, c->st_rec1, c->zL, c->zOF};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (c->K[0]) hipFree(c->K[0]);  // the block of all seven
    void* rptrs[] = {c->Mt, c->Ek, c->Er, c->partial, c->pw, c->rargs, c->rctl};
    for (void* p : rptrs)
        if (p) hipFree(p);
    if (c->rgraph) hipGraphExecDestroy(c->rgraph);
    free(c->zgrid_h);
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
    if (!st) st = scratch_init(c->device);
    if (!st && hipStreamCreate(&c->stream) != hipSuccess) st = LH_E_HIP;
    const size_t N = (size_t)d->n_omega, NS = (size_t)c->n_state;
    if (!st) st = shared_upload(c->device, &c->omega_win, d->omega_win, N);
    if (!st && d->omega) st = shared_upload(c->device, &c->omega, d->omega, N);
    if (!st && d->transform == LH_TRANS_MODEAVG)
        st = shared_upload(c->device, &c->zsidx, d->sidx, N);
    if (!st) st = shared_upload(c->device, &c->t_win, d->t_win, (size_t)d->n_t);
    if (!st) st = shared_upload(c->device, &c->to_win, d->to_win, (size_t)d->n_to);
    if (!st) {
        // Output factor per bin: to_freq! crop scale, then norm! and ×ωwin on
        // sidx bins (mode-average, src/NonlinearRHS.jl:368-387), or the radial
        // factor on every bin (:473).
        const double sf = scale_to_freq(*d);
        const double2* nf = (const double2*)d->norm;
        std::vector<double2> of(NS);
        for (size_t i = 0; i < NS; i++) {
            size_t k = i % N;
            if (d->transform == LH_TRANS_RADIAL)
                of[i] = make_double2(sf * nf[i].x, sf * nf[i].y);
            else if (d->sidx[k])
                of[i] = make_double2((sf * nf[i].x) * d->omega_win[k], (sf * nf[i].y) * d->omega_win[k]);
            else
                of[i] = make_double2(sf, 0.0);
        }
        st = upload(&c->outfac, of.data(), NS);
    }
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
        st = shared_upload(c->device, &c->ppt, p.data(), p.size());
    }
    const bool big_env = (ctx_class(c) == LH_CLASS_ENV16384);
    const bool big_field = (ctx_class(c) == LH_CLASS_32768);
    const bool field_raman = (d->grid_kind == LH_GRID_REAL && d->transform == LH_TRANS_MODEAVG && d->raman);
    if (!st && big_field && d->raman) st = LH_E_ARG;  // no Raman convolution on the staged field grid
    if (!st && (big_env || big_field || d->raman)) {
        // envelope Raman: E(t) stash of Nto points; the big envelope grid:
        // [E(t) | W | T], Nto + Nto + Nto/2 points, with or without Raman;
        // field-grid Raman: [series | W | P], Nto/2 + Nto + Nto/2 points;
        // the big field grid: [W | T], Nto/2 + Nto/4 points
        c->escr_n = big_env     ? (size_t)d->n_to * 5 / 2
                    : big_field ? (size_t)d->n_to * 3 / 4
                    : field_raman ? (size_t)d->n_to * 2
                                  : (size_t)d->n_to;
        if (hipMalloc(&c->escr, sizeof(double2) * c->escr_n) != hipSuccess) st = LH_E_NOMEM;
    }
    if (!st && d->raman) {
        const double2* tw;
        double* dh = nullptr;
        st = twiddles(c->device, &tw);
        if (!st) st = upload(&dh, d->raman_h, (size_t)d->n_to);
        if (!st && hipMalloc(&c->hwdt, sizeof(double2) * (size_t)(d->n_to + 1)) != hipSuccess) st = LH_E_NOMEM;
        if (!st) {
            double2* work = big_env ? c->escr + d->n_to : field_raman ? c->escr + d->n_to / 2 : nullptr;
            launch_raman_hw(d->n_to, dh, d->dt_o, tw, c->hwdt, work, c->stream);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) st = LH_E_HIP;
        }
        if (dh) (void)hipFree(dh);
    }
    if (!st && d->transform == LH_TRANS_RADIAL) st = rad_setup(c, d);
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

int lh_ctx_status(lh_ctx* c) {
    if (!c) return LH_E_ARG;
    if (c->rhs_pending) {  // status of the last lh_rhs_device launch
        int kst = LH_OK;
        HIPCK(hipSetDevice(c->device));
        HIPCK(hipStreamSynchronize(c->rhs_stream ? c->rhs_stream : c->stream));
        const int* sp = (c->d.transform == LH_TRANS_RADIAL) ? &c->rctl[1].status : c->dstatus;
        HIPCK(hipMemcpy(&kst, sp, sizeof(int), hipMemcpyDeviceToHost));
        c->last_status = kst;
        c->rhs_pending = 0;
    }
    return c->last_status;
}

int lh_rhs_device(lh_ctx* c, void* nl_dev, const void* Ew_dev, double z, void* stream) {
    if (!c || !nl_dev || !Ew_dev) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const double2* tw;
    int st = twiddles(c->device, &tw);
    if (st) return st;
    if (c->d.transform == LH_TRANS_RADIAL) {
        RadArgs a = rad_args(c, 1);
        a.rhs_in = (const double2*)Ew_dev;
        a.rhs_out = (double2*)nl_dev;
        RadCtl h;
        rad_ctl_init(&h, nullptr);
        HIPCK(hipMemcpyAsync(c->rargs + 1, &a, sizeof(a), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(c->rctl + 1, &h, sizeof(h), hipMemcpyHostToDevice, s));
        rad_rhs_dispatch(c->rplan, c->d.grid_kind == LH_GRID_ENV, c->d.n_to, c->d.n_r, c->rargs + 1, tw, -1, s);
    } else {
        DInst in = base_inst(c);
        in.rhs_in = (const double2*)Ew_dev;
        in.rhs_out = (double2*)nl_dev;
        in.z0 = z;  // a gradient fill evaluates its constants at z
        HIPCK(hipMemcpyAsync(c->dinst, &in, sizeof(DInst), hipMemcpyHostToDevice, s));
        launch_rhs(ctx_class(c), c->dinst, tw, c->dstatus, s);
    }
    HIPCK(hipGetLastError());
    c->rhs_pending = 1;
    c->rhs_stream = s;
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
    const int* sp = (c->d.transform == LH_TRANS_RADIAL) ? &c->rctl[1].status : c->dstatus;
    HIPCK(hipMemcpyAsync(&kst, sp, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    c->last_status = kst;
    c->rhs_pending = 0;
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

// lh_solve_stream of a radial context: the host replays the steps (RadHook).
static int rad_solve_stream(lh_ctx* c, double* Ew, const lh_rk_opts* o, lh_output_cb cb, void* user,
                            lh_stats* stats) {
    const size_t NS = (size_t)c->n_state;
    const size_t nsave = (size_t)(o->save_n > 0 ? o->save_n : 0);
    double2* y_h = nullptr;
    double2* sv_h = nullptr;
    HIPCK(hipHostMalloc((void**)&y_h, sizeof(double2) * NS, 0));
    if (nsave && hipHostMalloc((void**)&sv_h, sizeof(double2) * NS * nsave, 0) != hipSuccess) {
        (void)hipHostFree(y_h);
        return LH_E_NOMEM;
    }
    std::vector<double> zg;
    save_grid(o, zg);
    DInst in;
    int st = prepare_instance(c, o, true, &in, c->dstats, c->stream);
    RadHook hk{cb, user, o->every_step != 0, y_h, sv_h, zg.data()};
    if (!st) st = rad_solve(c, o, true, c->stream, stats, &hk);
    if (!st && hipMemcpy(Ew, c->Y, sizeof(double2) * NS, hipMemcpyDeviceToHost) != hipSuccess) st = LH_E_HIP;
    (void)hipHostFree(y_h);
    if (sv_h) (void)hipHostFree(sv_h);
    if (st) return st;
    c->last_status = stats->status;
    c->have_result = 1;
    c->result_save_n = 0;  // the saves went to the callback
    c->result_nrec = 0;
    c->have_records = 1;
    return stats->status;
}

// Page-locked bytes of the every-step ring of state slots (clamped to 64..4096
// slots): 1023 slots of a 4097-bin state.
static constexpr size_t LH_RING_BYTES = (size_t)64 << 20;

int lh_solve_stream(lh_ctx* c, double* Ew, const lh_rk_opts* o, lh_output_cb cb, void* user, lh_stats* stats) {
    if (!c || !Ew || !o || !cb || !stats || !opts_ok(o)) return LH_E_ARG;
    const int cls = ctx_class(c);
    if (cls < 0 || c->zg_on) return LH_E_ARG;
    int st = lh_set_state(c, Ew);
    if (st) return st;
    if (cls == LH_CLASS_RADIAL) return rad_solve_stream(c, Ew, o, cb, user, stats);
    const double2* tw;
    st = twiddles(c->device, &tw);
    if (st) return st;
    // one page-locked, device-mapped, coherent block: publication counter and
    // stop flag, slot info and state slots (two, or the every-step ring), every
    // save and its position, the statistics records
    const size_t NS = (size_t)c->n_state;
    const size_t nsave = (size_t)(o->save_n > 0 ? o->save_n : 0);
    const size_t ncap = (size_t)(c->st_on ? c->st_cap : 0);
    const bool every = o->every_step != 0;
    size_t nslot = 2;
    if (every) {
        nslot = LH_RING_BYTES / (NS * sizeof(double2));
        nslot = nslot < 64 ? 64 : nslot > 4096 ? 4096 : nslot;
    }
    const size_t off_info = 256, off_y = (off_info + nslot * 4 * sizeof(double) + 255) / 256 * 256,
                 off_sv = off_y + nslot * NS * sizeof(double2), off_rec = off_sv + nsave * NS * sizeof(double2),
                 total = off_rec + ncap * LH_NSTAT * sizeof(double);
    unsigned char* hb = nullptr;
    HIPCK(hipHostMalloc((void**)&hb, total, hipHostMallocMapped | hipHostMallocCoherent));
    unsigned char* db = nullptr;
    if (hipHostGetDevicePointer((void**)&db, hb, 0) != hipSuccess) {
        (void)hipHostFree(hb);
        return LH_E_HIP;
    }
    memset(hb, 0, 256);
    unsigned long long* pub = (unsigned long long*)hb;
    int* stop = (int*)(hb + 64);
    const double* info = (const double*)(hb + off_info);
    const double2* ys = (const double2*)(hb + off_y);
    const double2* sv = (const double2*)(hb + off_sv);
    const double* recs = (const double*)(hb + off_rec);
    DInst in;
    st = prepare_instance(c, o, false, &in, c->dstats, c->stream);
    std::vector<double> zg;
    save_grid(o, zg);
    if (!st) {
        in.saves = nsave ? (double2*)(db + off_sv) : nullptr;
        if (ncap) in.steprec = (double*)(db + off_rec);
        in.pub = (unsigned long long*)db;
        in.pub_stop = (const int*)(db + 64);
        in.pub_info = (double*)(db + off_info);
        in.pub_y = (double2*)(db + off_y);
        in.pub_ring = every ? (int)nslot : 0;
        if (hipMemcpyAsync(c->dinst, &in, sizeof(DInst), hipMemcpyHostToDevice, c->stream) != hipSuccess) st = LH_E_HIP;
    }
    if (!st) {
        launch_solve_stream(cls, c->d.grid_kind == LH_GRID_REAL && c->d.raman, c->dinst, tw, c->stream);
        if (hipGetLastError() != hipSuccess) st = LH_E_HIP;
    }
    // Drain the publications while the kernel runs.  A publication's state
    // slot is rewritten two publications later, so a copy of it counts only
    // if the counter has not moved on meanwhile; saves and records have slots
    // of their own.  In every-step mode each publication is handed over in
    // turn from the ring, which the kernel rewrites nslot publications later:
    // the copy of publication q counts if the counter has not passed
    // q + nslot − 2 after it (the kernel writes publication j into its slot
    // only once j publications are out), else the host has fallen behind and
    // the run ends with LH_E_HOST_WAIT.  After the callback asks to stop,
    // nothing more is handed over (the run ends at its next publication).
    std::vector<double2> ycopy(NS);
    unsigned long long done = 0;
    int saves_done = o->save_first;
    long long recs_done = 0;
    bool stop_req = false, lapped = false;
    while (!st) {
        const hipError_t q = hipStreamQuery(c->stream);
        if (q != hipSuccess && q != hipErrorNotReady) {
            st = LH_E_HIP;
            break;
        }
        unsigned long long p = __atomic_load_n(pub, __ATOMIC_ACQUIRE);
        while (every && p > done && !lapped) {
            const unsigned long long q = done + 1;
            const int slot = (int)((q - 1) % nslot);
            const double z = info[4 * slot], dz = info[4 * slot + 1];
            const int nsaved = (int)info[4 * slot + 2];
            const long long nacc = (long long)info[4 * slot + 3];
            memcpy(ycopy.data(), ys + (size_t)slot * NS, NS * sizeof(double2));
            if (__atomic_load_n(pub, __ATOMIC_ACQUIRE) > q + nslot - 2) {
                lapped = true;
                __atomic_store_n(stop, 1, __ATOMIC_RELEASE);
                break;
            }
            const long long nrec = ncap ? (nacc < (long long)ncap ? nacc : (long long)ncap) : 0;
            if (!stop_req) {
                lh_step_out out;
                out.first_save = saves_done;
                out.n_saves = nsaved - saves_done;
                out.saves = (const double*)(sv + (size_t)saves_done * NS);
                out.zsaves = zg.data() + saves_done;
                out.z = z;
                out.dz = dz;
                out.y = (const double*)ycopy.data();
                out.accepted = nacc;
                out.first_record = (int32_t)recs_done;
                out.n_records = (int32_t)(nrec - recs_done);
                out.records = recs + (size_t)recs_done * LH_NSTAT;
                if (cb(user, &out) != 0) {
                    stop_req = true;
                    __atomic_store_n(stop, 1, __ATOMIC_RELEASE);
                }
            }
            done = q;
            saves_done = nsaved;
            recs_done = nrec;
        }
        if (!every && p > done && !stop_req) {
            double z = 0, dz = 0, fsaved = 0, facc = 0;
            for (;;) {
                const int slot = (int)((p - 1) & 1);
                z = info[4 * slot];
                dz = info[4 * slot + 1];
                fsaved = info[4 * slot + 2];
                facc = info[4 * slot + 3];
                memcpy(ycopy.data(), ys + (size_t)slot * NS, NS * sizeof(double2));
                const unsigned long long p2 = __atomic_load_n(pub, __ATOMIC_ACQUIRE);
                if (p2 == p) break;
                p = p2;
            }
            const int nsaved = (int)fsaved;
            const long long nacc = (long long)facc;
            const long long nrec = ncap ? (nacc < (long long)ncap ? nacc : (long long)ncap) : 0;
            lh_step_out out;
            out.first_save = saves_done;
            out.n_saves = nsaved - saves_done;
            out.saves = (const double*)(sv + (size_t)saves_done * NS);
            out.zsaves = zg.data() + saves_done;
            out.z = z;
            out.dz = dz;
            out.y = (const double*)ycopy.data();
            out.accepted = nacc;
            out.first_record = (int32_t)recs_done;
            out.n_records = (int32_t)(nrec - recs_done);
            out.records = recs + (size_t)recs_done * LH_NSTAT;
            if (cb(user, &out) != 0) {
                stop_req = true;
                __atomic_store_n(stop, 1, __ATOMIC_RELEASE);
            }
            done = p;
            saves_done = nsaved;
            recs_done = nrec;
        }
        if (q == hipSuccess) break;  // p was read after the kernel had ended
        struct timespec ts = {0, 20000};
        nanosleep(&ts, nullptr);
    }
    if (!st) {
        if (hipMemcpy(stats, c->dstats, sizeof(lh_stats), hipMemcpyDeviceToHost) != hipSuccess) st = LH_E_HIP;
        if (!st && lapped) stats->status = LH_E_HOST_WAIT;
    } else {
        (void)hipStreamSynchronize(c->stream);  // never free a block the kernel may still write
    }
    if (!st && hipMemcpy(Ew, c->Y, sizeof(double2) * NS, hipMemcpyDeviceToHost) != hipSuccess) st = LH_E_HIP;
    if (!st && ncap && c->steprec &&
        hipMemcpy(c->steprec, recs, sizeof(double) * LH_NSTAT * ncap, hipMemcpyHostToDevice) != hipSuccess)
        st = LH_E_HIP;
    (void)hipHostFree(hb);
    if (st) return st;
    c->last_status = stats->status;
    c->have_result = 1;
    c->result_save_n = 0;  // the saves went to the callback
    c->result_nrec = ncap ? (int)(stats->accepted < (long long)ncap ? stats->accepted : (long long)ncap) : 0;
    c->have_records = 1;
    return stats->status;
}

int lh_set_state(lh_ctx* c, const double* Ew) {
    if (!c || !Ew) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    int st = ensure_solver_buffers(c, 0, false);
    if (st) return st;
    HIPCK(hipMemcpy(c->Yin, Ew, sizeof(double2) * (size_t)c->n_state, hipMemcpyHostToDevice));
    c->have_state = 1;
    return LH_OK;
}

int lh_get_state(lh_ctx* c, double* Ew) {
    if (!c || !Ew || !c->have_result) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    HIPCK(hipMemcpy(Ew, c->Y, sizeof(double2) * (size_t)c->n_state, hipMemcpyDeviceToHost));
    return LH_OK;
}

int lh_get_saves(lh_ctx* c, double* saves, double* zsaves) {
    if (!c || !c->have_result) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    int n = c->result_save_n;
    if (saves && n > 0)
        HIPCK(hipMemcpy(saves, c->saves, sizeof(double2) * (size_t)c->n_state * n, hipMemcpyDeviceToHost));
    if (zsaves && n > 0) HIPCK(hipMemcpy(zsaves, c->zgrid, sizeof(double) * n, hipMemcpyDeviceToHost));
    return LH_OK;
}

// One stream per (device, size class) for the batched launches, created on
// first use and kept for the process (default priority).
static std::mutex g_cls_mu;
static hipStream_t g_cls_stream[64][LH_NCLASS];

static int class_stream(int dev, int c, hipStream_t* out) {
    if (dev < 0 || dev >= 64) return LH_E_ARG;
    std::lock_guard<std::mutex> lk(g_cls_mu);
    if (!g_cls_stream[dev][c] && hipStreamCreate(&g_cls_stream[dev][c]) != hipSuccess) {
        g_cls_stream[dev][c] = nullptr;
        return LH_E_HIP;
    }
    *out = g_cls_stream[dev][c];
    return LH_OK;
}

// The resident batch with options per instance: op[i] for instance i.
static int batch_resident(lh_ctx* const* ctxs, int n, const lh_rk_opts* const* op, int want_saves, lh_stats* stats) {
    if (!ctxs || n <= 0 || !op || !stats) return LH_E_ARG;
    for (int i = 0; i < n; i++)
        if (!op[i]) return LH_E_ARG;
    const int dev = ctxs[0]->device;
    for (int i = 0; i < n; i++)
        if (!ctxs[i] || ctxs[i]->device != dev || ctx_class(ctxs[i]) < 0) return LH_E_ARG;
    for (int i = 0; i < n; i++)  // the same context twice would share buffers
        for (int j = i + 1; j < n; j++)
            if (ctxs[i] == ctxs[j]) return LH_E_ARG;
    HIPCK(hipSetDevice(dev));
    hipStream_t s0 = ctxs[0]->stream;
    const double2* tw;
    int st = twiddles(dev, &tw);
    if (st) return st;
    // group instances by size class; each class is one persistent launch on
    // its own stream so the classes run concurrently
    int count[LH_NCLASS] = {0}, first[LH_NCLASS] = {0};
    for (int i = 0; i < n; i++) count[ctx_class(ctxs[i])]++;
    for (int c = 1; c < LH_NCLASS; c++) first[c] = first[c - 1] + count[c - 1];
    std::vector<int> pos(n);
    {
        int fill[LH_NCLASS] = {0};
        for (int i = 0; i < n; i++) {
            int c = ctx_class(ctxs[i]);
            pos[i] = first[c] + fill[c]++;
        }
    }
    DInst* dinst = nullptr;
    lh_stats* dstats = nullptr;
    HIPCK(hipMalloc(&dinst, sizeof(DInst) * (size_t)n));
    HIPCK(hipMalloc(&dstats, sizeof(lh_stats) * (size_t)n));
    std::vector<DInst> h(n);
    for (int i = 0; i < n && !st; i++)
        st = prepare_instance(ctxs[i], op[i], want_saves != 0, &h[pos[i]], dstats + pos[i], s0);
    if (!st && hipMemcpyAsync(dinst, h.data(), sizeof(DInst) * (size_t)n, hipMemcpyHostToDevice, s0) != hipSuccess)
        st = LH_E_HIP;
    std::vector<lh_stats> rad_stats(n);
    hipEvent_t ready = nullptr, done[LH_NCLASS] = {nullptr}, t0[LH_NCLASS] = {nullptr}, t1[LH_NCLASS] = {nullptr};
    hipStream_t cs[LH_NCLASS] = {nullptr};
    if (!st && hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess) st = LH_E_HIP;
    if (!st && hipEventRecord(ready, s0) != hipSuccess) st = LH_E_HIP;
    for (int c = 0; c < LH_NCLASS && !st; c++) {
        if (!count[c]) continue;
        if (class_stream(dev, c, &cs[c]) != LH_OK || hipEventCreateWithFlags(&done[c], hipEventDisableTiming) != hipSuccess ||
            hipStreamWaitEvent(cs[c], ready, 0) != hipSuccess) {
            st = LH_E_HIP;
            break;
        }
        if (hipEventCreate(&t0[c]) != hipSuccess || hipEventCreate(&t1[c]) != hipSuccess ||
            hipEventRecord(t0[c], cs[c]) != hipSuccess) {
            st = LH_E_HIP;
            break;
        }
        if (c == LH_CLASS_RADIAL) {
            // context by context (each fills the GPU); field classes run meanwhile
            for (int i = 0; i < n && !st; i++)
                if (ctx_class(ctxs[i]) == LH_CLASS_RADIAL)
                    st = rad_solve(ctxs[i], op[i], want_saves != 0, cs[c], &rad_stats[i]);
            if (st) break;
        } else {
            bool zdep = false, raman = false;
            for (int i = 0; i < n; i++)
                if (ctx_class(ctxs[i]) == c) {
                    if (ctxs[i]->zg_on) zdep = true;
                    if (ctxs[i]->d.grid_kind == LH_GRID_REAL && ctxs[i]->d.raman) raman = true;
                }
            if (zdep && raman) {
                // one class holding both z-dependent fills and Raman instances
                // would need two kernels: refused
                st = LH_E_ARG;
                break;
            }
            launch_solve(c, count[c], dinst + first[c], tw, cs[c], zdep, raman);
        }
        if (hipGetLastError() != hipSuccess || hipEventRecord(t1[c], cs[c]) != hipSuccess ||
            hipEventRecord(done[c], cs[c]) != hipSuccess || hipStreamWaitEvent(s0, done[c], 0) != hipSuccess)
            st = LH_E_HIP;
    }
    std::vector<lh_stats> hs(n);
    if (!st && hipMemcpyAsync(hs.data(), dstats, sizeof(lh_stats) * (size_t)n, hipMemcpyDeviceToHost, s0) != hipSuccess)
        st = LH_E_HIP;
    if (hipStreamSynchronize(s0) != hipSuccess && !st) st = LH_E_HIP;
    for (int c = 0; c < LH_NCLASS; c++) {
        float ms = 0.0f;
        if (!st && t0[c] && t1[c] && hipEventElapsedTime(&ms, t0[c], t1[c]) == hipSuccess) {
            g_last_ms[c] = ms;
            g_last_n[c] = count[c];
        } else if (!count[c]) {
            g_last_ms[c] = 0.0f;
            g_last_n[c] = 0;
        }
        if (t0[c]) (void)hipEventDestroy(t0[c]);
        if (t1[c]) (void)hipEventDestroy(t1[c]);
        if (done[c]) (void)hipEventDestroy(done[c]);
    }
    if (ready) (void)hipEventDestroy(ready);
    (void)hipFree(dinst);
    (void)hipFree(dstats);
    if (st) return st;
    for (int i = 0; i < n; i++) {
        stats[i] = (ctx_class(ctxs[i]) == LH_CLASS_RADIAL) ? rad_stats[i] : hs[pos[i]];
        ctxs[i]->last_status = stats[i].status;
        ctxs[i]->have_result = 1;
        ctxs[i]->result_nrec = ctxs[i]->st_on ? (int)(stats[i].accepted < ctxs[i]->st_cap ? stats[i].accepted : ctxs[i]->st_cap) : 0;
        ctxs[i]->have_records = 1;
    }
    return LH_OK;
}

int lh_batch_solve_resident(lh_ctx* const* ctxs, int n, const lh_rk_opts* o, int want_saves, lh_stats* stats) {
    if (!o || n <= 0) return LH_E_ARG;
    std::vector<const lh_rk_opts*> op((size_t)n, o);
    return batch_resident(ctxs, n, op.data(), want_saves, stats);
}

int lh_batch_solve_opts(lh_ctx* const* ctxs, int n, double* const* Ew, const lh_rk_opts* opts, double* const* saves,
                        double* const* zsaves, lh_stats* stats) {
    if (!ctxs || n <= 0 || !Ew || !opts || !stats) return LH_E_ARG;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i] || !Ew[i]) return LH_E_ARG;
        int st = lh_set_state(ctxs[i], Ew[i]);
        if (st) return st;
    }
    int want = 0;
    for (int i = 0; i < n; i++)
        if (saves && saves[i]) want = 1;
    std::vector<const lh_rk_opts*> op((size_t)n);
    for (int i = 0; i < n; i++) op[i] = opts + i;
    int st = batch_resident(ctxs, n, op.data(), want, stats);
    if (st) return st;
    for (int i = 0; i < n && !st; i++) {
        st = lh_get_state(ctxs[i], Ew[i]);
        if (!st && ((saves && saves[i]) || (zsaves && zsaves[i])))
            st = lh_get_saves(ctxs[i], saves ? saves[i] : nullptr, zsaves ? zsaves[i] : nullptr);
    }
    return st;
}

int lh_set_step_stats(lh_ctx* c, const lh_stats_desc* sd) {
    if (!c) return LH_E_ARG;
    if (!sd) {
        c->st_on = 0;
        return LH_OK;
    }
    if (c->d.transform != LH_TRANS_MODEAVG || ctx_class(c) == LH_CLASS_ENV16384 || ctx_class(c) == LH_CLASS_32768)
        return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    if (sd->capacity < 0 || !sd->energy_w || !sd->t || !(sd->aeff > 0) || !c->omega) return LH_E_ARG;
    if (sd->n_windows < 0 || sd->n_windows > LH_STAT_MAX_WINDOWS || (sd->n_windows > 0 && !sd->window_w))
        return LH_E_ARG;
    double* ew = nullptr;
    double* tg = nullptr;
    double* wn = nullptr;
    int st = shared_upload(c->device, &ew, sd->energy_w, (size_t)c->d.n_omega);
    if (!st) st = shared_upload(c->device, &tg, sd->t, (size_t)c->d.n_t);
    if (!st && sd->n_windows > 0)
        st = shared_upload(c->device, &wn, sd->window_w, (size_t)sd->n_windows * (size_t)c->d.n_omega);
    if (!st && !c->st_rec1 && hipMalloc(&c->st_rec1, sizeof(double) * LH_NSTAT) != hipSuccess) st = LH_E_NOMEM;
    if (!st && sd->capacity > c->st_cap) {
        if (c->steprec) (void)hipFree(c->steprec);
        c->steprec = nullptr;
        c->st_cap = 0;
        if (hipMalloc(&c->steprec, sizeof(double) * LH_NSTAT * (size_t)sd->capacity) != hipSuccess) st = LH_E_NOMEM;
    }
    if (st) {
        shared_release(ew);
        shared_release(tg);
        shared_release(wn);
        return st;
    }
    shared_release(c->st_ew);
    shared_release(c->st_t);
    shared_release(c->st_win);
    c->st_ew = ew;
    c->st_t = tg;
    c->st_win = wn;
    c->st_nwin = sd->n_windows;
    c->st_cap = sd->capacity;
    c->st_aeff = sd->aeff;
    c->st_ediv = sqrt(EPS0 * CLIGHT * sd->aeff / 2);  // sqrt(ε_0*c*aeff(z)/2), src/Stats.jl:236
    c->st_on = 1;
    c->result_nrec = 0;
    return LH_OK;
}

static void zg_release(lh_ctx* c) {
    for (const void* p : {(const void*)c->zg_Z, (const void*)c->zg_P, (const void*)c->zfd, (const void*)c->zgam,
                          (const void*)c->zd, (const void*)c->znwg, (const void*)c->zpre, (const void*)c->zU,
                          (const void*)c->zV, (const void*)c->zfdt, (const void*)c->zt_Z, (const void*)c->zt_A})
        shared_release(p);
    c->zt_Z = c->zt_A = nullptr;
    c->zt_n = 0;
    c->zg_Z = c->zg_P = c->zfd = c->zgam = nullptr;
    c->zd = nullptr;
    c->znwg = c->zpre = nullptr;
    c->zU = c->zfdt = nullptr;
    c->zV = nullptr;
    c->zg_taper = 0;
}

int lh_set_gradient(lh_ctx* c, const lh_gradient_desc* g) {
    if (!c) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    if (!g) {
        c->zg_on = 0;
        zg_release(c);
        return LH_OK;
    }
    const lh_desc& d = c->d;
    // (a Raman response depends on the density; with ρ(z) it would change along z)
    if (d.transform != LH_TRANS_MODEAVG || d.grid_kind != LH_GRID_REAL || !c->omega || !c->zsidx || d.raman ||
        ctx_class(c) == LH_CLASS_32768)
        return LH_E_ARG;
    if (g->n_points < 2 || !g->Z || !g->P || g->dens_n < 4 || !g->dens_p || !g->dens_rho || !g->dens_d ||
        !g->gamma || !g->nwg || !g->pre || !g->fd || !(g->fd_h > 0) || !(g->sqrt_aeff > 0) ||
        (g->model != 0 && g->model != 1) || (g->two_point && g->n_points != 2) || g->taper < 0 || g->taper > 2)
        return LH_E_ARG;
    if (g->taper && (!(g->sqrt_aeff_intg > 0) || !g->wgU || !g->wgV || !g->fdt)) return LH_E_ARG;
    if (g->taper == 1 && (!(g->taper_a0 > 0) || !(g->taper_a1 > 0) || !(g->taper_L > 0))) return LH_E_ARG;
    if (g->taper == 2) {
        if (g->taper_n < 2 || !g->taper_Z || !g->taper_A) return LH_E_ARG;
        for (int i = 0; i < g->taper_n; i++)
            if (!(g->taper_A[i] > 0) || (i > 0 && !(g->taper_Z[i - 1] < g->taper_Z[i]))) return LH_E_ARG;
    }
    for (int i = 0; i + 1 < g->n_points; i++)
        if (!(g->Z[i] < g->Z[i + 1])) return LH_E_ARG;
    if (g->two_point && !(g->Z[0] == 0.0)) return LH_E_ARG;
    const size_t N = (size_t)d.n_omega;
    std::vector<double4> spl((size_t)g->dens_n);
    for (int i = 0; i < g->dens_n; i++) spl[i] = make_double4(g->dens_p[i], g->dens_rho[i], g->dens_d[i], 0.0);
    zg_release(c);
    int st = LH_OK;
    if (!st) st = shared_upload(c->device, &c->zg_Z, g->Z, (size_t)g->n_points);
    if (!st) st = shared_upload(c->device, &c->zg_P, g->P, (size_t)g->n_points);
    if (!st) st = shared_upload(c->device, &c->zd, spl.data(), spl.size());
    if (!st) st = shared_upload(c->device, &c->zgam, g->gamma, N);
    if (!st) st = shared_upload(c->device, &c->znwg, g->nwg, N);
    if (!st) st = shared_upload(c->device, &c->zpre, g->pre, N);
    if (!st) st = shared_upload(c->device, &c->zfd, g->fd, (size_t)36);
    if (!st && g->taper) st = shared_upload(c->device, &c->zU, g->wgU, N);
    if (!st && g->taper) st = shared_upload(c->device, &c->zV, g->wgV, N);
    if (!st && g->taper) st = shared_upload(c->device, &c->zfdt, g->fdt, (size_t)27);
    if (!st && g->taper == 2) st = shared_upload(c->device, &c->zt_Z, g->taper_Z, (size_t)g->taper_n);
    if (!st && g->taper == 2) st = shared_upload(c->device, &c->zt_A, g->taper_A, (size_t)g->taper_n);
    if (!st && !c->zL && hipMalloc(&c->zL, sizeof(double2) * N) != hipSuccess) st = LH_E_NOMEM;
    if (!st && !c->zOF && hipMalloc(&c->zOF, sizeof(double2) * N) != hipSuccess) st = LH_E_NOMEM;
    if (st) {
        zg_release(c);
        c->zg_on = 0;
        return st;
    }
    c->zg_np = g->n_points;
    c->zg_two = g->two_point;
    c->zg_model = g->model;
    c->zg_loss = g->loss;
    c->zd_n = g->dens_n;
    c->zsqrt_aeff = g->sqrt_aeff;
    c->zfd_h = g->fd_h;
    c->zg_taper = g->taper;
    c->zg_a0 = g->taper_a0;
    c->zg_a1 = g->taper_a1;
    c->zg_aL = g->taper_L;
    c->zg_sqrt_intg = g->sqrt_aeff_intg;
    c->zt_n = (g->taper == 2) ? g->taper_n : 0;
    c->zg_on = 1;
    return LH_OK;
}

int lh_stats_eval(lh_ctx* c, const double* Ew, double z, double dz, double* rec) {
    if (!c || !Ew || !rec || !c->st_on) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    const size_t bytes = sizeof(double2) * (size_t)c->n_state;
    if (!c->rhs_in) {
        HIPCK(hipMalloc(&c->rhs_in, bytes));
        HIPCK(hipMalloc(&c->rhs_out, bytes));
    }
    const double2* tw;
    int st = twiddles(c->device, &tw);
    if (st) return st;
    DInst in = base_inst(c);
    in.rhs_in = c->rhs_in;
    in.steprec = c->st_rec1;
    HIPCK(hipMemcpyAsync(c->rhs_in, Ew, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->dinst, &in, sizeof(DInst), hipMemcpyHostToDevice, c->stream));
    launch_stats(ctx_class(c), c->dinst, tw, z, dz, c->dstatus, c->stream);
    HIPCK(hipGetLastError());
    int kst = LH_OK;
    HIPCK(hipMemcpyAsync(rec, c->st_rec1, sizeof(double) * LH_NSTAT, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(&kst, c->dstatus, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return kst;
}

int lh_get_step_stats(lh_ctx* c, double* rec, int capacity, int* n) {
    if (!c || !n || capacity < 0 || (!rec && capacity > 0) || !c->have_records) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    int k = c->result_nrec < capacity ? c->result_nrec : capacity;
    if (k > 0) HIPCK(hipMemcpy(rec, c->steprec, sizeof(double) * LH_NSTAT * (size_t)k, hipMemcpyDeviceToHost));
    *n = k;
    return LH_OK;
}

int lh_last_launch_times(double* ms, int* n_inst, int nclass) {
    if (!ms || !n_inst || nclass < LH_NCLASS) return LH_E_ARG;  // LH_NCLASS = 6
    for (int c = 0; c < LH_NCLASS; c++) {
        ms[c] = g_last_ms[c];
        n_inst[c] = g_last_n[c];
    }
    return LH_OK;
}

int lh_host_alloc(size_t bytes, void** ptr) {
    if (!ptr || bytes == 0) return LH_E_ARG;
    *ptr = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LH_E_NODEVICE;
    if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
        *ptr = nullptr;
        return LH_E_NOMEM;
    }
    return LH_OK;
}

int lh_host_free(void* ptr) {
    if (!ptr) return LH_OK;
    return hipHostFree(ptr) == hipSuccess ? LH_OK : LH_E_HIP;
}

int lh_get_plan(const lh_ctx* c, lh_plan* p) {
    if (!c || !p) return LH_E_ARG;
    if (c->d.transform != LH_TRANS_RADIAL) return LH_E_ARG;
    p->fft_radix_plan = c->rplan.fft;
    p->gemm_tiles = c->rplan.mt;
    p->rhs_us = 0.0;
    p->agrees = 1;
    return LH_OK;
}

int lh_set_plan(lh_ctx* c, const lh_plan* p) {
    if (!c || !p || c->d.transform != LH_TRANS_RADIAL) return LH_E_ARG;
    const RadPlan np = {p->fft_radix_plan, p->gemm_tiles};
    if (!rad_plan_ok(np)) return LH_E_ARG;
    if (np.fft != c->rplan.fft || np.mt != c->rplan.mt) {
        HIPCK(hipSetDevice(c->device));
        HIPCK(hipStreamSynchronize(c->stream));
        if (c->rgraph) {  // captured with the old plan
            (void)hipGraphExecDestroy(c->rgraph);
            c->rgraph = nullptr;
        }
        c->rplan = np;
    }
    return LH_OK;
}

// Measures every plan on one RHS of ctx's shape (the field Ew, or zero when
// NULL) and installs the fastest in ctx and in the wisdom.  Each plan runs
// once untimed, then reps times between two events on the context's stream;
// a plan whose RHS differs from the default plan's by more than 1e-12
// (relative L2) is never chosen.
int lh_autotune(lh_ctx* c, const double* Ew, int reps, lh_plan* tried, int cap, int* n_tried) {
    if (!c || reps < 1 || (cap > 0 && !tried) || c->d.transform != LH_TRANS_RADIAL) return LH_E_ARG;
    HIPCK(hipSetDevice(c->device));
    const double2* tw;
    int st = twiddles(c->device, &tw);
    if (st) return st;
    hipStream_t s = c->stream;
    const size_t n = (size_t)c->n_state;
    double2 *din = nullptr, *dout = nullptr;
    if (hipMalloc(&din, sizeof(double2) * n) != hipSuccess) return LH_E_NOMEM;
    if (hipMalloc(&dout, sizeof(double2) * n) != hipSuccess) {
        (void)hipFree(din);
        return LH_E_NOMEM;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<double2> ref(n), out(n);
    RadPlan plans[6];
    int np = 0;
    plans[np++] = RAD_DEFAULT_PLAN;
    for (int f = 0; f < 2; f++)
        for (int mt : {1, 2, 4})
            if (!(f == RAD_DEFAULT_PLAN.fft && mt == RAD_DEFAULT_PLAN.mt)) plans[np++] = {f, mt};
    RadPlan best = RAD_DEFAULT_PLAN;
    double best_us = 0.0;
    st = LH_OK;
    do {
        if (Ew) {
            if (hipMemcpy(din, Ew, sizeof(double2) * n, hipMemcpyHostToDevice) != hipSuccess) { st = LH_E_HIP; break; }
        } else if (hipMemset(din, 0, sizeof(double2) * n) != hipSuccess) { st = LH_E_HIP; break; }
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { st = LH_E_HIP; break; }
        RadArgs a = rad_args(c, 1);
        a.rhs_in = din;
        a.rhs_out = dout;
        RadCtl h;
        rad_ctl_init(&h, nullptr);
        if (hipStreamSynchronize(s) != hipSuccess ||
            hipMemcpy(c->rargs + 1, &a, sizeof(a), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->rctl + 1, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) { st = LH_E_HIP; break; }
        double refn = 0.0;
        for (int i = 0; i < np && !st; i++) {
            const RadPlan p = plans[i];
            rad_rhs_dispatch(p, c->d.grid_kind == LH_GRID_ENV, c->d.n_to, c->d.n_r, c->rargs + 1, tw, -1, s);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) { st = LH_E_HIP; break; }
            if (hipMemcpy(i == 0 ? ref.data() : out.data(), dout, sizeof(double2) * n, hipMemcpyDeviceToHost) !=
                hipSuccess) { st = LH_E_HIP; break; }
            double dn = 0.0;
            if (i == 0) {
                for (size_t k = 0; k < n; k++) refn += ref[k].x * ref[k].x + ref[k].y * ref[k].y;
            } else {
                for (size_t k = 0; k < n; k++) {
                    const double dx = out[k].x - ref[k].x, dy = out[k].y - ref[k].y;
                    dn += dx * dx + dy * dy;
                }
            }
            const bool agrees = !(dn > 1e-24 * refn);
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps; r++) rad_rhs_dispatch(p, c->d.grid_kind == LH_GRID_ENV, c->d.n_to, c->d.n_r, c->rargs + 1, tw, -1, s);
            (void)hipEventRecord(e1, s);
            float ms = 0.0f;
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { st = LH_E_HIP; break; }
            const double us = 1e3 * (double)ms / reps;
            if (i < cap) {
                tried[i].fft_radix_plan = p.fft;
                tried[i].gemm_tiles = p.mt;
                tried[i].rhs_us = us;
                tried[i].agrees = agrees ? 1 : 0;
            }
            if (agrees && (i == 0 || us < best_us)) {
                best = p;
                best_us = us;
            }
        }
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(din);
    (void)hipFree(dout);
    if (st) return st;
    if (n_tried) *n_tried = np;
    lh_plan bp = {best.fft, best.mt, best_us, 1};
    st = lh_set_plan(c, &bp);
    if (!st) wisdom_record(c->d.n_to, c->d.n_r, best, best_us);
    return st;
}

// Wisdom as text: a header line, then one line per shape
// "radial <n_to> <n_r> <fft_radix_plan> <gemm_tiles> <rhs_us>".
int lh_wisdom_export(char* buf, size_t cap, size_t* len) {
    std::string t = "luna-hip-wisdom 1\n";
    {
        std::lock_guard<std::mutex> lk(g_wisdom_mu);
        char line[128];
        for (const WisdomEntry& w : g_wisdom) {
            snprintf(line, sizeof(line), "radial %d %d %d %d %.3f\n", w.n_to, w.n_r, w.plan.fft, w.plan.mt, w.us);
            t += line;
        }
    }
    if (len) *len = t.size() + 1;
    if (!buf) return LH_OK;
    if (cap < t.size() + 1) return LH_E_ARG;
    memcpy(buf, t.c_str(), t.size() + 1);
    return LH_OK;
}

int lh_wisdom_import(const char* text) {
    if (!text) return LH_E_ARG;
    std::vector<WisdomEntry> got;
    const char* p = text;
    int lineno = 0;
    while (*p) {
        const char* e = strchr(p, '\n');
        std::string line = e ? std::string(p, e - p) : std::string(p);
        p = e ? e + 1 : p + line.size();
        if (lineno++ == 0) {
            if (line != "luna-hip-wisdom 1") return LH_E_ARG;
            continue;
        }
        if (line.empty()) continue;
        WisdomEntry w;
        char kind[16];
        int used = 0;
        if (sscanf(line.c_str(), "%15s %d %d %d %d %lf%n", kind, &w.n_to, &w.n_r, &w.plan.fft, &w.plan.mt, &w.us,
                   &used) != 6 || strcmp(kind, "radial") != 0 || (size_t)used != line.size() || !rad_plan_ok(w.plan) ||
            !rad_size_ok(w.n_to) || w.n_r < 64 || w.n_r % 64 != 0)
            return LH_E_ARG;
        got.push_back(w);
    }
    if (lineno == 0) return LH_E_ARG;
    for (const WisdomEntry& w : got) wisdom_record(w.n_to, w.n_r, w.plan, w.us);
    return LH_OK;
}

int lh_wisdom_forget(void) {
    std::lock_guard<std::mutex> lk(g_wisdom_mu);
    g_wisdom.clear();
    return LH_OK;
}

// Offsets of the working memory of one lh_solve_noalloc call inside `scratch`
// (each part 256-byte aligned): instance record, run counters, save grid, the
// second state buffer YN, the stage-vector block (8 slots as in
// ensure_solver_buffers) and the envelope Raman stash.  The first state buffer
// is the caller's Ew_dev.
struct NoallocLayout {
    size_t inst, stats, zgrid, yn, k, escr, total, kstride;
};
static NoallocLayout noalloc_layout(size_t ns, int save_n, size_t n_escr) {
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    NoallocLayout L;
    L.kstride = (ns + 15) & ~(size_t)15;
    L.inst = 0;
    L.stats = up(sizeof(DInst));
    L.zgrid = L.stats + up(sizeof(lh_stats));
    L.yn = L.zgrid + up(sizeof(double) * (size_t)(save_n > 0 ? save_n : 1));
    L.k = L.yn + up(sizeof(double2) * ns);
    L.escr = L.k + sizeof(double2) * L.kstride * 8;
    L.total = L.escr + up(sizeof(double2) * n_escr);
    return L;
}

int lh_solve_noalloc(lh_ctx* c, void* Ew_dev, const lh_rk_opts* o, void* saves_dev, double* zsaves, lh_stats* stats,
                     void* stream) {
    if (!c || !Ew_dev || !o || !stats || !opts_ok(o)) return LH_E_ARG;
    const int cls = ctx_class(c);
    if (c->d.transform != LH_TRANS_MODEAVG || cls < 0 || cls == LH_CLASS_RADIAL) return LH_E_ARG;
    const size_t NS = (size_t)c->n_state;
    const NoallocLayout L = noalloc_layout(NS, o->save_n, c->escr ? c->escr_n : 0);
    if (L.total > LH_NOALLOC_SCRATCH_BYTES) return LH_E_TOO_LARGE;
    HIPCK(hipSetDevice(c->device));
    unsigned char* base = scratch_existing(c->device);
    const double2* tw = twiddles_existing(c->device);
    if (!base || !tw) return LH_E_HIP;  // set by lh_create on this device
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    DInst in = base_inst(c);
    in.Yin = (const double2*)Ew_dev;  // read once into Y, the same buffer
    in.Y = (double2*)Ew_dev;          // receives the final state
    in.YN = (double2*)(base + L.yn);
    for (int j = 0; j < 7; j++) in.K[j] = (double2*)(base + L.k) + (size_t)j * L.kstride;
    in.kstride = (long long)L.kstride;
    if (c->escr) in.escr = (double2*)(base + L.escr);
    in.saves = (double2*)saves_dev;
    in.zgrid = (const double*)(base + L.zgrid);
    in.stats = (lh_stats*)(base + L.stats);
    inst_options(&in, o);
    // Nothing in the working memory needs clearing: the kernel writes every
    // state and stage vector before reading it, and the counters at the end.
    // Host staging lives on this stack frame; each copy is waited for before
    // the frame's buffer is reused or left.
    HIPCK(hipMemcpyAsync(base + L.inst, &in, sizeof(DInst), hipMemcpyHostToDevice, s));
    const int zn = o->save_n > 0 ? o->save_n : 1;
    double chunk[256];
    for (int i0 = 0; i0 < zn; i0 += 256) {
        const int m = (zn - i0 < 256) ? zn - i0 : 256;
        for (int i = 0; i < m; i++) chunk[i] = save_grid_at(o, i0 + i);
        HIPCK(hipMemcpyAsync(base + L.zgrid + sizeof(double) * (size_t)i0, chunk, sizeof(double) * (size_t)m,
                             hipMemcpyHostToDevice, s));
        HIPCK(hipStreamSynchronize(s));
    }
    launch_solve(cls, 1, (const DInst*)(base + L.inst), tw, s, c->zg_on != 0,
                 c->d.grid_kind == LH_GRID_REAL && c->d.raman);
    HIPCK(hipGetLastError());
    lh_stats hs;
    HIPCK(hipMemcpyAsync(&hs, base + L.stats, sizeof(lh_stats), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    if (zsaves && o->save_n > 0) save_grid_into(o, zsaves);
    *stats = hs;
    c->last_status = hs.status;
    // the state and saves went to the caller's buffers, not the context's
    c->have_result = 0;
    c->result_nrec = c->st_on ? (int)(hs.accepted < c->st_cap ? hs.accepted : c->st_cap) : 0;
    c->have_records = 1;
    return hs.status;
}

int lh_batch_solve(lh_ctx* const* ctxs, int n, double* const* Ew, const lh_rk_opts* o, double* const* saves,
                   double* const* zsaves, lh_stats* stats) {
    if (!ctxs || n <= 0 || !Ew || !o || !stats) return LH_E_ARG;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i] || !Ew[i]) return LH_E_ARG;
        int st = lh_set_state(ctxs[i], Ew[i]);
        if (st) return st;
    }
    int want = 0;
    for (int i = 0; i < n; i++)
        if (saves && saves[i]) want = 1;
    int st = lh_batch_solve_resident(ctxs, n, o, want, stats);
    if (st) return st;
    for (int i = 0; i < n && !st; i++) {
        st = lh_get_state(ctxs[i], Ew[i]);
        if (!st && ((saves && saves[i]) || (zsaves && zsaves[i])))
            st = lh_get_saves(ctxs[i], saves ? saves[i] : nullptr, zsaves ? zsaves[i] : nullptr);
    }
    return st;
}

}  // extern "C"
