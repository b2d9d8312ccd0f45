// lh_radial.hip -- radially symmetric propagation (TransRadial) on gfx950.
//
// Included by lh_kernels.hip (shares its device helpers).  The state is
// Eω(ω, k) with Nω rows and Nr columns (column-major, ω fastest).  One RHS
// (src/NonlinearRHS.jl:466-474) is a sequence of launches:
//   A  per k-column: stage prologue, zero-pad, c2r (LDS)      → Ek (t, k)
//   G  FP64 MFMA GEMM, k → r: Er = (K/R)/Nto · Ek · Mᵀ        → Er (t, r)
//   C  per r-column: Kerr + PPT plasma scans, ×towin (LDS)    → Er (in place)
//   G  FP64 MFMA GEMM, r → k: Ek = (R/K) · Er · Mᵀ            → Ek (t, k)
//   E  per k-column: r2c, crop, ×output factor, ×exp(L hb)  → k_{i+1}
// and a step appends a per-column combination with partial weaknorm sums,
// a one-workgroup controller, the per-column window filter and the saves.
// Every kernel reads the step scalars from a device-side control block, so
// one step is a fixed launch sequence that is captured once as a hipGraph and
// replayed; once the run is done every kernel returns immediately.
//
// QDHT (Hankel.jl, after Guizar-Sicairos & Gutiérrez-Vega, JOSA A 21, 53
// (2004)): j_n zeros of J0, S = j_{N+1}, r_n = j_n R/S, k_n = j_n/R, K = S/R,
// M_nm = (2/S) J0(j_n j_m/S) / J1(j_m)²;  r → k:  Y = (R/K) M A,
// k → r:  A = (K/R) M Y  (M² = I to the accuracy of the discrete orthogonality).

typedef double lh_d4 __attribute__((ext_vector_type(4)));

struct RadCtl {
    double t, tn, dt, dtn, errlast, err, h0;
    long long steps, accepted, rejected, rhs_evals;
    int status, repeated, saved, save_from, save_to, ok, done, ycur;
    int perm[8];
};

// Stage vector j of one column: slot perm[j] of the K block, at the column's offset.
struct RadK {
    double2* base;
    long long stride;
    const int* perm;
    size_t off;
    __device__ __forceinline__ gptr<double2> operator[](int j) const { return gl(base + (long long)perm[j] * stride + off); }
};

struct RadArgs {
    int n_omega, n_t, n_to, n_r;
    int col;  // doubles per column of Ek/Er: Nto real samples (field grid), 2·Nto (envelope)
    double sc_t, gemm_kr, gemm_rk;
    const double2* linop;   // [Nr][Nω]
    const double2* outfac;  // [Nr][Nω]
    const double* Mt;       // [Nr][Nr] transposed QDHT matrix: Mt[k*Nr + n] = M[n][k]
    const DInst* pw;        // scalar parameters + windows + PPT table (e_scale = 1)
    double2* Y[2];
    double2* K[7];          // one block: K[j] = K[0] + j·kstride
    long long kstride;
    double* Ek;             // [Nr][Nto] (t fastest)
    double* Er;             // [Nr][Nto]
    double* partial;        // [Nr][3]
    double2* saves;         // [save_n][Nr][Nω] or null
    const double* zgrid;
    RadCtl* ctl;
    double zmax, rtol, atol, safety, max_dz, min_dz;
    int repeat_limit, save_n;
    double zstop;         // > 0: stop after the first accepted step with tn > zstop
    long long max_steps;  // cap on step attempts
    // single RHS mode: input/output columns
    const double2* rhs_in;
    double2* rhs_out;
};

// ---------------------------------------------------------------- A: prologue + c2r
// PLAN selects the radix plan of the column transforms (fft_lds), measured
// per shape by lh_autotune.
template <int MO, int PLAN>
__global__ __launch_bounds__(LH_T) void rad_stage_a(const RadArgs* __restrict__ ap, const double2* __restrict__ twg,
                                                    int stage) {
    const RadArgs& a = *ap;
    const RadCtl& ctl = *gl(a.ctl);
    if (ctl.done) return;
    __shared__ SmemT<MO> sm;
    tw_load(sm.tw, twg);
    const int c = blockIdx.x, N = a.n_omega;
    const size_t off = (size_t)N * c;
    const RadK Kc{gl(a.K[0]), a.kstride, ctl.perm, off};
    if (stage < 0) {  // plain RHS of the input column (PreconStepper init / single RHS)
        const double2* src = (gl(a.rhs_in) ? gl(a.rhs_in) : gl(a.Y[ctl.ycur])) + off;
        for (int k = threadIdx.x; k < N; k += LH_T) sm.buf[pad(k)] = cscal(a.sc_t, src[k]);
        for (int k = N + threadIdx.x; k <= MO; k += LH_T) sm.buf[pad(k)] = make_double2(0.0, 0.0);
    } else {
        const double t = ctl.t, dt = ctl.dt;
        const double t2 = t + c_nodes[stage] * dt;
        const double h0 = (stage == 0) ? ctl.h0 : 0.0;
        stage_prologue_cols<MO, SmemT<MO>, RadK, false>(gl(a.linop) + off, a.sc_t, sm, gl(a.Y[ctl.ycur]) + off, Kc, stage, dt, h0,
                                                 t2 - t);
    }
    __syncthreads();
    c2r_prepack<MO>(sm.buf, sm.tw);
    __syncthreads();
    fft_lds<MO, true, PLAN>(sm.buf, sm.tw);
    double* out = gl(a.Ek) + (size_t)a.col * c;
    for (int j = threadIdx.x; j < MO; j += LH_T) {
        double2 z = sm.buf[pad(j)];
        out[2 * j] = z.x;
        out[2 * j + 1] = z.y;
    }
}

// ---------------------------------------------------------------- G: FP64 MFMA GEMM
// C[M×N] = scale · A[M×K] · B[K×N]; A and C column-major (t fastest, M = Nto),
// B[k][n] = Mt[k*Nr + n] = M[n][k].  The MFMA computes the transposed tile
// Cᵀ = Bᵀ·Aᵀ so that its column index runs along t: each 16-lane group then
// loads and stores 16 consecutive doubles.  v_mfma_f64_16x16x4f64 lane layout
// (measured): operand 1 lane l = X[l%16][l/16], operand 2 lane l = Y[l/16][l%16],
// result register r of lane l = D[4r + l/16][l%16].
// Workgroup: 2×2 waves; a wave owns 32 r-columns × 16·MT t-rows (2 × MT MFMA
// tiles), operands of 16 consecutive k double-buffered in registers.
template <int MT>
__global__ __launch_bounds__(256) void rad_gemm(const RadArgs* __restrict__ ap, int dir) {
    const RadArgs& a = *ap;
    if (gl(a.ctl)->done) return;
    const double* __restrict__ A = dir == 0 ? gl(a.Ek) : gl(a.Er);
    double* __restrict__ C = dir == 0 ? gl(a.Er) : gl(a.Ek);
    const double* __restrict__ B = gl(a.Mt);
    const double scale = dir == 0 ? a.gemm_kr : a.gemm_rk;
    const int M = a.col, K = a.n_r;  // K = N, a multiple of 64; M a multiple of 32·MT
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n0 = blockIdx.y * 64 + (wave & 1) * 32;
    const int m0 = blockIdx.x * (32 * MT) + (wave >> 1) * (16 * MT);
    const int lr = lane & 15, lk = lane >> 4;
    lh_d4 acc[2][MT];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int j = 0; j < MT; j++) acc[t][j] = (lh_d4){0, 0, 0, 0};
    const double* Bp = B + (n0 + lr) + (size_t)K * lk;
    const double* Ap = A + (m0 + lr) + (size_t)M * lk;
    double rb[2][4][2], ra[2][4][MT];
#define RAD_LOAD(buf, kbase)                                          \
    _Pragma("unroll") for (int s4 = 0; s4 < 4; s4++) {                \
        const size_t kb = (size_t)K * ((kbase) + 4 * s4);             \
        const size_t ka = (size_t)M * ((kbase) + 4 * s4);             \
        _Pragma("unroll") for (int t = 0; t < 2; t++) rb[buf][s4][t] = Bp[kb + 16 * t];  \
        _Pragma("unroll") for (int j = 0; j < MT; j++) ra[buf][s4][j] = Ap[ka + 16 * j]; \
    }
#define RAD_MFMA(buf)                                                                              \
    _Pragma("unroll") for (int s4 = 0; s4 < 4; s4++)                                               \
        _Pragma("unroll") for (int t = 0; t < 2; t++)                                              \
            _Pragma("unroll") for (int j = 0; j < MT; j++)                                         \
                acc[t][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(rb[buf][s4][t], ra[buf][s4][j], acc[t][j], 0, 0, 0);
    RAD_LOAD(0, 0)
#pragma unroll 1
    for (int k0 = 0; k0 < K; k0 += 32) {
        RAD_LOAD(1, k0 + 16)
        RAD_MFMA(0)
        if (k0 + 32 < K) {
            RAD_LOAD(0, k0 + 32)
        }
        RAD_MFMA(1)
    }
#undef RAD_LOAD
#undef RAD_MFMA
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int j = 0; j < MT; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int n = n0 + 16 * t + 4 * r + lk;
                const int m = m0 + 16 * j + lr;
                C[m + (size_t)M * n] = scale * acc[t][j][r];
            }
}

// ---------------------------------------------------------------- C: responses per r
template <int MO>
__global__ __launch_bounds__(LH_T) void rad_pointwise(const RadArgs* __restrict__ ap) {
    const RadArgs& a = *ap;
    if (gl(a.ctl)->done) return;
    __shared__ SmemT<MO> sm;
    if (threadIdx.x == 0) sm.flag = LH_OK;
    double* col = gl(a.Er) + (size_t)a.col * blockIdx.x;
    for (int j = threadIdx.x; j < MO; j += LH_T) sm.buf[pad(j)] = make_double2(col[2 * j], col[2 * j + 1]);
    __syncthreads();
    pointwise_field<MO, SmemT<MO>>(*gl(a.pw), sm, gl(a.pw)->density, gl(a.pw)->kerr_fac);
    __syncthreads();
    for (int j = threadIdx.x; j < MO; j += LH_T) {
        double2 z = sm.buf[pad(j)];
        col[2 * j] = z.x;
        col[2 * j + 1] = z.y;
    }
    if (threadIdx.x == 0 && sm.flag != LH_OK) atomicCAS(&gl(a.ctl)->status, LH_OK, sm.flag);
}

// ---------------------------------------------------------------- E: r2c + epilogue
template <int MO, int PLAN>
__global__ __launch_bounds__(LH_T) void rad_stage_e(const RadArgs* __restrict__ ap, const double2* __restrict__ twg,
                                                    int stage) {
    const RadArgs& a = *ap;
    const RadCtl& ctl = *gl(a.ctl);
    if (ctl.done) return;
    __shared__ SmemT<MO> sm;
    tw_load(sm.tw, twg);
    const int c = blockIdx.x, N = a.n_omega;
    const size_t off = (size_t)N * c;
    const double* col = gl(a.Ek) + (size_t)a.col * c;
    for (int j = threadIdx.x; j < MO; j += LH_T) sm.buf[pad(j)] = make_double2(col[2 * j], col[2 * j + 1]);
    __syncthreads();
    fft_lds<MO, false, PLAN>(sm.buf, sm.tw);
    const double2* of = gl(a.outfac) + off;
    if (stage < 0) {
        double2* out = (gl(a.rhs_out) ? gl(a.rhs_out) : gl(a.K[ctl.perm[0]])) + off;
        for (int k = threadIdx.x; k < N; k += LH_T) out[k] = cmul(r2c_bin<MO>(sm.buf, sm.tw, k), of[k]);
    } else {
        const double t = ctl.t, t2 = t + c_nodes[stage] * ctl.dt, hb = t - t2;
        double2* out = gl(a.K[ctl.perm[stage + 1]]) + off;
        const double2* L = gl(a.linop) + off;
        for (int k = threadIdx.x; k < N; k += LH_T)
            out[k] = cmul(cmul(r2c_bin<MO>(sm.buf, sm.tw, k), of[k]), cexp_lh(L[k], hb));
    }
}

// ---------------------------------------------------------------- envelope grid
// TransRadial on an EnvGrid (src/NonlinearRHS.jl:457-459, 466-474): Nω = Nt
// bins per column, one complex transform of NO = Nto points; a column of Ek/Er
// holds Nto complex samples (2·Nto doubles), on which the real QDHT GEMMs act
// as on 2·Nto rows.  State bin k sits at slot k (k < Nt/2) or k + Nto − Nt of
// the oversampled spectrum (copy_scale_both!, src/NonlinearRHS.jl:89-95).
__device__ __forceinline__ int rad_env_slot(int k, int n, int no) { return (k < n / 2) ? k : k + (no - n); }

template <int NO>
__global__ __launch_bounds__(LH_T) void rad_stage_a_env(const RadArgs* __restrict__ ap,
                                                        const double2* __restrict__ twg, int stage) {
    const RadArgs& a = *ap;
    const RadCtl& ctl = *gl(a.ctl);
    if (ctl.done) return;
    __shared__ SmemT<NO> sm;
    tw_load(sm.tw, twg);
    const int c = blockIdx.x, N = a.n_omega;
    const size_t off = (size_t)N * c;
    for (int j = threadIdx.x; j < NO; j += LH_T) sm.buf[pad(j)] = make_double2(0.0, 0.0);
    __syncthreads();
    if (stage < 0) {  // plain RHS of the input column
        const double2* src = (gl(a.rhs_in) ? gl(a.rhs_in) : gl(a.Y[ctl.ycur])) + off;
        for (int k = threadIdx.x; k < N; k += LH_T) sm.buf[pad(rad_env_slot(k, N, NO))] = cscal(a.sc_t, src[k]);
    } else {
        // yb = y + dt·Σ B[stage][j]·k_j, then exp(L·(t2 − t)); at stage 0 k1 is
        // first carried to t by exp(L·h0) (src/RK45.jl:209-223, 271-281)
        const RadK Kc{gl(a.K[0]), a.kstride, ctl.perm, off};
        const double dt = ctl.dt, hf = c_nodes[stage] * dt;
        const double h0 = (stage == 0) ? ctl.h0 : 0.0;
        const double2* Y = gl(a.Y[ctl.ycur]) + off;
        const double2* L = gl(a.linop) + off;
        for (int k = threadIdx.x; k < N; k += LH_T) {
            const double2 Lk = L[k];
            if (h0 != 0.0) Kc[0][k] = cmul(Kc[0][k], cexp_lh(Lk, h0));
            double2 yb = Y[k];
            for (int j = 0; j <= stage; j++) {
                const double cj = dt * c_B[stage][j];
                if (cj != 0.0) yb = cadd(yb, cscal(cj, Kc[j][k]));
            }
            if (hf != 0.0) yb = cmul(yb, cexp_lh(Lk, hf));
            sm.buf[pad(rad_env_slot(k, N, NO))] = cscal(a.sc_t, yb);
        }
    }
    __syncthreads();
    fft_lds<NO, true>(sm.buf, sm.tw);
    double2* out = reinterpret_cast<double2*>(gl(a.Ek) + (size_t)a.col * c);
    for (int j = threadIdx.x; j < NO; j += LH_T) out[j] = sm.buf[pad(j)];
}

// Kerr_env (src/Nonlinear.jl:47-49): P = ¾ρε0γ3|E|²E, ×towin, in place.
__global__ __launch_bounds__(LH_T) void rad_pointwise_env(const RadArgs* __restrict__ ap) {
    const RadArgs& a = *ap;
    if (gl(a.ctl)->done) return;
    const DInst& in = *gl(a.pw);
    double2* col = reinterpret_cast<double2*>(gl(a.Er) + (size_t)a.col * blockIdx.x);
    const double k34 = in.kerr34;
    const bool kerr = (in.kerr == LH_KERR_ENV);
    const double* tw = gl(in.to_win);
    for (int j = threadIdx.x; j < a.n_to; j += LH_T) {
        const double2 e = col[j];
        double2 p = make_double2(0.0, 0.0);
        if (kerr) {
            const double s = k34 * (e.x * e.x + e.y * e.y);
            p = make_double2(s * e.x, s * e.y);
        }
        col[j] = make_double2(p.x * tw[j], p.y * tw[j]);
    }
}

template <int NO>
__global__ __launch_bounds__(LH_T) void rad_stage_e_env(const RadArgs* __restrict__ ap,
                                                        const double2* __restrict__ twg, int stage) {
    const RadArgs& a = *ap;
    const RadCtl& ctl = *gl(a.ctl);
    if (ctl.done) return;
    __shared__ SmemT<NO> sm;
    tw_load(sm.tw, twg);
    const int c = blockIdx.x, N = a.n_omega;
    const size_t off = (size_t)N * c;
    const double2* col = reinterpret_cast<const double2*>(gl(a.Ek) + (size_t)a.col * c);
    for (int j = threadIdx.x; j < NO; j += LH_T) sm.buf[pad(j)] = col[j];
    __syncthreads();
    fft_lds<NO, false>(sm.buf, sm.tw);
    const double2* of = gl(a.outfac) + off;
    if (stage < 0) {
        double2* out = (gl(a.rhs_out) ? gl(a.rhs_out) : gl(a.K[ctl.perm[0]])) + off;
        for (int k = threadIdx.x; k < N; k += LH_T) out[k] = cmul(sm.buf[pad(rad_env_slot(k, N, NO))], of[k]);
    } else {
        const double t = ctl.t, t2 = t + c_nodes[stage] * ctl.dt, hb = t - t2;
        double2* out = gl(a.K[ctl.perm[stage + 1]]) + off;
        const double2* L = gl(a.linop) + off;
        for (int k = threadIdx.x; k < N; k += LH_T)
            out[k] = cmul(cmul(sm.buf[pad(rad_env_slot(k, N, NO))], of[k]), cexp_lh(L[k], hb));
    }
}

// stepfun per column on the envelope grid (src/Luna.jl:366-370 with a c2c
// FT): YN ← fft(twin · ifft(ωwin · YN·exp(L h))), NT = Nt points.
template <int NT>
__global__ __launch_bounds__(LH_T) void rad_stepfun_env(const RadArgs* __restrict__ ap,
                                                        const double2* __restrict__ twg) {
    const RadArgs& a = *ap;
    const RadCtl& ctl = *gl(a.ctl);
    if (ctl.done || !ctl.ok) return;
    __shared__ SmemT<NT> sm;
    tw_load(sm.tw, twg);
    const DInst& in = *gl(a.pw);
    const size_t off = (size_t)a.n_omega * blockIdx.x;
    double2* YN = gl(a.Y[ctl.ycur ^ 1]) + off;
    const double2* L = gl(a.linop) + off;
    const double h = ctl.tn - ctl.t;
    for (int k = threadIdx.x; k < NT; k += LH_T) {
        double2 v = YN[k];
        if (h != 0.0) v = cmul(v, cexp_lh(L[k], h));
        sm.buf[pad(k)] = cscal(gl(in.omega_win)[k], v);
    }
    __syncthreads();
    fft_lds<NT, true>(sm.buf, sm.tw);
    const double inv_nt = in.inv_nt;
    for (int j = threadIdx.x; j < NT; j += LH_T) {
        const double2 z = sm.buf[pad(j)];
        const double w = gl(in.t_win)[j];
        sm.buf[pad(j)] = make_double2((inv_nt * z.x) * w, (inv_nt * z.y) * w);
    }
    __syncthreads();
    fft_lds<NT, false>(sm.buf, sm.tw);
    for (int k = threadIdx.x; k < NT; k += LH_T) YN[k] = sm.buf[pad(k)];
}

// ---------------------------------------------------------------- combination
__global__ __launch_bounds__(LH_T) void rad_combine(const RadArgs* __restrict__ ap) {
    const RadArgs& a = *ap;
    const RadCtl& ctl = *gl(a.ctl);
    if (ctl.done) return;
    __shared__ double scratch[3 * LH_NW];
    const int c = blockIdx.x, N = a.n_omega;
    const size_t off = (size_t)N * c;
    const double dt = ctl.dt;
    const double2* Y = gl(a.Y[ctl.ycur]) + off;
    double2* YN = gl(a.Y[ctl.ycur ^ 1]) + off;
    double sy = 0.0, syn = 0.0, se = 0.0;
    for (int k = threadIdx.x; k < N; k += LH_T) {
        double2 yv = Y[k], yn = yv, ye = make_double2(0.0, 0.0);
        for (int j = 0; j < 7; j++) {
            if (j == 1) continue;
            double2 kv = gl(a.K[ctl.perm[j]])[off + k];
            yn = cadd(yn, cscal(dt * c_b5[j], kv));
            ye = cadd(ye, cscal(c_b5[j] - c_b4[j], cscal(dt, kv)));
        }
        sy += yv.x * yv.x + yv.y * yv.y;
        syn += yn.x * yn.x + yn.y * yn.y;
        se += ye.x * ye.x + ye.y * ye.y;
        YN[k] = yn;
    }
    block_sum3(sy, syn, se, scratch);
    if (threadIdx.x == 0) {
        gl(a.partial)[3 * c + 0] = sy;
        gl(a.partial)[3 * c + 1] = syn;
        gl(a.partial)[3 * c + 2] = se;
    }
}

// ---------------------------------------------------------------- controller
// weaknorm over all columns (fixed order), stepcontrolPI!, steplims!, FSAL,
// save bookkeeping (src/RK45.jl:167-192, 343-406; src/Output.jl:497-502).
// One wave of 64 lanes: lane l sums columns l, l+64, ... in order, then a
// fixed xor tree (deterministic); lane 0 runs the scalar controller.
__global__ __launch_bounds__(64) void rad_control(const RadArgs* __restrict__ ap) {
    const RadArgs& a = *ap;
    RadCtl* ctl = gl(a.ctl);
    if (ctl->done) return;
    double sy = 0.0, syn = 0.0, se = 0.0;
    for (int c = threadIdx.x; c < a.n_r; c += 64) {
        sy += gl(a.partial)[3 * c + 0];
        syn += gl(a.partial)[3 * c + 1];
        se += gl(a.partial)[3 * c + 2];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sy += __shfl_xor(sy, off, 64);
        syn += __shfl_xor(syn, off, 64);
        se += __shfl_xor(se, off, 64);
    }
    if (threadIdx.x != 0) return;
    RadCtl s = *ctl;
    s.rhs_evals += 6;
    if (s.status != LH_OK) {
        s.done = 1;
        *ctl = s;
        return;
    }
    double errwt = fmax(fmax(sqrt(sy), sqrt(syn)), a.atol);
    double err = sqrt(se) / a.rtol / errwt;
    bool ok = err <= 1.0;
    const double beta1 = 3.0 / 5 / 5, beta2 = -1.0 / 5 / 5, eps = 0.8;
    double dtn;
    if (ok) {
        if (s.errlast == 0.0) s.errlast = err;
        double fac = (err == 0.0) ? 1.5 : a.safety * pow(eps / err, beta1) * pow(eps / s.errlast, beta2);
        dtn = fac * s.dt;
        s.errlast = err;
    } else {
        dtn = !isfinite(err) ? s.dt / 2 : s.dt * fmax(0.1, a.safety * pow(err, -1.0 / 5));
    }
    if (dtn > a.max_dz)
        dtn = a.max_dz;
    else if (dtn < a.min_dz) {
        dtn = a.min_dz;
        ok = true;
    }
    s.err = err;
    s.dtn = dtn;
    s.steps++;
    s.ok = ok;
    if (ok) {
        s.tn = s.t + s.dt;
        int tmp = s.perm[0];  // ks[1] .= ks[7]
        s.perm[0] = s.perm[6];
        s.perm[6] = tmp;
        s.accepted++;
        s.repeated = 0;
        s.save_from = s.saved;
        while (s.saved < a.save_n && gl(a.zgrid)[s.saved] <= s.tn) s.saved++;
        s.save_to = s.saved;
    } else {
        s.rejected++;
        s.repeated++;
        s.save_from = s.save_to = s.saved;
        if (s.repeated > a.repeat_limit) {
            s.status = LH_E_REPEAT_LIMIT;
            s.done = 1;
        }
    }
    *ctl = s;
}

// ---------------------------------------------------------------- stepfun per column
template <int MC>
__global__ __launch_bounds__(LH_T) void rad_stepfun(const RadArgs* __restrict__ ap, const double2* __restrict__ twg) {
    const RadArgs& a = *ap;
    const RadCtl& ctl = *gl(a.ctl);
    if (ctl.done || !ctl.ok) return;
    __shared__ SmemT<MC> sm;
    tw_load(sm.tw, twg);
    __syncthreads();
    const size_t off = (size_t)a.n_omega * blockIdx.x;
    stepfun_filter_cols<MC>(*gl(a.pw), gl(a.linop) + off, sm, gl(a.Y[ctl.ycur ^ 1]) + off, ctl.tn - ctl.t);
}

// ---------------------------------------------------------------- saves per column
__global__ __launch_bounds__(LH_T) void rad_saves(const RadArgs* __restrict__ ap) {
    const RadArgs& a = *ap;
    const RadCtl& ctl = *gl(a.ctl);
    if (ctl.done || !ctl.ok || !gl(a.saves)) return;
    const int c = blockIdx.x, N = a.n_omega;
    const size_t off = (size_t)N * c, nst = (size_t)N * a.n_r;
    const double2* Y = gl(a.Y[ctl.ycur]) + off;
    const double2* YN = gl(a.Y[ctl.ycur ^ 1]) + off;
    const double t = ctl.t, tn = ctl.tn, dt = ctl.dt;
    for (int sidx = ctl.save_from; sidx < ctl.save_to; sidx++) {
        const double ti = gl(a.zgrid)[sidx];
        double2* out = gl(a.saves) + nst * sidx + off;
        if (ti == t) {
            for (int k = threadIdx.x; k < N; k += LH_T) out[k] = Y[k];
        } else if (ti == tn) {
            for (int k = threadIdx.x; k < N; k += LH_T) out[k] = YN[k];
        } else {
            double sg = (ti - t) / dt;
            double sp[4] = {sg, pow(sg, 2.0), pow(sg, 3.0), pow(sg, 4.0)};
            double b[7];
            for (int j = 0; j < 7; j++) {
                double acc = sp[0] * c_iC[0][j];
                for (int p = 1; p < 4; p++) acc = acc + sp[p] * c_iC[p][j];
                b[j] = acc;
            }
            const double h = ti - t;
            for (int k = threadIdx.x; k < N; k += LH_T) {
                double2 yi = make_double2(0.0, 0.0);
                for (int j = 0; j < 7; j++) {
                    if (j == 1) continue;
                    const double2* kj = gl(a.K[ctl.perm[j == 6 ? 0 : j]]) + off;
                    yi = cadd(yi, cscal(b[j], kj[k]));
                }
                out[k] = cmul(cadd(Y[k], cscal(dt, yi)), cexp_lh(gl(a.linop)[off + k], h));
            }
        }
    }
}

// ---------------------------------------------------------------- step end
__global__ void rad_step_end(const RadArgs* __restrict__ ap) {
    const RadArgs& a = *ap;
    RadCtl* ctl = gl(a.ctl);
    if (threadIdx.x != 0 || ctl->done) return;
    RadCtl s = *ctl;
    if (s.ok) s.ycur ^= 1;  // y ← yn
    s.h0 = s.tn - s.t;      // k1 is carried to the new t (zero after a rejection)
    s.t = s.tn;
    s.dt = s.dtn;
    if (!(s.tn <= a.zmax)) {
        s.done = 1;
    } else if (s.ok && a.zstop > 0.0 && s.tn > a.zstop) {
        s.status = LH_E_STOPPED;
        s.done = 1;
    } else if (s.steps >= a.max_steps) {
        s.status = LH_E_MAXSTEPS;
        s.done = 1;
    }
    *ctl = s;
}

// Launch helpers for one RHS and one step (also used during graph capture).
// A plan is one variant of the RHS launch sequence: the radix plan of the
// column transforms (0: radix 8, radix 16 from 8192 points; 1: radix 16
// throughout) and the MFMA tiles per wave along t of the two GEMMs (1, 2 or
// 4; n_to is a multiple of 32·4).  Every plan computes the same transform;
// the last bits of the result differ with the radix plan.
struct RadPlan {
    int fft;  // 0 or 1
    int mt;   // 1, 2 or 4
};
static constexpr RadPlan RAD_DEFAULT_PLAN = {0, 2};
static bool rad_plan_ok(const RadPlan& p) { return (p.fft == 0 || p.fft == 1) && (p.mt == 1 || p.mt == 2 || p.mt == 4); }

template <int MO, int PLAN, int MT>
static void rad_launch_rhs(const RadArgs* dargs, const double2* tw, int n_to, int n_r, int stage, hipStream_t s) {
    dim3 g(n_to / (32 * MT), n_r / 64);
    hipLaunchKernelGGL((rad_stage_a<MO, PLAN>), dim3(n_r), dim3(LH_T), 0, s, dargs, tw, stage);
    hipLaunchKernelGGL((rad_gemm<MT>), g, dim3(256), 0, s, dargs, 0);
    hipLaunchKernelGGL((rad_pointwise<MO>), dim3(n_r), dim3(LH_T), 0, s, dargs);
    hipLaunchKernelGGL((rad_gemm<MT>), g, dim3(256), 0, s, dargs, 1);
    hipLaunchKernelGGL((rad_stage_e<MO, PLAN>), dim3(n_r), dim3(LH_T), 0, s, dargs, tw, stage);
}

template <int MO, int MC, int PLAN, int MT>
static void rad_launch_step(const RadArgs* dargs, const double2* tw, int n_to, int n_r, hipStream_t s) {
    for (int st = 0; st < 6; st++) rad_launch_rhs<MO, PLAN, MT>(dargs, tw, n_to, n_r, st, s);
    hipLaunchKernelGGL(rad_combine, dim3(n_r), dim3(LH_T), 0, s, dargs);
    hipLaunchKernelGGL(rad_control, dim3(1), dim3(64), 0, s, dargs);
    hipLaunchKernelGGL((rad_stepfun<MC>), dim3(n_r), dim3(LH_T), 0, s, dargs, tw);
    hipLaunchKernelGGL(rad_saves, dim3(n_r), dim3(LH_T), 0, s, dargs);
    hipLaunchKernelGGL(rad_step_end, dim3(1), dim3(64), 0, s, dargs);
}

// Envelope grid, Nt = Nto = NO: the GEMMs run over 2·Nto rows.
template <int NO, int MT>
static void rad_launch_rhs_env(const RadArgs* dargs, const double2* tw, int n_r, int stage, hipStream_t s) {
    dim3 g(2 * NO / (32 * MT), n_r / 64);
    hipLaunchKernelGGL((rad_stage_a_env<NO>), dim3(n_r), dim3(LH_T), 0, s, dargs, tw, stage);
    hipLaunchKernelGGL((rad_gemm<MT>), g, dim3(256), 0, s, dargs, 0);
    hipLaunchKernelGGL(rad_pointwise_env, dim3(n_r), dim3(LH_T), 0, s, dargs);
    hipLaunchKernelGGL((rad_gemm<MT>), g, dim3(256), 0, s, dargs, 1);
    hipLaunchKernelGGL((rad_stage_e_env<NO>), dim3(n_r), dim3(LH_T), 0, s, dargs, tw, stage);
}

template <int NO, int MT>
static void rad_run_env(bool step, const RadArgs* a, const double2* tw, int n_r, int stage, hipStream_t s) {
    if (!step) {
        rad_launch_rhs_env<NO, MT>(a, tw, n_r, stage, s);
        return;
    }
    for (int st = 0; st < 6; st++) rad_launch_rhs_env<NO, MT>(a, tw, n_r, st, s);
    hipLaunchKernelGGL(rad_combine, dim3(n_r), dim3(LH_T), 0, s, a);
    hipLaunchKernelGGL(rad_control, dim3(1), dim3(64), 0, s, a);
    hipLaunchKernelGGL((rad_stepfun_env<NO>), dim3(n_r), dim3(LH_T), 0, s, a, tw);
    hipLaunchKernelGGL(rad_saves, dim3(n_r), dim3(LH_T), 0, s, a);
    hipLaunchKernelGGL(rad_step_end, dim3(1), dim3(64), 0, s, a);
}

template <int NO>
static void rad_run_env_mt(const RadPlan& p, bool step, const RadArgs* a, const double2* tw, int n_r, int stage,
                           hipStream_t s) {
    switch (p.mt) {
        case 1: rad_run_env<NO, 1>(step, a, tw, n_r, stage, s); break;
        case 4: rad_run_env<NO, 4>(step, a, tw, n_r, stage, s); break;
        default: rad_run_env<NO, 2>(step, a, tw, n_r, stage, s); break;
    }
}

// Oversampled lengths with compiled radial kernels (field grid: Nt = Nto/2,
// Nω = Nt/2+1; envelope grid: Nω = Nt = Nto).
static bool rad_size_ok(int n_to) { return n_to == 2048 || n_to == 4096 || n_to == 8192 || n_to == 16384; }
static bool rad_env_size_ok(int n_to) { return n_to >= 256 && n_to <= 8192 && (n_to & (n_to - 1)) == 0; }

// One RHS (stage < 0: of rhs_in or the current state) or one step, under plan p.
template <int MO, int MC, int PLAN, int MT>
static void rad_run(bool step, const RadArgs* a, const double2* tw, int n_to, int n_r, int stage, hipStream_t s) {
    if (step)
        rad_launch_step<MO, MC, PLAN, MT>(a, tw, n_to, n_r, s);
    else
        rad_launch_rhs<MO, PLAN, MT>(a, tw, n_to, n_r, stage, s);
}

template <int MO, int MC, int PLAN>
static void rad_run_mt(const RadPlan& p, bool step, const RadArgs* a, const double2* tw, int n_to, int n_r, int stage,
                       hipStream_t s) {
    switch (p.mt) {
        case 1: rad_run<MO, MC, PLAN, 1>(step, a, tw, n_to, n_r, stage, s); break;
        case 4: rad_run<MO, MC, PLAN, 4>(step, a, tw, n_to, n_r, stage, s); break;
        default: rad_run<MO, MC, PLAN, 2>(step, a, tw, n_to, n_r, stage, s); break;
    }
}

template <int MO, int MC>
static void rad_run_plan(const RadPlan& p, bool step, const RadArgs* a, const double2* tw, int n_to, int n_r,
                         int stage, hipStream_t s) {
    if (p.fft == 1)
        rad_run_mt<MO, MC, 1>(p, step, a, tw, n_to, n_r, stage, s);
    else
        rad_run_mt<MO, MC, 0>(p, step, a, tw, n_to, n_r, stage, s);
}

static void rad_dispatch(const RadPlan& p, bool step, bool env, int n_to, int n_r, const RadArgs* a,
                         const double2* tw, int stage, hipStream_t s) {
    if (env) {  // the radix plan of the column transforms is the default one
        switch (n_to) {
            case 8192: rad_run_env_mt<8192>(p, step, a, tw, n_r, stage, s); break;
            case 4096: rad_run_env_mt<4096>(p, step, a, tw, n_r, stage, s); break;
            case 2048: rad_run_env_mt<2048>(p, step, a, tw, n_r, stage, s); break;
            case 1024: rad_run_env_mt<1024>(p, step, a, tw, n_r, stage, s); break;
            case 512: rad_run_env_mt<512>(p, step, a, tw, n_r, stage, s); break;
            default: rad_run_env_mt<256>(p, step, a, tw, n_r, stage, s); break;
        }
        return;
    }
    switch (n_to) {
        case 16384: rad_run_plan<8192, 4096>(p, step, a, tw, n_to, n_r, stage, s); break;
        case 8192: rad_run_plan<4096, 2048>(p, step, a, tw, n_to, n_r, stage, s); break;
        case 4096: rad_run_plan<2048, 1024>(p, step, a, tw, n_to, n_r, stage, s); break;
        default: rad_run_plan<1024, 512>(p, step, a, tw, n_to, n_r, stage, s); break;
    }
}

static void rad_rhs_dispatch(const RadPlan& p, bool env, int n_to, int n_r, const RadArgs* a, const double2* tw,
                             int stage, hipStream_t s) {
    rad_dispatch(p, false, env, n_to, n_r, a, tw, stage, s);
}

static void rad_step_dispatch(const RadPlan& p, bool env, int n_to, int n_r, const RadArgs* a, const double2* tw,
                              hipStream_t s) {
    rad_dispatch(p, true, env, n_to, n_r, a, tw, -1, s);
}
