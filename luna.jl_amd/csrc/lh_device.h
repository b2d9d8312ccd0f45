// lh_device.h -- device building blocks for the propagation kernels (gfx950).
//
// Complex arithmetic on double2, LDS-resident Stockham FFTs, block-wide
// reductions.  One workgroup = LH_T threads (8 wave64s) works on one
// transform; the whole oversampled time series of one RHS lives in LDS.
// 512 threads leave 256 registers per lane for the scan/spline phase.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef LH_T
#define LH_T 512                // threads per workgroup
#endif
#define LH_NW (LH_T / 64)       // wavefronts per workgroup
#define LH_TWN 16384            // size of the global twiddle table

// ---------------------------------------------------------------- complex ops
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscal(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// a * (-i)
__device__ __forceinline__ double2 cmul_mi(double2 a) { return make_double2(a.y, -a.x); }
// a * (+i)
__device__ __forceinline__ double2 cmul_pi(double2 a) { return make_double2(-a.y, a.x); }

// exp(L*h) for complex L and real h, Julia's exp(::Complex) for finite input.
__device__ __forceinline__ double2 cexp_lh(double2 L, double h) {
    double zr = L.x * h, zi = L.y * h;
    double er = exp(zr);
    if (zi == 0.0) return make_double2(er, zi);
    double s, c;
    sincos(zi, &s, &c);
    return make_double2(er * c, er * s);
}

// exp(L*h) as cexp_lh, and in `inv` exp(-L*h) from the same exp and sincos
// (cos is even and sin odd, so only 1/exp(Re) stands in for exp(-Re): one
// rounding apart).
__device__ __forceinline__ double2 cexp_lh_inv(double2 L, double h, double2& inv) {
    double zr = L.x * h, zi = L.y * h;
    double er = exp(zr);
    double ir = 1.0 / er;
    if (zi == 0.0) {
        inv = make_double2(ir, -zi);
        return make_double2(er, zi);
    }
    double s, c;
    sincos(zi, &s, &c);
    inv = make_double2(ir * c, -(ir * s));
    return make_double2(er * c, er * s);
}

// Global-memory view of a pointer.  Pointers the kernels read out of their
// parameter structures are generic to the compiler, so accesses through them
// become flat instructions: those need a 64-bit address per lane and also
// count on the LDS wait counter (an LDS wait then waits for outstanding HBM
// traffic too).  Every pointer here is to device (or device-mapped host)
// memory; the address-space cast lets the compiler emit global_*
// instructions (the pointer converts back to a plain one where a function
// takes one, and the compiler still knows where it points).
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = __attribute__((address_space(1))) T*;
#else  // the host pass only type-checks device code
template <class T>
using gptr = T*;
#endif
template <class T>
__device__ __forceinline__ gptr<T> gl(T* p) {
    return (gptr<T>)p;
}

// LDS padding: one spare double2 every 8 keeps radix-8 strided accesses
// conflict-free for 16-byte lanes.
__device__ __forceinline__ int pad(int i) { return i + (i >> 3); }
__host__ __device__ constexpr int padded_len(int n) { return n + (n >> 3) + 1; }

// ------------------------------------------------------------- register DFTs
template <bool INV>
__device__ __forceinline__ void dft2(double2* v) {
    double2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
}

template <bool INV>
__device__ __forceinline__ void dft4(double2& v0, double2& v1, double2& v2, double2& v3) {
    double2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = csub(v1, v3);
    v0 = cadd(a0, a2);
    v2 = csub(a0, a2);
    if (!INV) {
        v1 = cadd(a1, cmul_mi(a3));
        v3 = cadd(a1, cmul_pi(a3));
    } else {
        v1 = cadd(a1, cmul_pi(a3));
        v3 = cadd(a1, cmul_mi(a3));
    }
}

template <bool INV>
__device__ __forceinline__ void dft4v(double2* v) { dft4<INV>(v[0], v[1], v[2], v[3]); }

template <bool INV>
__device__ __forceinline__ void dft8(double2* v) {
    const double c = 0.70710678118654752440;
    double2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    double2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    double2 w1, w2, w3;
    if (!INV) {
        w1 = make_double2(c * (o1.x + o1.y), c * (o1.y - o1.x));
        w2 = cmul_mi(o2);
        w3 = make_double2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
    } else {
        w1 = make_double2(c * (o1.x - o1.y), c * (o1.x + o1.y));
        w2 = cmul_pi(o2);
        w3 = make_double2(-c * (o3.x + o3.y), c * (o3.x - o3.y));
    }
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1); v[5] = csub(e1, w1);
    v[2] = cadd(e2, w2); v[6] = csub(e2, w2);
    v[3] = cadd(e3, w3); v[7] = csub(e3, w3);
}

// 16 points as 4 × 4: DFT-4 over n1 for each n2 (x[4 n1 + n2]), the twiddles
// W16^(n2 k1), DFT-4 over n2 for each k1, then X[k1 + 4 k2] into natural order.
template <bool INV>
__device__ __forceinline__ void dft16(double2* v) {
    const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
    const double sg = INV ? 1.0 : -1.0;  // W16^m = cos(πm/8) + sg·i sin(πm/8)
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) dft4<INV>(v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]);
    // v[n2 + 4 k1] holds Y[n2][k1]
    v[5] = cmul(v[5], make_double2(c1, sg * s1));    // m = 1
    v[9] = cmul(v[9], make_double2(h, sg * h));      // m = 2
    v[13] = cmul(v[13], make_double2(s1, sg * c1));  // m = 3
    v[6] = cmul(v[6], make_double2(h, sg * h));      // m = 2
    v[10] = INV ? cmul_pi(v[10]) : cmul_mi(v[10]);   // m = 4
    v[14] = cmul(v[14], make_double2(-h, sg * h));   // m = 6
    v[7] = cmul(v[7], make_double2(s1, sg * c1));    // m = 3
    v[11] = cmul(v[11], make_double2(-h, sg * h));   // m = 6
    v[15] = cmul(v[15], make_double2(-c1, -sg * s1));  // m = 9
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) dft4<INV>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    // v[4 k1 + k2] holds X[k1 + 4 k2]
    double2 o[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++)
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = o[k];
}

template <int R, bool INV>
__device__ __forceinline__ void dft(double2* v) {
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 4) dft4v<INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
    else dft16<INV>(v);
}

// Thread index that the compiler must re-read at each use site.  Index and
// LDS-address arithmetic derived from it then stays local to each FFT pass or
// scan instead of being hoisted out of the solver's loops, where dozens of
// loop-invariant addresses would otherwise be kept live (and spilled) across
// the whole right-hand side.
__device__ __forceinline__ int fresh_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// ------------------------------------------------------------ twiddles
// Two-level table in LDS: W^m = hi[m >> 7] · lo[m & 127], W = exp(-2πi/LH_TWN).
// Filled once per workgroup from the global table (4 KB of LDS).
struct TwLds {
    double2 lo[128];
    double2 hi[128];
};

__device__ __forceinline__ void tw_load(TwLds& t, const double2* __restrict__ tw) {
    for (int i = threadIdx.x; i < 128; i += LH_T) {
        t.lo[i] = tw[i];
        t.hi[i] = tw[128 * i];
    }
}

__device__ __forceinline__ double2 tw_get(const TwLds& t, int m) { return cmul(t.hi[m >> 7], t.lo[m & 127]); }

// ------------------------------------------------------------ Stockham FFT
// One radix-R Stockham pass over M complex points in LDS (padded layout).
// Active threads: TA = min(M/8, LH_T); each does M/(R*TA) butterflies (one for a
// radix-16 pass of 8192 points).  The
// twiddle w = exp(∓2πi k/(NS·R)) comes from the LDS table; its powers w^r by
// repeated multiplication (error ≤ R ulp, far below the 1e-12 parity budget).
template <int M, int R, int NS, bool INV, int PLAN = 0>
__device__ __forceinline__ void fft_pass(double2* buf, const TwLds& tw) {
    // plan 0: min(M/8, LH_T) threads (one radix-8 butterfly each where they
    // fit); plan 1: one thread per radix-R butterfly up to LH_T
    constexpr int NB = M / R;
    constexpr int TA0 = (PLAN == 1) ? ((M / R < LH_T) ? M / R : LH_T) : ((M / 8 < LH_T) ? M / 8 : LH_T);
    constexpr int TA = (TA0 < NB) ? TA0 : NB;  // never more threads than butterflies
    constexpr int BPT = NB / TA;
    static_assert(BPT >= 1 && NB % TA == 0, "FFT pass split");
    constexpr int TSTRIDE = LH_TWN / (NS * R);
    const int tid = fresh_tid();
    double2 v[BPT][R];
    if (tid < TA) {
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            int j = tid + b * TA;
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = buf[pad(j + r * NB)];
        }
    }
    __syncthreads();
    if (tid < TA) {
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            int j = tid + b * TA;
            int k = j & (NS - 1);
            if constexpr (NS > 1) {
                double2 w1 = tw_get(tw, k * TSTRIDE);
                if (INV) w1.y = -w1.y;
                double2 w = w1;
#pragma unroll
                for (int r = 1; r < R; r++) {
                    v[b][r] = cmul(v[b][r], w);
                    if (r + 1 < R) w = cmul(w, w1);
                }
            }
            dft<R, INV>(v[b]);
            int idx = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[pad(idx + r * NS)] = v[b][r];
        }
    }
    __syncthreads();
}

// Radix choice per pass.  Plan 0: 8, except that transforms of 8192 points
// and more take radix 16 after the first pass (8·16·16·4 for 8192: four LDS
// passes instead of five, each still one butterfly per thread).  Plan 1:
// radix 16 wherever 16 points remain (16·16·16 for 4096: three passes with
// half the threads active instead of four with all).  Which plan is faster
// for a given kernel is measured, not assumed (lh_autotune).
template <int M, int NS, int PLAN = 0>
struct PassRadix {
    static constexpr int R =
        (PLAN == 1) ? ((M / NS >= 16) ? 16 : (M / NS))
                    : ((M >= 8192 && NS > 1 && M / NS >= 16) ? 16 : ((M / NS >= 8) ? 8 : (M / NS)));
};

template <int M, int NS, bool INV, int PLAN = 0>
__device__ __forceinline__ void fft_passes(double2* buf, const TwLds& tw) {
    if constexpr (NS < M) {
        constexpr int R = PassRadix<M, NS, PLAN>::R;
        fft_pass<M, R, NS, INV, PLAN>(buf, tw);
        fft_passes<M, NS * R, INV, PLAN>(buf, tw);
    }
}

// Unnormalised complex FFT of length M held in LDS (natural order in and out).
// INV=false: exp(-2πi jk/M) (FFTW_FORWARD); INV=true: exp(+2πi jk/M).
// Caller must __syncthreads() before (data complete); returns synchronised.
template <int M, bool INV, int PLAN = 0>
__device__ __forceinline__ void fft_lds(double2* buf, const TwLds& tw) {
    fft_passes<M, 1, INV, PLAN>(buf, tw);
}

// Half-length real-transform packing.  For a real series x of length 2M held
// as z[j] = x[2j] + i x[2j+1]:
//   c2r pre-step: X[0..M] → Z[0..M-1] with Z = A + iB,
//     A = X[k] + conj(X[M-k]),  B = (X[k] - conj(X[M-k]))·conj(W_{2M}^k)
//   (imag of X[0] and X[M] ignored as in FFTW's c2r), then z = IFFT_M(Z).
template <int M>
__device__ __forceinline__ void c2r_prepack(double2* buf, const TwLds& tw) {
    constexpr int TS = LH_TWN / (2 * M);
    for (int k = fresh_tid(); k <= M / 2; k += LH_T) {
        int m = M - k;
        double2 Xk = buf[pad(k)], Xm = buf[pad(m)];
        if (k == 0) {
            Xk.y = 0.0;
            Xm.y = 0.0;
            double2 A = cadd(Xk, cconj(Xm));
            double2 B = csub(Xk, cconj(Xm));
            buf[pad(0)] = cadd(A, cmul_pi(B));
        } else if (k == M / 2) {
            double2 A = cadd(Xk, cconj(Xk));
            double2 B = cmul(csub(Xk, cconj(Xk)), cconj(tw_get(tw, k * TS)));
            buf[pad(k)] = cadd(A, cmul_pi(B));
        } else {
            double2 wk = cconj(tw_get(tw, k * TS)), wm = cconj(tw_get(tw, m * TS));
            double2 Ak = cadd(Xk, cconj(Xm));
            double2 Bk = cmul(csub(Xk, cconj(Xm)), wk);
            double2 Am = cadd(Xm, cconj(Xk));
            double2 Bm = cmul(csub(Xm, cconj(Xk)), wm);
            buf[pad(k)] = cadd(Ak, cmul_pi(Bk));
            buf[pad(m)] = cadd(Am, cmul_pi(Bm));
        }
    }
}

// c2r pre-pack of one bin when X is zero above M/2 (a field spectrum of
// M/2 + 1 bins zero-padded to M): the pair (k, M − k) then depends on X[k]
// alone, so the owner of bin k writes Z[k] and Z[M − k] itself, by the same
// operations as c2r_prepack with X[M − k] = 0.
template <int M>
__device__ __forceinline__ void c2r_pack_bin(double2* buf, const TwLds& tw, int k, double2 Xk) {
    constexpr int TS = LH_TWN / (2 * M);
    const double2 Xm = make_double2(0.0, 0.0);
    if (k == 0) {
        Xk.y = 0.0;
        double2 A = cadd(Xk, cconj(Xm));
        double2 B = csub(Xk, cconj(Xm));
        buf[pad(0)] = cadd(A, cmul_pi(B));
    } else if (k == M / 2) {
        double2 A = cadd(Xk, cconj(Xk));
        double2 B = cmul(csub(Xk, cconj(Xk)), cconj(tw_get(tw, k * TS)));
        buf[pad(k)] = cadd(A, cmul_pi(B));
    } else {
        const int m = M - k;
        double2 wk = cconj(tw_get(tw, k * TS)), wm = cconj(tw_get(tw, m * TS));
        double2 Ak = cadd(Xk, cconj(Xm));
        double2 Bk = cmul(csub(Xk, cconj(Xm)), wk);
        double2 Am = cadd(Xm, cconj(Xk));
        double2 Bm = cmul(csub(Xm, cconj(Xk)), wm);
        buf[pad(k)] = cadd(Ak, cmul_pi(Bk));
        buf[pad(m)] = cadd(Am, cmul_pi(Bm));
    }
}

// r2c post-step for bin k (0 <= k <= M): Z = FFT_M(z) in LDS →
//   X[k] = E + W_{2M}^k O,  E = ½(Z[k] + conj(Z[M-k])),  O = -½ i (Z[k] - conj(Z[M-k])).
template <int M>
__device__ __forceinline__ double2 r2c_bin(const double2* buf, const TwLds& tw, int k) {
    constexpr int TS = LH_TWN / (2 * M);
    double2 zk = buf[pad(k == M ? 0 : k)];
    double2 zc = cconj(buf[pad(k == 0 ? 0 : M - k)]);
    double2 E = cscal(0.5, cadd(zk, zc));
    double2 d = csub(zk, zc);
    double2 O = make_double2(0.5 * d.y, -(0.5 * d.x));
    double2 w = (k == M) ? make_double2(-1.0, 0.0) : tw_get(tw, k * TS);
    return cadd(E, cmul(w, O));
}

// ------------------------------------------------------- lane movement (DPP)
// Data-parallel-primitive moves run on the vector ALU (a move with a lane
// pattern, no LDS round trip as a lane shuffle takes).  A double moves as its
// two 32-bit halves; a lane whose source lies outside its row (or whose row
// ROW_MASK leaves out) receives `fill`'s bits, 0.0 by default.
enum : int {
    DPP_ROW_SHR1 = 0x111,  // lane l ← lane l−1 of the same row of 16
    DPP_ROW_SHR2 = 0x112,
    DPP_ROW_SHR4 = 0x114,
    DPP_ROW_SHR8 = 0x118,
    DPP_WAVE_SHR1 = 0x138,  // lane l ← lane l−1 of the wavefront
    DPP_ROW_BCAST15 = 0x142,  // row r+1 ← lane 15 of row r
    DPP_ROW_BCAST31 = 0x143,  // rows 2, 3 ← lane 31
};

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double x, double fill = 0.0) {
    const long long b = __double_as_longlong(x), f = __double_as_longlong(fill);
    const int lo = __builtin_amdgcn_update_dpp((int)f, (int)b, CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(f >> 32), (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_i32(int x, int fill) {
    return __builtin_amdgcn_update_dpp(fill, x, CTRL, ROW_MASK, 0xf, false);
}

// Value of x in lane `lane` (wave-uniform) for every lane.
__device__ __forceinline__ double lane_value(double x, int lane) {
    const long long b = __double_as_longlong(x);
    const int l = __builtin_amdgcn_readfirstlane(lane);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Inclusive prefix sum over the 64 lanes: within rows of 16 by row shifts,
// then across rows by the two row broadcasts.
__device__ __forceinline__ double wave_incl_scan(double x) {
    x += dpp_f64<DPP_ROW_SHR1>(x);
    x += dpp_f64<DPP_ROW_SHR2>(x);
    x += dpp_f64<DPP_ROW_SHR4>(x);
    x += dpp_f64<DPP_ROW_SHR8>(x);
    x += dpp_f64<DPP_ROW_BCAST15, 0xa>(x);
    x += dpp_f64<DPP_ROW_BCAST31, 0xc>(x);
    return x;
}

// Inclusive prefix sum over lanes 0..LH_NW−1 (row 0; the cross-wave step of
// the workgroup scans, one value per wave).
__device__ __forceinline__ double lanes_incl_scan(double x) {
    static_assert(LH_NW <= 16, "one value per wave in row 0");
    x += dpp_f64<DPP_ROW_SHR1>(x);
    x += dpp_f64<DPP_ROW_SHR2>(x);
    x += dpp_f64<DPP_ROW_SHR4>(x);
    if constexpr (LH_NW > 8) x += dpp_f64<DPP_ROW_SHR8>(x);
    return x;
}

// Minimum of lo and maximum of hi over the 64 lanes, in every lane.
__device__ __forceinline__ void wave_minmax(int& lo, int& hi) {
    constexpr int BIG = 0x7fffffff, SMALL = -0x7fffffff - 1;
    lo = min(lo, dpp_i32<DPP_ROW_SHR1>(lo, BIG));
    hi = max(hi, dpp_i32<DPP_ROW_SHR1>(hi, SMALL));
    lo = min(lo, dpp_i32<DPP_ROW_SHR2>(lo, BIG));
    hi = max(hi, dpp_i32<DPP_ROW_SHR2>(hi, SMALL));
    lo = min(lo, dpp_i32<DPP_ROW_SHR4>(lo, BIG));
    hi = max(hi, dpp_i32<DPP_ROW_SHR4>(hi, SMALL));
    lo = min(lo, dpp_i32<DPP_ROW_SHR8>(lo, BIG));
    hi = max(hi, dpp_i32<DPP_ROW_SHR8>(hi, SMALL));
    lo = min(lo, dpp_i32<DPP_ROW_BCAST15, 0xa>(lo, BIG));
    hi = max(hi, dpp_i32<DPP_ROW_BCAST15, 0xa>(hi, SMALL));
    lo = min(lo, dpp_i32<DPP_ROW_BCAST31, 0xc>(lo, BIG));
    hi = max(hi, dpp_i32<DPP_ROW_BCAST31, 0xc>(hi, SMALL));
    lo = __builtin_amdgcn_readlane(lo, 63);
    hi = __builtin_amdgcn_readlane(hi, 63);
}

// ------------------------------------------------------- block reductions
// Sum of four values over the workgroup (fixed order → deterministic);
// scratch: 4·LH_NW doubles.  Ends synchronised.
__device__ __forceinline__ void block_sum4(double& a, double& b, double& c, double& d, double* scratch) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
        d += __shfl_xor(d, off, 64);
    }
    if (lane == 0) {
        scratch[4 * wid + 0] = a;
        scratch[4 * wid + 1] = b;
        scratch[4 * wid + 2] = c;
        scratch[4 * wid + 3] = d;
    }
    __syncthreads();
    double sa = 0, sb = 0, sc = 0, sd = 0;
    for (int w = 0; w < LH_NW; w++) {
        sa += scratch[4 * w + 0];
        sb += scratch[4 * w + 1];
        sc += scratch[4 * w + 2];
        sd += scratch[4 * w + 3];
    }
    __syncthreads();
    a = sa;
    b = sb;
    c = sc;
    d = sd;
}

// Sum of three values over the workgroup (fixed tree → deterministic).
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double* scratch) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    if (lane == 0) {
        scratch[3 * wid + 0] = a;
        scratch[3 * wid + 1] = b;
        scratch[3 * wid + 2] = c;
    }
    __syncthreads();
    double sa = 0, sb = 0, sc = 0;
    for (int w = 0; w < LH_NW; w++) {
        sa += scratch[3 * w + 0];
        sb += scratch[3 * w + 1];
        sc += scratch[3 * w + 2];
    }
    __syncthreads();
    a = sa;
    b = sb;
    c = sc;
}
