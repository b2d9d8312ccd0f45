// lh_fno.hip -- 1-D Fourier neural operator surrogate on gfx950.
//
// Every matrix product (lift, the truncated forward/inverse DFTs of the
// spectral convolution, the per-mode complex channel mixing, the pointwise
// channel mixing, the projection, and all their gradients) is one call of a
// generic strided, batched GEMM on BF16 matrix cores with FP32 accumulation.
// The truncated DFT is a dense (2·modes × n_x) product: with 32 modes of a
// 1024-point series it is cheaper as a matrix-core GEMM than as an FFT, and
// it keeps the whole network on the MFMA path.  Elementwise work (GELU and its
// derivative, loss, AdamW) runs in FP32.  Reductions over the batch (bias
// gradients, split-K partial sums, the loss) use fixed orders, so results are
// deterministic run to run.

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/luna_fno.h"
#include "../../include/luna_hip.h"

typedef float fno_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 fno_bf8 __attribute__((ext_vector_type(8)));

static const int FNO_MAXL = 8;
static const int FNO_T = 256;

#define FNOCK(x)                                                                     \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            fprintf(stderr, "luna_fno: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return (e_ == hipErrorOutOfMemory) ? LH_E_NOMEM : LH_E_HIP;              \
        }                                                                            \
    } while (0)

// ---------------------------------------------------------------------------
// Generic GEMM: C[b] = alpha · A[b] · B[b] (+ beta · C[b]) (+ bias[n])
//   A(m, k) = A[b·sa_b + m·sa_m + k·sa_k],  B(k, n) = B[b·sb_b + k·sb_k + n·sb_n],
//   C(m, n) = C[b·sc_b + m·sc_m + n·sc_n].
// Workgroup tile 64×64, K step 32; four wave64s, each 32×32 = 2×2 tiles of
// v_mfma_f32_16x16x32_bf16.  Lane l supplies A[l%16][8(l/16)..8(l/16)+7] and
// B[8(l/16)..+7][l%16]; accumulator register r of lane l is D[4(l/16)+r][l%16].
// Operands are converted to BF16 (round to nearest even) on their way to LDS.
// With ksplit > 1 each z-slice sums a K range into work[], and a second
// kernel adds the slices in order.
// ---------------------------------------------------------------------------
struct GemmArgs {
    int M, N, K, batch, ksplit, kchunk;
    const void* A;
    long long sa_m, sa_k, sa_b;
    const void* B;
    long long sb_k, sb_n, sb_b;
    float* C;
    long long sc_m, sc_n, sc_b;
    float alpha, beta;
    const float* bias;
    float* work;
};

__device__ __forceinline__ __bf16 ld_bf(const float* p) { return (__bf16)(*p); }
__device__ __forceinline__ __bf16 ld_bf(const __bf16* p) { return *p; }

template <class TA, class TB>
__global__ __launch_bounds__(FNO_T) void fno_gemm(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) __bf16 As[64][40];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[64][40];  // B tile as [n][k]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int n0 = blockIdx.x * 64, m0 = blockIdx.y * 64;
    const int bz = blockIdx.z, b = bz / g.ksplit, ks = bz % g.ksplit;
    const int kbeg = ks * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    const TA* A = (const TA*)g.A + (long long)b * g.sa_b;
    const TB* B = (const TB*)g.B + (long long)b * g.sb_b;
    const bool a_kc = (g.sa_k == 1), b_nc = (g.sb_n == 1);
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int lr = lane & 15, lq = lane >> 4;
    fno_f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = (fno_f4){0.f, 0.f, 0.f, 0.f};
    const __bf16 zero = (__bf16)0.0f;
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
        if (a_kc) {
            const int kk = tid & 31, k = k0 + kk;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int mm = (tid >> 5) + 8 * i, m = m0 + mm;
                As[mm][kk] = (m < g.M && k < kend) ? ld_bf(A + (long long)m * g.sa_m + k) : zero;
            }
        } else {
            const int mm = tid & 63, m = m0 + mm;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int kk = (tid >> 6) + 4 * i, k = k0 + kk;
                As[mm][kk] = (m < g.M && k < kend) ? ld_bf(A + (long long)m * g.sa_m + (long long)k * g.sa_k) : zero;
            }
        }
        if (b_nc) {
            const int nn = tid & 63, n = n0 + nn;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int kk = (tid >> 6) + 4 * i, k = k0 + kk;
                Bs[nn][kk] = (n < g.N && k < kend) ? ld_bf(B + (long long)k * g.sb_k + n) : zero;
            }
        } else {
            const int kk = tid & 31, k = k0 + kk;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int nn = (tid >> 5) + 8 * i, n = n0 + nn;
                Bs[nn][kk] = (n < g.N && k < kend) ? ld_bf(B + (long long)k * g.sb_k + (long long)n * g.sb_n) : zero;
            }
        }
        __syncthreads();
        fno_bf8 af[2], bfr[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            af[t] = *(const fno_bf8*)&As[wm + 16 * t + lr][8 * lq];
            bfr[t] = *(const fno_bf8*)&Bs[wn + 16 * t + lr][8 * lq];
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * lq + r, n = n0 + wn + 16 * j + lr;
                if (m >= g.M || n >= g.N) continue;
                const float v = acc[i][j][r];
                if (g.ksplit > 1) {
                    g.work[((long long)bz * g.M + m) * g.N + n] = v;
                } else {
                    const long long ci = (long long)b * g.sc_b + (long long)m * g.sc_m + (long long)n * g.sc_n;
                    float o = g.alpha * v;
                    if (g.beta != 0.0f) o += g.beta * g.C[ci];
                    if (g.bias) o += g.bias[n];
                    g.C[ci] = o;
                }
            }
}

// FP32 variant for the projection head: same tiling, K step 16 in LDS,
// v_mfma_f32_16x16x4f32 (lane l supplies A[l%16][l/16] and B[l/16][l%16];
// accumulator register r of lane l is D[4(l/16)+r][l%16]).
__global__ __launch_bounds__(FNO_T) void fno_gemm32(GemmArgs g) {
    __shared__ float As[64][17];
    __shared__ float Bs[64][17];  // B tile as [n][k]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int n0 = blockIdx.x * 64, m0 = blockIdx.y * 64;
    const int bz = blockIdx.z, b = bz / g.ksplit, ks = bz % g.ksplit;
    const int kbeg = ks * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    const float* A = (const float*)g.A + (long long)b * g.sa_b;
    const float* B = (const float*)g.B + (long long)b * g.sb_b;
    const bool a_kc = (g.sa_k == 1), b_nc = (g.sb_n == 1);
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int lr = lane & 15, lq = lane >> 4;
    fno_f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = (fno_f4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        if (a_kc) {
            const int kk = tid & 15, k = k0 + kk;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int mm = (tid >> 4) + 16 * i, m = m0 + mm;
                As[mm][kk] = (m < g.M && k < kend) ? A[(long long)m * g.sa_m + k] : 0.0f;
            }
        } else {
            const int mm = tid & 63, m = m0 + mm;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int kk = (tid >> 6) + 4 * i, k = k0 + kk;
                As[mm][kk] = (m < g.M && k < kend) ? A[(long long)m * g.sa_m + (long long)k * g.sa_k] : 0.0f;
            }
        }
        if (b_nc) {
            const int nn = tid & 63, n = n0 + nn;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int kk = (tid >> 6) + 4 * i, k = k0 + kk;
                Bs[nn][kk] = (n < g.N && k < kend) ? B[(long long)k * g.sb_k + n] : 0.0f;
            }
        } else {
            const int kk = tid & 15, k = k0 + kk;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int nn = (tid >> 4) + 16 * i, n = n0 + nn;
                Bs[nn][kk] = (n < g.N && k < kend) ? B[(long long)k * g.sb_k + (long long)n * g.sb_n] : 0.0f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int s4 = 0; s4 < 4; s4++) {
            float af[2], bfr[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                af[t] = As[wm + 16 * t + lr][4 * s4 + lq];
                bfr[t] = Bs[wn + 16 * t + lr][4 * s4 + lq];
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * lq + r, n = n0 + wn + 16 * j + lr;
                if (m >= g.M || n >= g.N) continue;
                const float v = acc[i][j][r];
                if (g.ksplit > 1) {
                    g.work[((long long)bz * g.M + m) * g.N + n] = v;
                } else {
                    const long long ci = (long long)b * g.sc_b + (long long)m * g.sc_m + (long long)n * g.sc_n;
                    float o = g.alpha * v;
                    if (g.beta != 0.0f) o += g.beta * g.C[ci];
                    if (g.bias) o += g.bias[n];
                    g.C[ci] = o;
                }
            }
}

// The k-slices of 32 outputs per workgroup: eight groups of threads take the
// slices part, part + 8, ... of each output in order (independent loads, so
// many are in flight), and the eight group sums are added in a fixed tree.
constexpr int RED_OUT = 32, RED_PART = FNO_T / RED_OUT;
static_assert(RED_PART == 8, "the reduction tree below adds eight group sums");
__global__ __launch_bounds__(FNO_T) void fno_splitk_reduce(GemmArgs g) {
    __shared__ float red[RED_PART][RED_OUT + 1];
    const long long MN = (long long)g.M * g.N;
    const int ol = threadIdx.x % RED_OUT, part = threadIdx.x / RED_OUT;
    const long long idx = (long long)blockIdx.x * RED_OUT + ol;
    const bool live = idx < MN * g.batch;
    const int b = live ? (int)(idx / MN) : 0;
    const long long rem = idx - (long long)b * MN;
    float s = 0.0f;
    if (live) {
        const float* w = g.work + (long long)b * g.ksplit * MN + rem;
#pragma unroll 8
        for (int ks = part; ks < g.ksplit; ks += RED_PART) s += w[(long long)ks * MN];
    }
    red[part][ol] = s;
    __syncthreads();
    if (part != 0 || !live) return;
    const float t = ((red[0][ol] + red[1][ol]) + (red[2][ol] + red[3][ol])) +
                    ((red[4][ol] + red[5][ol]) + (red[6][ol] + red[7][ol]));
    const int m = (int)(rem / g.N), n = (int)(rem - (long long)m * g.N);
    const long long ci = (long long)b * g.sc_b + (long long)m * g.sc_m + (long long)n * g.sc_n;
    float o = g.alpha * t;
    if (g.beta != 0.0f) o += g.beta * g.C[ci];
    if (g.bias) o += g.bias[n];
    g.C[ci] = o;
}

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

__global__ void fno_cvt_bf16(const float* __restrict__ src, __bf16* __restrict__ dst, long long n) {
    long long i = (long long)blockIdx.x * FNO_T + threadIdx.x;
    if (i < n) dst[i] = (__bf16)src[i];
}

// Rblk[m] = [[Rre, Rim], [-Rim, Rre]] (2C × 2C, rows: re/im of the input
// channel, columns: re/im of the output channel), so that one real product
// [Vre, Vim] · Rblk gives [Re(V R), Im(V R)].
__global__ void fno_build_rblk(const float* __restrict__ Rre, const float* __restrict__ Rim, __bf16* __restrict__ Rblk,
                               int C, int modes) {
    const long long n = (long long)modes * 4 * C * C;
    long long i = (long long)blockIdx.x * FNO_T + threadIdx.x;
    if (i >= n) return;
    const int C2 = 2 * C;
    const int m = (int)(i / ((long long)C2 * C2));
    const int rem = (int)(i - (long long)m * C2 * C2);
    const int kr = rem / C2, nc = rem % C2;
    const long long base = (long long)m * C * C;
    float v;
    if (kr < C && nc < C)
        v = Rre[base + kr * C + nc];
    else if (kr < C)
        v = Rim[base + kr * C + (nc - C)];
    else if (nc < C)
        v = -Rim[base + (kr - C) * C + nc];
    else
        v = Rre[base + (kr - C) * C + (nc - C)];
    Rblk[i] = (__bf16)v;
}

// dR from dRblk: dRre = dRblk[:C,:C] + dRblk[C:,C:], dRim = dRblk[:C,C:] - dRblk[C:,:C]
__global__ void fno_drblk_to_dr(const float* __restrict__ dRblk, float* __restrict__ dRre, float* __restrict__ dRim,
                                int C, int modes) {
    const long long n = (long long)modes * C * C;
    long long i = (long long)blockIdx.x * FNO_T + threadIdx.x;
    if (i >= n) return;
    const int C2 = 2 * C;
    const int m = (int)(i / ((long long)C * C));
    const int rem = (int)(i - (long long)m * C * C);
    const int r = rem / C, c = rem % C;
    const float* D = dRblk + (long long)m * C2 * C2;
    dRre[i] = D[r * C2 + c] + D[(r + C) * C2 + (c + C)];
    dRim[i] = D[r * C2 + (c + C)] - D[(r + C) * C2 + c];
}

__global__ void fno_gelu_fwd(const float* __restrict__ Z, float* __restrict__ V, long long n) {
    long long i = (long long)blockIdx.x * FNO_T + threadIdx.x;
    if (i < n) V[i] = gelu_f(Z[i]);
}

__global__ void fno_gelu_bwd(const float* __restrict__ dV, const float* __restrict__ Z, float* __restrict__ dZ,
                             long long n) {
    long long i = (long long)blockIdx.x * FNO_T + threadIdx.x;
    if (i < n) dZ[i] = dV[i] * gelu_d(Z[i]);
}

// dZ1[i][j] = dy[i] · Q2[j] · gelu'(Z1[i][j])
__global__ void fno_dz1(const float* __restrict__ dy, const float* __restrict__ Q2, const float* __restrict__ Z1,
                        float* __restrict__ dZ1, long long rows, int H) {
    long long i = (long long)blockIdx.x * FNO_T + threadIdx.x;
    if (i >= rows * H) return;
    const long long r = i / H;
    const int j = (int)(i - r * H);
    dZ1[i] = dy[r] * Q2[j] * gelu_d(Z1[i]);
}

// d = y - t;  loss = Σ w d² / n;  dy = 2 w d / n;  per-block partial sums (fixed tree)
__global__ void fno_loss_fwd(const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ w,
                             float* __restrict__ dy, float* __restrict__ partial, long long n) {
    __shared__ float red[FNO_T];
    long long i = (long long)blockIdx.x * FNO_T + threadIdx.x;
    float d = 0.0f, wi = 0.0f;
    if (i < n) {
        d = y[i] - t[i];
        wi = w ? w[i] : 1.0f;
        dy[i] = 2.0f * wi * d / (float)n;
    }
    red[threadIdx.x] = wi * d * d;
    __syncthreads();
    for (int s = FNO_T / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void fno_loss_reduce(const float* __restrict__ partial, int nb, long long n, float* __restrict__ out) {
    __shared__ float red[FNO_T];
    float s = 0.0f;
    for (int i = threadIdx.x; i < nb; i += FNO_T) s += partial[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = FNO_T / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] / (float)n;
}

// Column sums of X[rows][N] (bias gradients): stage 1 per 128-row chunk,
// stage 2 adds the chunks in order.
static const int FNO_CS_ROWS = 128;
__global__ void fno_colsum_part(const float* __restrict__ X, long long rows, int N, float* __restrict__ part) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), lane4 = threadIdx.x >> 6;
    const long long r0 = (long long)blockIdx.y * FNO_CS_ROWS;
    const long long r1 = (r0 + FNO_CS_ROWS < rows) ? r0 + FNO_CS_ROWS : rows;
    float s = 0.0f;
    if (c < N)
        for (long long r = r0 + lane4; r < r1; r += 4) s += X[r * N + c];
    red[lane4][threadIdx.x & 63] = s;
    __syncthreads();
    if (lane4 == 0 && c < N)
        part[(long long)blockIdx.y * N + c] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// One workgroup per column: thread t adds chunks t, t + 256, ... in order,
// then a fixed tree over the threads (the same order on every call).
__global__ __launch_bounds__(FNO_T) void fno_colsum_final(const float* __restrict__ part, int nchunk, int N,
                                                          float* __restrict__ out) {
    __shared__ float red[FNO_T];
    const int c = blockIdx.x, t = threadIdx.x;
    float s = 0.0f;
    for (int k = t; k < nchunk; k += FNO_T) s += part[(long long)k * N + c];
    red[t] = s;
    __syncthreads();
#pragma unroll
    for (int w = FNO_T / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) out[c] = red[0];
}

// AdamW on the FP32 master copy; refreshes the BF16 copy.
__global__ void fno_adamw(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, __bf16* __restrict__ pbf, long long n, const float* __restrict__ lr_sched,
                          int step_idx, float lr0, float b1, float b2, float eps, float wd, float bc1, float bc2) {
    long long i = (long long)blockIdx.x * FNO_T + threadIdx.x;
    if (i >= n) return;
    const float lr = lr_sched ? lr_sched[step_idx] : lr0;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    const float pi = p[i] - lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
    p[i] = pi;
    pbf[i] = (__bf16)pi;
}

__global__ void fno_gather(const float* __restrict__ xs, const float* __restrict__ ts, const float* __restrict__ ws,
                           const int32_t* __restrict__ idx, float* __restrict__ xb, float* __restrict__ tb,
                           float* __restrict__ wb, int B, int X, int cin) {
    const long long per_x = (long long)X * cin;
    const long long nx = (long long)B * per_x, nt = (long long)B * X;
    long long i = (long long)blockIdx.x * FNO_T + threadIdx.x;
    if (i < nx) {
        const int b = (int)(i / per_x);
        xb[i] = xs[(long long)idx[b] * per_x + (i - (long long)b * per_x)];
    }
    if (i < nt) {
        const int b = (int)(i / X);
        const long long src = (long long)idx[b] * X + (i - (long long)b * X);
        tb[i] = ts[src];
        if (ws) wb[i] = ws[src];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct FnoLayout {
    long long P, bP, W[FNO_MAXL], b[FNO_MAXL], Rre[FNO_MAXL], Rim[FNO_MAXL], Q1, bq1, Q2, bq2, total;
};

struct lh_fno {
    lh_fno_cfg cfg;
    int device;
    hipStream_t stream;
    FnoLayout L;
    float* params;
    float* grads;
    float* adam_m;
    float* adam_v;
    __bf16* pbf;
    __bf16* Rblk;  // [layers][modes][2C][2C]
    __bf16* F;     // [2·modes][n_x]
    __bf16* Ginv;  // [n_x][2·modes]
    float* xb;
    float* tb;
    float* wb;       // loss weights of the batch
    const float* wcur;  // wb, or null for unit weights
    float* V[FNO_MAXL];     // layer inputs v_0 .. v_{L-1}
    float* Z[FNO_MAXL];     // pre-activations z_0 .. z_{L-1}  (v_L = z_{L-1})
    float* Vhat[FNO_MAXL];  // [B][2·modes][C]
    float* Ohat;
    float* Z1;
    float* Hh;
    float* y;
    float* dy;
    float* dVa;
    float* dVb;
    float* dZ;
    float* dOhat;
    float* dVhat;
    float* dZ1;
    float* dRblk;
    float* work;
    long long work_cap;
    float* part;
    long long part_cap;
    float* loss_dev;  // [max steps per train call]
    int loss_cap;
    float* data_x;
    float* data_t;
    float* data_w;
    int data_n;
    int32_t* idx_dev;
    long long idx_cap;
    float* lr_dev;
    long long adam_t;
    std::vector<std::pair<const char*, size_t>>* allocs;  // every device buffer (base, bytes)
};

// Host-side bounds check of a strided operand before any launch: the largest
// element offset must stay inside the allocation the pointer points into.
static bool fno_in_bounds(const lh_fno* f, const void* p, long long max_off, size_t esize) {
    if (!p || max_off < 0) return false;
    const char* c = (const char*)p;
    for (const auto& a : *f->allocs)
        if (c >= a.first && c < a.first + a.second) return c + (size_t)(max_off + 1) * esize <= a.first + a.second;
    return false;
}

static FnoLayout fno_layout(const lh_fno_cfg& c) {
    FnoLayout L;
    memset(&L, 0, sizeof(L));
    const long long C = c.channels, H = c.hidden, Mo = c.modes;
    long long o = 0;
    L.P = o; o += C * c.cin;
    L.bP = o; o += C;
    for (int l = 0; l < c.layers; l++) {
        L.W[l] = o; o += C * C;
        L.b[l] = o; o += C;
        L.Rre[l] = o; o += Mo * C * C;
        L.Rim[l] = o; o += Mo * C * C;
    }
    L.Q1 = o; o += H * C;
    L.bq1 = o; o += H;
    L.Q2 = o; o += H;
    L.bq2 = o; o += 1;
    L.total = o;
    return L;
}

static inline unsigned nblk(long long n) { return (unsigned)((n + FNO_T - 1) / FNO_T); }

template <class T>
static const void* cv(const T* p) { return (const void*)p; }

// Launch one GEMM; ta/tb: 0 = FP32 operand, 1 = BF16 operand (both on BF16
// matrix cores); ta = tb = 2: FP32 operands on FP32 matrix cores.
static int fno_gemm_launch(lh_fno* f, GemmArgs g, int ta, int tb) {
    if (g.M < 1 || g.N < 1 || g.K < 1 || g.batch < 1 || g.sa_m < 0 || g.sa_k < 0 || g.sa_b < 0 || g.sb_k < 0 ||
        g.sb_n < 0 || g.sb_b < 0 || g.sc_m < 0 || g.sc_n < 0 || g.sc_b < 0)
        return LH_E_ARG;
    const long long amax = (long long)(g.batch - 1) * g.sa_b + (long long)(g.M - 1) * g.sa_m + (long long)(g.K - 1) * g.sa_k;
    const long long bmax = (long long)(g.batch - 1) * g.sb_b + (long long)(g.K - 1) * g.sb_k + (long long)(g.N - 1) * g.sb_n;
    const long long cmax = (long long)(g.batch - 1) * g.sc_b + (long long)(g.M - 1) * g.sc_m + (long long)(g.N - 1) * g.sc_n;
    if (!fno_in_bounds(f, g.A, amax, ta == 1 ? 2 : 4) || !fno_in_bounds(f, g.B, bmax, tb == 1 ? 2 : 4) ||
        !fno_in_bounds(f, g.C, cmax, 4) || (g.bias && !fno_in_bounds(f, g.bias, g.N - 1, 4))) {
        fprintf(stderr, "luna_fno: GEMM operand out of bounds (M=%d N=%d K=%d batch=%d)\n", g.M, g.N, g.K, g.batch);
        return LH_E_ARG;
    }
    const long long tiles = (long long)g.batch * ((g.M + 63) / 64) * ((g.N + 63) / 64);
    int ksplit = 1;
    if (tiles < 256 && g.K >= 1024 && !g.bias) {
        // at least ~512 workgroups (two per CU) in k-chunks of at least 128
        long long want = 512 / (tiles > 0 ? tiles : 1);
        long long maxk = (g.K + 127) / 128;
        ksplit = (int)(want < maxk ? want : maxk);
        if (ksplit < 1) ksplit = 1;
        while (ksplit > 1 && (long long)g.batch * ksplit * g.M * g.N > f->work_cap) ksplit--;
    }
    int kchunk = (int)(((g.K + ksplit - 1) / ksplit + 31) / 32 * 32);
    if (kchunk < 32) kchunk = 32;
    ksplit = (g.K + kchunk - 1) / kchunk;
    if (ksplit < 1) ksplit = 1;
    g.ksplit = ksplit;
    g.kchunk = kchunk;
    g.work = f->work;
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, g.batch * ksplit);
    hipStream_t s = f->stream;
    if (ta == 2 && tb == 2)
        hipLaunchKernelGGL(fno_gemm32, grid, dim3(FNO_T), 0, s, g);
    else if (ta == 0 && tb == 0)
        hipLaunchKernelGGL((fno_gemm<float, float>), grid, dim3(FNO_T), 0, s, g);
    else if (ta == 0 && tb == 1)
        hipLaunchKernelGGL((fno_gemm<float, __bf16>), grid, dim3(FNO_T), 0, s, g);
    else if (ta == 1 && tb == 0)
        hipLaunchKernelGGL((fno_gemm<__bf16, float>), grid, dim3(FNO_T), 0, s, g);
    else
        hipLaunchKernelGGL((fno_gemm<__bf16, __bf16>), grid, dim3(FNO_T), 0, s, g);
    if (ksplit > 1) {
        const long long n = (long long)g.batch * g.M * g.N;
        hipLaunchKernelGGL(fno_splitk_reduce, dim3((unsigned)((n + RED_OUT - 1) / RED_OUT)), dim3(FNO_T), 0, s, g);
    }
    FNOCK(hipGetLastError());
    return LH_OK;
}

static GemmArgs G(int M, int N, int K, int batch, const void* A, long long sa_m, long long sa_k, long long sa_b,
                  const void* B, long long sb_k, long long sb_n, long long sb_b, float* C, long long sc_m,
                  long long sc_n, long long sc_b, float beta = 0.0f, const float* bias = nullptr) {
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.M = M; g.N = N; g.K = K; g.batch = batch;
    g.A = A; g.sa_m = sa_m; g.sa_k = sa_k; g.sa_b = sa_b;
    g.B = B; g.sb_k = sb_k; g.sb_n = sb_n; g.sb_b = sb_b;
    g.C = C; g.sc_m = sc_m; g.sc_n = sc_n; g.sc_b = sc_b;
    g.alpha = 1.0f; g.beta = beta; g.bias = bias;
    return g;
}

static int colsum(lh_fno* f, const float* X, long long rows, int N, float* out) {
    const int nchunk = (int)((rows + FNO_CS_ROWS - 1) / FNO_CS_ROWS);
    if ((long long)nchunk * N > f->part_cap) return LH_E_ARG;
    hipLaunchKernelGGL(fno_colsum_part, dim3((N + 63) / 64, nchunk), dim3(FNO_T), 0, f->stream, X, rows, N, f->part);
    hipLaunchKernelGGL(fno_colsum_final, dim3(N), dim3(FNO_T), 0, f->stream, f->part, nchunk, N, out);
    FNOCK(hipGetLastError());
    return LH_OK;
}

static void refresh_bf16(lh_fno* f) {
    const lh_fno_cfg& c = f->cfg;
    hipLaunchKernelGGL(fno_cvt_bf16, dim3(nblk(f->L.total)), dim3(FNO_T), 0, f->stream, f->params, f->pbf, f->L.total);
    const long long rb = (long long)c.modes * 4 * c.channels * c.channels;
    for (int l = 0; l < c.layers; l++)
        hipLaunchKernelGGL(fno_build_rblk, dim3(nblk(rb)), dim3(FNO_T), 0, f->stream, f->params + f->L.Rre[l],
                           f->params + f->L.Rim[l], f->Rblk + (long long)l * rb, c.channels, c.modes);
}

// Forward pass on f->xb (batch B) → f->y; keeps every activation for backward.
static int fno_forward_dev(lh_fno* f, int B) {
    const lh_fno_cfg& c = f->cfg;
    const int X = c.n_x, C = c.channels, Mo2 = 2 * c.modes, H = c.hidden, cin = c.cin;
    const long long BX = (long long)B * X;
    const FnoLayout& L = f->L;
    const __bf16* pb = f->pbf;
    const float* pf = f->params;
    const long long rb = (long long)c.modes * 4 * C * C;
    int st;
    // lift: v0 = x Pᵀ + bP
    st = fno_gemm_launch(f, G((int)BX, C, cin, 1, cv(f->xb), cin, 1, 0, cv(pb + L.P), 1, cin, 0, f->V[0], C, 1, 0, 0.0f,
                              pf + L.bP), 0, 1);
    if (st) return st;
    for (int l = 0; l < c.layers; l++) {
        const float* v = f->V[l];
        // Vhat_b = F · v_b                              [2·modes × C] per sample
        st = fno_gemm_launch(f, G(Mo2, C, X, B, cv(f->F), X, 1, 0, cv(v), C, 1, (long long)X * C, f->Vhat[l], C, 1,
                                  (long long)Mo2 * C), 1, 0);
        if (st) return st;
        // per mode: [B × 2C] · Rblk[m] (2C × 2C)
        st = fno_gemm_launch(f, G(B, 2 * C, 2 * C, c.modes, cv(f->Vhat[l]), (long long)Mo2 * C, 1, 2 * C,
                                  cv(f->Rblk + (long long)l * rb), 2 * C, 1, 4LL * C * C, f->Ohat, (long long)Mo2 * C, 1,
                                  2 * C), 0, 1);
        if (st) return st;
        // z_b = Ginv · Ohat_b                          [X × C] per sample
        st = fno_gemm_launch(f, G(X, C, Mo2, B, cv(f->Ginv), Mo2, 1, 0, cv(f->Ohat), C, 1, (long long)Mo2 * C, f->Z[l],
                                  C, 1, (long long)X * C), 1, 0);
        if (st) return st;
        // z += v Wᵀ + b
        st = fno_gemm_launch(f, G((int)BX, C, C, 1, cv(v), C, 1, 0, cv(pb + L.W[l]), 1, C, 0, f->Z[l], C, 1, 0, 1.0f,
                                  pf + L.b[l]), 0, 1);
        if (st) return st;
        if (l + 1 < c.layers)
            hipLaunchKernelGGL(fno_gelu_fwd, dim3(nblk(BX * C)), dim3(FNO_T), 0, f->stream, f->Z[l], f->V[l + 1], BX * C);
    }
    const float* vL = f->Z[c.layers - 1];
    // projection head on FP32 matrix cores: its output is the log spectrum
    st = fno_gemm_launch(f, G((int)BX, H, C, 1, cv(vL), C, 1, 0, cv(pf + L.Q1), 1, C, 0, f->Z1, H, 1, 0, 0.0f, pf + L.bq1),
                         2, 2);
    if (st) return st;
    hipLaunchKernelGGL(fno_gelu_fwd, dim3(nblk(BX * H)), dim3(FNO_T), 0, f->stream, f->Z1, f->Hh, BX * H);
    st = fno_gemm_launch(f, G((int)BX, 1, H, 1, cv(f->Hh), H, 1, 0, cv(pf + L.Q2), 1, H, 0, f->y, 1, 1, 0, 0.0f,
                              pf + L.bq2), 2, 2);
    if (st) return st;
    FNOCK(hipGetLastError());
    return LH_OK;
}

// Loss against f->tb and the full gradient into f->grads; loss → *loss_out (device).
static int fno_backward_dev(lh_fno* f, int B, float* loss_out) {
    const lh_fno_cfg& c = f->cfg;
    const int X = c.n_x, C = c.channels, Mo2 = 2 * c.modes, H = c.hidden, cin = c.cin;
    const long long BX = (long long)B * X;
    const FnoLayout& L = f->L;
    const __bf16* pb = f->pbf;
    float* gr = f->grads;
    const long long rb = (long long)c.modes * 4 * C * C;
    hipStream_t s = f->stream;
    int st;
    const unsigned nb = nblk(BX);
    if ((long long)nb > f->part_cap) return LH_E_ARG;
    hipLaunchKernelGGL(fno_loss_fwd, dim3(nb), dim3(FNO_T), 0, s, f->y, f->tb, f->wcur, f->dy, f->part, BX);
    hipLaunchKernelGGL(fno_loss_reduce, dim3(1), dim3(FNO_T), 0, s, f->part, (int)nb, BX, loss_out);
    // projection
    st = fno_gemm_launch(f, G(1, H, (int)BX, 1, cv(f->dy), 0, 1, 0, cv(f->Hh), H, 1, 0, gr + L.Q2, H, 1, 0), 2, 2);
    if (st) return st;
    st = colsum(f, f->dy, BX, 1, gr + L.bq2);
    if (st) return st;
    hipLaunchKernelGGL(fno_dz1, dim3(nblk(BX * H)), dim3(FNO_T), 0, s, f->dy, f->params + L.Q2, f->Z1, f->dZ1, BX, H);
    const float* vL = f->Z[c.layers - 1];
    st = fno_gemm_launch(f, G(H, C, (int)BX, 1, cv(f->dZ1), 1, H, 0, cv(vL), C, 1, 0, gr + L.Q1, C, 1, 0), 2, 2);
    if (st) return st;
    st = colsum(f, f->dZ1, BX, H, gr + L.bq1);
    if (st) return st;
    float* dV = f->dVa;  // gradient w.r.t. the current layer's output
    float* dVn = f->dVb;
    st = fno_gemm_launch(f, G((int)BX, C, H, 1, cv(f->dZ1), H, 1, 0, cv(f->params + L.Q1), C, 1, 0, dV, C, 1, 0), 2, 2);
    if (st) return st;
    for (int l = c.layers - 1; l >= 0; l--) {
        const float* dz = dV;
        if (l + 1 < c.layers) {
            hipLaunchKernelGGL(fno_gelu_bwd, dim3(nblk(BX * C)), dim3(FNO_T), 0, s, dV, f->Z[l], f->dZ, BX * C);
            dz = f->dZ;
        }
        const float* v = f->V[l];
        // dW = dzᵀ v, db = Σ dz
        st = fno_gemm_launch(f, G(C, C, (int)BX, 1, cv(dz), 1, C, 0, cv(v), C, 1, 0, gr + L.W[l], C, 1, 0), 0, 0);
        if (st) return st;
        st = colsum(f, dz, BX, C, gr + L.b[l]);
        if (st) return st;
        // dOhat_b = Ginvᵀ dz_b
        st = fno_gemm_launch(f, G(Mo2, C, X, B, cv(f->Ginv), 1, Mo2, 0, cv(dz), C, 1, (long long)X * C, f->dOhat, C, 1,
                                  (long long)Mo2 * C), 1, 0);
        if (st) return st;
        // dRblk_m = Vblk_mᵀ dOblk_m
        st = fno_gemm_launch(f, G(2 * C, 2 * C, B, c.modes, cv(f->Vhat[l]), 1, (long long)Mo2 * C, 2 * C, cv(f->dOhat),
                                  (long long)Mo2 * C, 1, 2 * C, f->dRblk, 2 * C, 1, 4LL * C * C), 0, 0);
        if (st) return st;
        const long long nr = (long long)c.modes * C * C;
        hipLaunchKernelGGL(fno_drblk_to_dr, dim3(nblk(nr)), dim3(FNO_T), 0, s, f->dRblk, gr + L.Rre[l], gr + L.Rim[l], C,
                           c.modes);
        // dVblk_m = dOblk_m Rblk_mᵀ
        st = fno_gemm_launch(f, G(B, 2 * C, 2 * C, c.modes, cv(f->dOhat), (long long)Mo2 * C, 1, 2 * C,
                                  cv(f->Rblk + (long long)l * rb), 1, 2 * C, 4LL * C * C, f->dVhat, (long long)Mo2 * C, 1,
                                  2 * C), 0, 1);
        if (st) return st;
        // dv = dz W  +  Fᵀ dVhat
        st = fno_gemm_launch(f, G((int)BX, C, C, 1, cv(dz), C, 1, 0, cv(pb + L.W[l]), C, 1, 0, dVn, C, 1, 0), 0, 1);
        if (st) return st;
        st = fno_gemm_launch(f, G(X, C, Mo2, B, cv(f->F), 1, X, 0, cv(f->dVhat), C, 1, (long long)Mo2 * C, dVn, C, 1,
                                  (long long)X * C, 1.0f), 1, 0);
        if (st) return st;
        float* t = dV;
        dV = dVn;
        dVn = t;
    }
    // lift: dP = dv0ᵀ x, dbP = Σ dv0
    st = fno_gemm_launch(f, G(C, cin, (int)BX, 1, cv(dV), 1, C, 0, cv(f->xb), cin, 1, 0, gr + L.P, cin, 1, 0), 0, 0);
    if (st) return st;
    st = colsum(f, dV, BX, C, gr + L.bP);
    if (st) return st;
    FNOCK(hipGetLastError());
    return LH_OK;
}

static int fno_adam_step(lh_fno* f, const lh_adam* o, const float* lr_dev, int step_idx) {
    f->adam_t++;
    const float bc1 = 1.0f - powf(o->beta1, (float)f->adam_t);
    const float bc2 = 1.0f - powf(o->beta2, (float)f->adam_t);
    hipLaunchKernelGGL(fno_adamw, dim3(nblk(f->L.total)), dim3(FNO_T), 0, f->stream, f->params, f->grads, f->adam_m,
                       f->adam_v, f->pbf, f->L.total, lr_dev, step_idx, o->lr, o->beta1, o->beta2, o->eps,
                       o->weight_decay, bc1, bc2);
    const lh_fno_cfg& c = f->cfg;
    const long long rb = (long long)c.modes * 4 * c.channels * c.channels;
    for (int l = 0; l < c.layers; l++)
        hipLaunchKernelGGL(fno_build_rblk, dim3(nblk(rb)), dim3(FNO_T), 0, f->stream, f->params + f->L.Rre[l],
                           f->params + f->L.Rim[l], f->Rblk + (long long)l * rb, c.channels, c.modes);
    FNOCK(hipGetLastError());
    return LH_OK;
}

extern "C" {

int lh_fno_destroy(lh_fno* f) {
    if (!f) return LH_OK;
    hipSetDevice(f->device);
    if (f->stream) hipStreamSynchronize(f->stream);
    void* ptrs[] = {f->params, f->grads, f->adam_m, f->adam_v, f->pbf, f->Rblk, f->F, f->Ginv, f->xb, f->tb, f->wb,
                    f->data_w,
                    f->Ohat, f->Z1, f->Hh, f->y, f->dy, f->dVa, f->dVb, f->dZ, f->dOhat, f->dVhat, f->dZ1,
                    f->dRblk, f->work, f->part, f->loss_dev, f->data_x, f->data_t, f->idx_dev, f->lr_dev};
    for (void* p : ptrs)
        if (p) hipFree(p);
    for (int l = 0; l < FNO_MAXL; l++) {
        if (f->V[l]) hipFree(f->V[l]);
        if (f->Z[l]) hipFree(f->Z[l]);
        if (f->Vhat[l]) hipFree(f->Vhat[l]);
    }
    if (f->stream) hipStreamDestroy(f->stream);
    delete f->allocs;
    free(f);
    return LH_OK;
}

int lh_fno_create(lh_fno** out, const lh_fno_cfg* cfg) {
    if (!out || !cfg) return LH_E_ARG;
    *out = nullptr;
    const lh_fno_cfg c = *cfg;
    if (c.n_x < 64 || c.n_x % 64 || c.cin < 1 || c.channels < 32 || c.channels % 32 || c.modes < 1 ||
        2 * c.modes > c.n_x || c.layers < 1 || c.layers > FNO_MAXL || c.hidden < 32 || c.hidden % 32 ||
        c.max_batch < 1)
        return LH_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LH_E_NODEVICE;
    lh_fno* f = (lh_fno*)calloc(1, sizeof(lh_fno));
    if (!f) return LH_E_NOMEM;
    f->cfg = c;
    f->L = fno_layout(c);
    int st = LH_OK;
    f->allocs = new std::vector<std::pair<const char*, size_t>>();
    auto alloc = [&](void** p, size_t bytes) {
        if (!st && hipMalloc(p, bytes) != hipSuccess) st = LH_E_NOMEM;
        if (!st) f->allocs->push_back({(const char*)*p, bytes});
    };
    if (hipGetDevice(&f->device) != hipSuccess || hipStreamCreate(&f->stream) != hipSuccess) st = LH_E_HIP;
    const long long P = f->L.total, X = c.n_x, C = c.channels, Mo2 = 2 * c.modes, H = c.hidden, B = c.max_batch;
    const long long BX = B * X;
    alloc((void**)&f->params, 4 * P);
    alloc((void**)&f->grads, 4 * P);
    alloc((void**)&f->adam_m, 4 * P);
    alloc((void**)&f->adam_v, 4 * P);
    alloc((void**)&f->pbf, 2 * P);
    alloc((void**)&f->Rblk, 2 * (size_t)c.layers * c.modes * 4 * C * C);
    alloc((void**)&f->F, 2 * Mo2 * X);
    alloc((void**)&f->Ginv, 2 * Mo2 * X);
    alloc((void**)&f->xb, 4 * BX * c.cin);
    alloc((void**)&f->tb, 4 * BX);
    alloc((void**)&f->wb, 4 * BX);
    for (int l = 0; l < c.layers; l++) {
        alloc((void**)&f->V[l], 4 * BX * C);
        alloc((void**)&f->Z[l], 4 * BX * C);
        alloc((void**)&f->Vhat[l], 4 * B * Mo2 * C);
    }
    alloc((void**)&f->Ohat, 4 * B * Mo2 * C);
    alloc((void**)&f->Z1, 4 * BX * H);
    alloc((void**)&f->Hh, 4 * BX * H);
    alloc((void**)&f->y, 4 * BX);
    alloc((void**)&f->dy, 4 * BX);
    alloc((void**)&f->dVa, 4 * BX * C);
    alloc((void**)&f->dVb, 4 * BX * C);
    alloc((void**)&f->dZ, 4 * BX * C);
    alloc((void**)&f->dOhat, 4 * B * Mo2 * C);
    alloc((void**)&f->dVhat, 4 * B * Mo2 * C);
    alloc((void**)&f->dZ1, 4 * BX * H);
    alloc((void**)&f->dRblk, 4 * (size_t)c.modes * 4 * C * C);
    f->work_cap = 4LL << 20;  // split-K partials (floats)
    alloc((void**)&f->work, 4 * f->work_cap);
    long long pc = (BX + FNO_T - 1) / FNO_T;
    long long pc2 = ((BX + FNO_CS_ROWS - 1) / FNO_CS_ROWS) * (H > C ? H : C);
    f->part_cap = pc > pc2 ? pc : pc2;
    alloc((void**)&f->part, 4 * f->part_cap);
    f->loss_cap = 1;
    if (!st && hipMalloc(&f->loss_dev, 4) != hipSuccess) st = LH_E_NOMEM;  // not a GEMM operand
    if (!st) {
        // truncated DFT matrices (exact angles from the integer m·x mod n_x)
        std::vector<float> F(Mo2 * X), Gi(Mo2 * X);
        const long double tau = 6.283185307179586476925286766559005768L;
        for (int m = 0; m < c.modes; m++)
            for (int x = 0; x < X; x++) {
                const long long q = ((long long)m * x) % X;
                const long double ang = tau * (long double)q / (long double)X;
                const double cs = (double)cosl(ang), sn = (double)sinl(ang);
                const double cm = (m == 0) ? 1.0 : 2.0;
                F[(2 * m) * X + x] = (float)cs;
                F[(2 * m + 1) * X + x] = (float)(-sn);
                Gi[x * Mo2 + 2 * m] = (float)(cm * cs / (double)X);
                Gi[x * Mo2 + 2 * m + 1] = (float)(-cm * sn / (double)X);
            }
        float* tmp = nullptr;
        if (hipMalloc(&tmp, 4 * Mo2 * X) != hipSuccess) st = LH_E_NOMEM;
        if (!st && hipMemcpy(tmp, F.data(), 4 * Mo2 * X, hipMemcpyHostToDevice) != hipSuccess) st = LH_E_HIP;
        if (!st) hipLaunchKernelGGL(fno_cvt_bf16, dim3(nblk(Mo2 * X)), dim3(FNO_T), 0, f->stream, tmp, f->F, Mo2 * X);
        if (!st && hipStreamSynchronize(f->stream) != hipSuccess) st = LH_E_HIP;
        if (!st && hipMemcpy(tmp, Gi.data(), 4 * Mo2 * X, hipMemcpyHostToDevice) != hipSuccess) st = LH_E_HIP;
        if (!st) hipLaunchKernelGGL(fno_cvt_bf16, dim3(nblk(Mo2 * X)), dim3(FNO_T), 0, f->stream, tmp, f->Ginv, Mo2 * X);
        if (!st && hipStreamSynchronize(f->stream) != hipSuccess) st = LH_E_HIP;
        if (tmp) hipFree(tmp);
    }
    if (st) {
        lh_fno_destroy(f);
        return st;
    }
    *out = f;
    return LH_OK;
}

int lh_fno_param_count(const lh_fno* f, int64_t* n) {
    if (!f || !n) return LH_E_ARG;
    *n = f->L.total;
    return LH_OK;
}

int lh_fno_set_params(lh_fno* f, const float* params) {
    if (!f || !params) return LH_E_ARG;
    FNOCK(hipSetDevice(f->device));
    const size_t bytes = 4 * (size_t)f->L.total;
    FNOCK(hipMemcpy(f->params, params, bytes, hipMemcpyHostToDevice));
    FNOCK(hipMemset(f->adam_m, 0, bytes));
    FNOCK(hipMemset(f->adam_v, 0, bytes));
    f->adam_t = 0;
    refresh_bf16(f);
    FNOCK(hipGetLastError());
    FNOCK(hipStreamSynchronize(f->stream));
    return LH_OK;
}

int lh_fno_get_params(lh_fno* f, float* params) {
    if (!f || !params) return LH_E_ARG;
    FNOCK(hipSetDevice(f->device));
    FNOCK(hipStreamSynchronize(f->stream));
    FNOCK(hipMemcpy(params, f->params, 4 * (size_t)f->L.total, hipMemcpyDeviceToHost));
    return LH_OK;
}

int lh_fno_forward(lh_fno* f, const float* x, int batch, float* y) {
    if (!f || !x || !y || batch < 1) return LH_E_ARG;
    FNOCK(hipSetDevice(f->device));
    const lh_fno_cfg& c = f->cfg;
    for (int b0 = 0; b0 < batch; b0 += c.max_batch) {
        const int B = (batch - b0 < c.max_batch) ? batch - b0 : c.max_batch;
        const long long BX = (long long)B * c.n_x;
        FNOCK(hipMemcpyAsync(f->xb, x + (long long)b0 * c.n_x * c.cin, 4 * BX * c.cin, hipMemcpyHostToDevice, f->stream));
        int st = fno_forward_dev(f, B);
        if (st) return st;
        FNOCK(hipMemcpyAsync(y + (long long)b0 * c.n_x, f->y, 4 * BX, hipMemcpyDeviceToHost, f->stream));
        FNOCK(hipStreamSynchronize(f->stream));
    }
    return LH_OK;
}

int lh_fno_loss_grad(lh_fno* f, const float* x, const float* target, const float* weight, int batch, float* loss,
                     float* grad) {
    if (!f || !x || !target || !loss || batch < 1 || batch > f->cfg.max_batch) return LH_E_ARG;
    FNOCK(hipSetDevice(f->device));
    const lh_fno_cfg& c = f->cfg;
    const long long BX = (long long)batch * c.n_x;
    FNOCK(hipMemcpyAsync(f->xb, x, 4 * BX * c.cin, hipMemcpyHostToDevice, f->stream));
    FNOCK(hipMemcpyAsync(f->tb, target, 4 * BX, hipMemcpyHostToDevice, f->stream));
    f->wcur = nullptr;
    if (weight) {
        FNOCK(hipMemcpyAsync(f->wb, weight, 4 * BX, hipMemcpyHostToDevice, f->stream));
        f->wcur = f->wb;
    }
    int st = fno_forward_dev(f, batch);
    if (!st) st = fno_backward_dev(f, batch, f->loss_dev);
    if (st) return st;
    FNOCK(hipMemcpyAsync(loss, f->loss_dev, 4, hipMemcpyDeviceToHost, f->stream));
    if (grad) FNOCK(hipMemcpyAsync(grad, f->grads, 4 * (size_t)f->L.total, hipMemcpyDeviceToHost, f->stream));
    FNOCK(hipStreamSynchronize(f->stream));
    return LH_OK;
}

int lh_fno_set_data(lh_fno* f, const float* x, const float* target, const float* weight, int n) {
    if (!f || !x || !target || n < 1) return LH_E_ARG;
    FNOCK(hipSetDevice(f->device));
    const lh_fno_cfg& c = f->cfg;
    for (void* old : {(void*)f->data_x, (void*)f->data_t}) {
        if (!old) continue;
        for (size_t i = 0; i < f->allocs->size(); i++)
            if ((*f->allocs)[i].first == (const char*)old) {
                f->allocs->erase(f->allocs->begin() + i);
                break;
            }
        FNOCK(hipFree(old));
    }
    f->data_x = f->data_t = nullptr;
    f->data_n = 0;
    const size_t nx = (size_t)n * c.n_x * c.cin, nt = (size_t)n * c.n_x;
    FNOCK(hipMalloc(&f->data_x, 4 * nx));
    f->allocs->push_back({(const char*)f->data_x, 4 * nx});
    FNOCK(hipMalloc(&f->data_t, 4 * nt));
    f->allocs->push_back({(const char*)f->data_t, 4 * nt});
    FNOCK(hipMemcpy(f->data_x, x, 4 * nx, hipMemcpyHostToDevice));
    FNOCK(hipMemcpy(f->data_t, target, 4 * nt, hipMemcpyHostToDevice));
    if (f->data_w) FNOCK(hipFree(f->data_w));
    f->data_w = nullptr;
    if (weight) {
        FNOCK(hipMalloc(&f->data_w, 4 * nt));
        FNOCK(hipMemcpy(f->data_w, weight, 4 * nt, hipMemcpyHostToDevice));
    }
    f->data_n = n;
    return LH_OK;
}

int lh_fno_train(lh_fno* f, const int32_t* batch_idx, int n_steps, int batch, const lh_adam* opt,
                 const float* lr_sched, float* losses) {
    if (!f || !batch_idx || n_steps < 1 || batch < 1 || batch > f->cfg.max_batch || !opt || !f->data_x) return LH_E_ARG;
    const long long nidx = (long long)n_steps * batch;
    for (long long i = 0; i < nidx; i++)
        if (batch_idx[i] < 0 || batch_idx[i] >= f->data_n) return LH_E_ARG;
    FNOCK(hipSetDevice(f->device));
    const lh_fno_cfg& c = f->cfg;
    if (nidx > f->idx_cap) {
        if (f->idx_dev) FNOCK(hipFree(f->idx_dev));
        f->idx_dev = nullptr;
        FNOCK(hipMalloc(&f->idx_dev, 4 * nidx));
        f->idx_cap = nidx;
    }
    if (n_steps > f->loss_cap) {
        if (f->loss_dev) FNOCK(hipFree(f->loss_dev));
        f->loss_dev = nullptr;
        FNOCK(hipMalloc(&f->loss_dev, 4 * (size_t)n_steps));
        f->loss_cap = n_steps;
    }
    if (f->lr_dev) FNOCK(hipFree(f->lr_dev));
    f->lr_dev = nullptr;
    if (lr_sched) {
        FNOCK(hipMalloc(&f->lr_dev, 4 * (size_t)n_steps));
        FNOCK(hipMemcpy(f->lr_dev, lr_sched, 4 * (size_t)n_steps, hipMemcpyHostToDevice));
    }
    FNOCK(hipMemcpy(f->idx_dev, batch_idx, 4 * nidx, hipMemcpyHostToDevice));
    const long long ng = (long long)batch * c.n_x * (c.cin > 1 ? c.cin : 1);
    for (int s = 0; s < n_steps; s++) {
        hipLaunchKernelGGL(fno_gather, dim3(nblk(ng)), dim3(FNO_T), 0, f->stream, f->data_x, f->data_t, f->data_w,
                           f->idx_dev + (long long)s * batch, f->xb, f->tb, f->wb, batch, c.n_x, c.cin);
        f->wcur = f->data_w ? f->wb : nullptr;
        int st = fno_forward_dev(f, batch);
        if (!st) st = fno_backward_dev(f, batch, f->loss_dev + s);
        if (!st) st = fno_adam_step(f, opt, f->lr_dev, s);
        if (st) return st;
    }
    if (losses) FNOCK(hipMemcpyAsync(losses, f->loss_dev, 4 * (size_t)n_steps, hipMemcpyDeviceToHost, f->stream));
    FNOCK(hipStreamSynchronize(f->stream));
    return LH_OK;
}

}  // extern "C"
