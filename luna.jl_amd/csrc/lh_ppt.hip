// lh_ppt.hip -- PPT ionisation-rate table on the GPU (set-up of the plasma
// response: makePPTcache, src/Ionisation.jl:185-200, with ionrate_fun_PPT,
// :317-390, and φ, :403-429).  One thread per field value; each runs the
// reference's converge_series (src/Maths.jl:721-736) over the PPT sum on its
// own, so a 2^16-point table is one launch instead of a host loop.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/luna_hip.h"

namespace {

constexpr double kInvSqrtPi = 0.56418958354775628695;  // 1/√π
constexpr double kRybH = 0.25;                         // Rybicki sampling step
constexpr int kRybN = 13;                              // sample pairs
constexpr int kPptBlock = 256;
constexpr long long kMaxTerms = 1LL << 24;             // a series longer than this is reported

// Dawson's integral D(x) = exp(-x²)∫₀ˣ exp(t²) dt for x ≥ 0.
// x < 0.2: its Taylor series Σ (-2x²)^k x / (2k+1)!!.  Otherwise Rybicki's
// sampling formula D(x) = lim_{h→0} π^{-1/2} Σ_{n odd} exp(-(x - nh)²)/n
// (G. B. Rybicki, Computers in Physics 3, 85 (1989)), summed around the
// nearest even multiple of h; with h = 0.25 the sampling error is of order
// exp(-(π/2h)²) ≈ 7e-18 and 13 pairs leave out terms below exp(-42).
// c[i] = exp(-((2i+1)h)²).
__device__ __forceinline__ double dawson(double x, const double* c) {
    if (x < 0.2) {
        const double x2 = x * x;
        double term = x, sum = x;
        for (int k = 0; k < 12; k++) {
            term *= -2.0 * x2 / (double)(2 * k + 3);
            sum += term;
        }
        return sum;
    }
    const double n0 = 2.0 * rint(0.5 * x / kRybH);
    const double xp = x - n0 * kRybH;
    double e1 = exp(2.0 * xp * kRybH);
    const double e2 = e1 * e1;
    double d1 = n0 + 1.0, d2 = d1 - 2.0, sum = 0.0;
    for (int i = 0; i < kRybN; i++) {
        sum += c[i] * (e1 / d1 + 1.0 / (d2 * e1));
        d1 += 2.0;
        d2 -= 2.0;
        e1 *= e2;
    }
    return kInvSqrtPi * exp(-xp * xp) * sum;
}

// φ_m(x) = exp(-x²)∫₀ˣ (x² - y²)^|m| exp(y²) dy for |m| ≤ 1 (PPT, src/Ionisation.jl:403-429).
// φ_0 = D(x).  φ_1 = (x² + 1/2) D(x) - x/2 (integrating y² exp(y²) by parts),
// which cancels to ~1/(2x) for large x: above x = 10 its asymptotic series
// φ_1 = (1/2x) Σ_k (c_{k+1} + c_k/2) x^{-2k}, c_k = (2k-1)!!/2^k, is used
// (15 terms; the first omitted one is below 1e-16 of the sum there).
__device__ __forceinline__ double ppt_phi(int mabs, double x, const double* c) {
    if (mabs == 0) return dawson(x, c);
    if (x < 10.0) return (x * x + 0.5) * dawson(x, c) - 0.5 * x;
    const double ix2 = 1.0 / (x * x);
    double ck = 1.0, p = 1.0, sum = 0.0;
    for (int k = 0; k < 15; k++) {
        const double ck1 = ck * (2.0 * k + 1.0) * 0.5;
        sum += (ck1 + 0.5 * ck) * p;
        ck = ck1;
        p *= ix2;
    }
    return sum / (2.0 * x);
}

__device__ __forceinline__ double fact(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; i++) f *= (double)i;
    return f;
}

// ionrate_fun_PPT's rate at field E (src/Ionisation.jl:330-384), the same
// sequence of operations as the reference.
__device__ double ppt_rate_at(const lh_ppt_desc& p, double E, const double* c, int* status) {
    const double au_energy = p.au_energy, au_time = p.au_time, au_efield = p.au_efield;
    const double Ip_au = (p.ionpot + p.dalpha / 2 * (E * E)) / au_energy;
    const double ns = p.Z / sqrt(2 * Ip_au);
    const double ls = ns - 1;
    const double Cnl2 = (p.cnl > 0) ? p.cnl * p.cnl : pow(2.0, 2 * ns) / (ns * tgamma(ns + ls + 1) * tgamma(ns - ls));
    const double w0 = 2 * M_PI * p.c / p.lambda0;
    const double w0_au = au_time * w0;
    const double E0_au = pow(2 * Ip_au, 1.5);
    const double E_au = fabs(E) / au_efield;
    const double g = w0_au * sqrt(2 * Ip_au) / E_au;
    const double g2 = g * g;
    const double beta = 2 * g / sqrt(1 + g2);
    const double alpha = 2 * (asinh(g) - g / sqrt(1 + g2));
    const double Up_au = (E_au * E_au) / (4 * (w0_au * w0_au));
    const double v = (Ip_au + Up_au) / w0_au;
    double ret = 0.0;
    const int l = p.l, m0 = p.msum ? -l : 0, m1 = p.msum ? l : 0;
    for (int m = m0; m <= m1; m++) {
        const int mabs = abs(m);
        const double flm = ((2 * l + 1) * fact(l + mabs)) / (pow(2.0, (double)mabs) * fact(mabs) * fact(l - mabs));
        double lret = 4 * sqrt(2.0) / M_PI * Cnl2;
        lret *= pow(2 * E0_au / (E_au * sqrt(1 + g2)), 2 * ns - mabs - 1.5);
        lret *= flm / fact(mabs);
        lret *= exp(-2 * v * (asinh(g) - g * sqrt(1 + g2) / (1 + 2 * g2)));
        lret *= Ip_au * g2 / (1 + g2);
        if (!p.cycle_average) lret *= sqrt(M_PI * E0_au / (3 * E_au));
        // converge_series(0, n0 = ceil(v), rtol = sum_tol)
        double n = ceil(v), x0 = 0.0, x1 = 0.0;
        bool done = false;
        long long it = 0;
        while (!done && it < kMaxTerms) {
            const double diff = n - v;
            x1 = x0 + exp(-alpha * diff) * ppt_phi(mabs, sqrt(beta * diff), c);
            if (2 * fabs(x1 - x0) / fabs(x1 + x0) < p.sum_tol) done = true;
            n += 1;
            x0 = x1;
            it++;
        }
        if (!done) *status = LH_E_NONFINITE;
        lret *= x1;
        ret += p.occupancy * lret;
    }
    if (p.alpha_ion_au != 0) ret *= exp(-2 * p.alpha_ion_au * E_au);
    return ret / au_time;
}

__global__ __launch_bounds__(kPptBlock) void lh_ppt_kernel(lh_ppt_desc p, const double* __restrict__ E, int n,
                                                         double* __restrict__ rate, int* status) {
    __shared__ double c[kRybN];
    if (threadIdx.x < kRybN) {
        const double a = (2 * threadIdx.x + 1) * kRybH;
        c[threadIdx.x] = exp(-a * a);
    }
    __syncthreads();
    const int i = blockIdx.x * kPptBlock + threadIdx.x;
    if (i >= n) return;
    int st = LH_OK;
    const double r = ppt_rate_at(p, E[i], c, &st);
    rate[i] = r;
    if (st != LH_OK) atomicMax(status, st);
}

}  // namespace

#define PPT_CK(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            fprintf(stderr, "luna_hip: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            st = (e_ == hipErrorOutOfMemory) ? LH_E_NOMEM : LH_E_HIP;                \
            goto done;                                                               \
        }                                                                            \
    } while (0)

extern "C" int lh_ppt_table(const lh_ppt_desc* p, int n, const double* E, double* rate, double* ms) {
    if (!p || !E || !rate || n <= 0) return LH_E_ARG;
    if (p->l < 0 || p->l > 1 || !(p->sum_tol > 0) || !(p->lambda0 > 0) || !(p->ionpot > 0)) return LH_E_ARG;
    for (int i = 0; i < n; i++)
        if (!(E[i] > 0)) return LH_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LH_E_NODEVICE;
    int st = LH_OK, kst = LH_OK;
    double *dE = nullptr, *dR = nullptr;
    int* dS = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    const size_t bytes = sizeof(double) * (size_t)n;
    PPT_CK(hipMalloc(&dE, bytes));
    PPT_CK(hipMalloc(&dR, bytes));
    PPT_CK(hipMalloc(&dS, sizeof(int)));
    PPT_CK(hipMemcpy(dE, E, bytes, hipMemcpyHostToDevice));
    PPT_CK(hipMemcpy(dS, &kst, sizeof(int), hipMemcpyHostToDevice));
    PPT_CK(hipEventCreate(&t0));
    PPT_CK(hipEventCreate(&t1));
    PPT_CK(hipEventRecord(t0, nullptr));
    hipLaunchKernelGGL(lh_ppt_kernel, dim3((n + kPptBlock - 1) / kPptBlock), dim3(kPptBlock), 0, nullptr, *p, dE, n,
                       dR, dS);
    PPT_CK(hipGetLastError());
    PPT_CK(hipEventRecord(t1, nullptr));
    PPT_CK(hipMemcpy(rate, dR, bytes, hipMemcpyDeviceToHost));
    PPT_CK(hipMemcpy(&kst, dS, sizeof(int), hipMemcpyDeviceToHost));
    if (ms) {
        float f = 0.0f;
        PPT_CK(hipEventElapsedTime(&f, t0, t1));
        *ms = f;
    }
    st = kst;
done:
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (dE) (void)hipFree(dE);
    if (dR) (void)hipFree(dR);
    if (dS) (void)hipFree(dS);
    return st;
}
