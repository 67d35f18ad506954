// nt_probe.hip -- kernels of the Buttiker-probe solve (negf/lesser.f90:12-45,
// 1199-1503): probe self-energies and shifts, grouped Frobenius products and
// traces of the full Green's function blocks, and the per-point combination
// into probe currents, Jacobian, contact currents and traces.
//
// The probes' self-energies are multiples of the identity, so every quantity
// the reference's lesser recursion produces is a sum over blocks G_ij of the
// full retarded function: Tr Gn_i = sum_j s_j |G_ij|_F^2 + contact terms,
// d Tr Gn_i / d T_j = ds_j |G_ij|_F^2.  The reductions here are HBM-bound
// (each block read once); the block products run on the batched GEMM.
#include "nt_kernels.h"

namespace nt {

namespace {

constexpr int NT = 256;
constexpr double BT_PI = 3.141592653589793238;
constexpr double BT_KB = 1.380649e-23;
constexpr double BT_HBAR = 1.0545718e-34;

inline unsigned blocks_for(long work, int per_block = NT)
{
    long b = (work + per_block - 1) / per_block;
    return (unsigned)(b < 1 ? 1 : b);
}

// bose_einstein and dbosedT (negf/util.f90:987-1013), same operation order
__device__ inline double bose(double freq, double T)
{
    const double x = exp(BT_HBAR * freq * 1.0e12 * 2.0 * BT_PI / BT_KB / T);
    return 1.0 / (x - 1.0);
}
__device__ inline double dbose(double freq, double T)
{
    const double x = exp(BT_HBAR * freq * 1.0e12 * 2.0 * BT_PI / BT_KB / T);
    return x / ((x - 1.0) * (x - 1.0)) * BT_HBAR * freq * 1.0e12 * 2.0 * BT_PI / BT_KB / (T * T);
}

// Im sigma_w and Im dsigma_w/dT of one layer (negf/lesser.f90:31-37); both are
// purely imaginary.
__device__ inline void probe_sigma(double E, double T, double B, double C, double* s_im, double* ds_im)
{
    const double tau = 1.0 / (B * E * T * exp(-C / T));
    *s_im = -2.0 * sqrt(E) / tau;
    *ds_im = -2.0 * sqrt(E) * (B * E * exp(-C / T) * (1.0 + C / T));
}

// zl[i][p] = zd[p] - sigma_w(i, p); gam[i][p] = i (sigma - sigma^*) = -2 Im sigma,
// dgam likewise from dsigma/dT.  Layer-major (L x P).
__global__ void k_probe_shifts(const double* E, const double* zd, const double* temp, int L, int P, double B,
                               double C, int lifetime, double* zl, double* gam, double* dgam)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (p >= P || i >= L) return;
    double s, ds;
    probe_sigma(E[p], temp[i], B, C, &s, &ds);
    if (!lifetime) s = 0.0;             // negf/lesser.f90:40-42 zeroes sigma_w only
    const long q = (long)i * P + p;
    zl[2 * q] = zd[2 * p];
    zl[2 * q + 1] = zd[2 * p + 1] - s;
    gam[q] = -2.0 * s;
    dgam[q] = -2.0 * ds;
}

// out[p, g] = Tr(X^+ Y) = sum_ab conj(X_ab) Y_ab over m x n blocks, as (re, im)
// at out + 2 (p * ostride + g * ogstride); grid (batch, groups).
template <typename T>
__global__ __launch_bounds__(NT) void k_frob_dot(int m, int n, BMat<T> X, BMat<T> Y, double* out, long ostride,
                                                 long ogstride)
{
    const int p = blockIdx.x, g = blockIdx.y;
    const T* __restrict__ x = bptr(X, p, g);
    const T* __restrict__ y = bptr(Y, p, g);
    double re = 0.0, im = 0.0;
    const long tot = (long)m * n;
    for (long e = threadIdx.x; e < tot; e += NT) {
        const int i = (int)(e % m), j = (int)(e / m);
        const T* a = x + 2 * ((long)i + (long)j * X.ld);
        const T* b = y + 2 * ((long)i + (long)j * Y.ld);
        const double ar = a[0], ai = a[1], br = b[0], bi = b[1];
        re += ar * br + ai * bi;
        im += ar * bi - ai * br;
    }
    __shared__ double red[2][NT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off);
        im += __shfl_xor(im, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = re;
        red[1][threadIdx.x >> 6] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sr = 0.0, si = 0.0;
        for (int w = 0; w < NT / 64; ++w) {
            sr += red[0][w];
            si += red[1][w];
        }
        double* o = out + 2 * ((long)p * ostride + (long)g * ogstride);
        o[0] = sr;
        o[1] = si;
    }
}

// out[p, g] = Tr X (n x n), as (re, im); grid (batch, groups)
template <typename T>
__global__ __launch_bounds__(NT) void k_trace_diag(int n, BMat<T> X, double* out, long ostride, long ogstride)
{
    const int p = blockIdx.x, g = blockIdx.y;
    const T* __restrict__ x = bptr(X, p, g);
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) {
        const T* a = x + 2 * ((long)i + (long)i * X.ld);
        re += a[0];
        im += a[1];
    }
    __shared__ double red[2][NT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off);
        im += __shfl_xor(im, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = re;
        red[1][threadIdx.x >> 6] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sr = 0.0, si = 0.0;
        for (int w = 0; w < NT / 64; ++w) {
            sr += red[0][w];
            si += red[1][w];
        }
        double* o = out + 2 * ((long)p * ostride + (long)g * ogstride);
        o[0] = sr;
        o[1] = si;
    }
}

// One thread per (point p, layer i).  acc holds per point (complex pairs):
// n[L*L] (n[i*L+j] = |G_ij|^2), cL[L], cR[L], m[L], r[L], trG[L], then
// aL, aR, xLL, xLR, xRL, xRR.
__global__ void k_probe_combine(ProbeCombineArgs a)
{
    const int p = blockIdx.x;
    const int i = blockIdx.y * blockDim.x + threadIdx.x;
    const int L = a.L, P = a.P;
    if (p >= P || i >= L) return;
    const double* acc = a.acc + 2 * (long)p * a.acc_stride;
    const double* n = acc;
    const double* cL = acc + 2 * (long)L * L;
    const double* cR = cL + 2 * L;
    const double* mm = cR + 2 * L;
    const double* rr = mm + 2 * L;
    const double* trG = rr + 2 * L;
    const double* misc = trG + 2 * L;
    const double freq = sqrt(a.E[p]);
    const double bL = bose(freq, a.t_left), bR = bose(freq, a.t_right);
    const bool first = a.first_mask && a.first_mask[p];
    // Tr Gn_i = sum_j s_j |G_ij|^2 + bL Tr(G_i0 gf G_i0^+) + bR Tr(G_iN gl G_iN^+)
    double gn = 0.0;
    for (int j = 0; j < L; ++j) {
        const double sj = a.gam[(long)j * P + p] * bose(freq, a.temp[j]);
        gn += sj * n[2 * ((long)i * L + j)];
    }
    gn += bL * cL[2 * i] + bR * cR[2 * i];
    const double gi = a.gam[(long)i * P + p], dgi = a.dgam[(long)i * P + p];
    const double Ti = a.temp[i];
    const double trA = -2.0 * trG[2 * i + 1];          // Tr i (G - G^+) = -2 Im Tr G
    const long po = (long)(a.p0 + p);
    double flux = -gi * gn + gi * trA * bose(freq, Ti);
    a.gn_sum[po * L + i] = first ? 0.0 : gn;
    a.a_sum[po * L + i] = first ? 0.0 : trA;
    a.flux[po * L + i] = first ? 0.0 : flux;
    // column-major per point: jac[i + j L] = d flux_i / d T_j
    double* jac = a.jac + po * L * L;
    for (int j = 0; j < L; ++j) {
        const double gj = a.gam[(long)j * P + p], dgj = a.dgam[(long)j * P + p];
        const double dsj = gj * dbose(freq, a.temp[j]) + dgj * bose(freq, a.temp[j]);
        double v = -gi * dsj * n[2 * ((long)i * L + j)];
        if (i == j) v += gi * trA * dbose(freq, Ti) + dgi * trA * bose(freq, Ti);
        jac[i + (long)j * L] = first ? 0.0 : v;
    }
    if (i == 0) {
        // contact currents (negf/lesser.f90:1336-1337, 1400-1401, 1464): kept at ie = 1
        double sl = 0.0, sr = 0.0;
        for (int j = 0; j < L; ++j) {
            const double sj = a.gam[(long)j * P + p] * bose(freq, a.temp[j]);
            sl += sj * mm[2 * j];
            sr += sj * rr[2 * j];
        }
        const double aL = -2.0 * misc[1], aR = -2.0 * misc[3];
        const double gnl = sl + bL * misc[4] + bR * misc[6];
        const double gnr = sr + bL * misc[8] + bR * misc[10];
        a.flux_left[po] = (aL * bL - gnl) * freq;
        a.flux_right[po] = (aR * bR - gnr) * freq;
    }
}

}  // namespace

hipError_t launch_probe_shifts(const double* E, const double* zd, const double* temp, int L, int P, double B,
                               double C, int lifetime, double* zl, double* gam, double* dgam, hipStream_t s)
{
    if (P <= 0 || L <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe_shifts, dim3(blocks_for(P), L), dim3(NT), 0, s, E, zd, temp, L, P, B, C, lifetime, zl,
                       gam, dgam);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_frob_dot(int m, int n, BMat<T> X, BMat<T> Y, double* out, long ostride, long ogstride, int batch,
                           int groups, hipStream_t s)
{
    if (batch <= 0 || groups <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_frob_dot<T>, dim3(batch, groups), dim3(NT), 0, s, m, n, X, Y, out, ostride, ogstride);
    return hipGetLastError();
}
template hipError_t launch_frob_dot<double>(int, int, BMat<double>, BMat<double>, double*, long, long, int, int,
                                            hipStream_t);

template <typename T>
hipError_t launch_trace_diag(int n, BMat<T> X, double* out, long ostride, long ogstride, int batch, int groups,
                             hipStream_t s)
{
    if (batch <= 0 || groups <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_diag<T>, dim3(batch, groups), dim3(NT), 0, s, n, X, out, ostride, ogstride);
    return hipGetLastError();
}
template hipError_t launch_trace_diag<double>(int, BMat<double>, double*, long, long, int, int, hipStream_t);

hipError_t launch_probe_combine(const ProbeCombineArgs& a, hipStream_t s)
{
    if (a.P <= 0 || a.L <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe_combine, dim3(a.P, blocks_for(a.L, 64)), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace nt
