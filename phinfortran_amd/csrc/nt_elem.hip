// nt_elem.hip -- small batched element-wise and reduction kernels of the
// transport path: energy shifts, decimation start and convergence bookkeeping,
// Gamma = i(Sigma - Sigma^+), adjoints, Caroli trace of a product, surface DOS.
// All are HBM-bound and touch each matrix once.
#include "nt_kernels.h"

namespace nt {

namespace {

constexpr int NT = 256;

inline unsigned blocks_for(long work, int per_block = NT)
{
    long b = (work + per_block - 1) / per_block;
    return (unsigned)(b < 1 ? 1 : b);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_copy_op(int m, int n, BMat<T> in, BMatW<T> out, int dagger, T sign,
                                                const int* active, int batch)
{
    // out is (dagger ? n x m : m x n); grid.y = item
    const int p = blockIdx.y;
    if (active && !active[p]) return;
    const T* __restrict__ A = bptr(in, p, 0);
    T* __restrict__ O = bptr(out, p, 0);
    const long e = (long)blockIdx.x * NT + threadIdx.x;
    if (dagger) {
        // O(j, i) = conj(A(i, j)); iterate over O column-major
        if (e >= (long)m * n) return;
        const int j = (int)(e % n), i = (int)(e / n);
        const T* a = A + 2 * ((long)i + (long)j * in.ld);
        T* o = O + 2 * ((long)j + (long)i * out.ld);
        o[0] = sign * a[0];
        o[1] = -sign * a[1];
    } else {
        if (e >= (long)m * n) return;
        const int i = (int)(e % m), j = (int)(e / m);
        const T* a = A + 2 * ((long)i + (long)j * in.ld);
        T* o = O + 2 * ((long)i + (long)j * out.ld);
        o[0] = sign * a[0];
        o[1] = sign * a[1];
    }
}

template <typename Ti, typename To>
__global__ __launch_bounds__(NT) void k_convert(int m, int n, BMat<Ti> in, BMatW<To> out, int dagger, To sign, int batch)
{
    const int p = blockIdx.y;
    const Ti* __restrict__ A = bptr(in, p, 0);
    To* __restrict__ O = bptr(out, p, 0);
    const long e = (long)blockIdx.x * NT + threadIdx.x;
    if (e >= (long)m * n) return;
    if (dagger) {
        const int j = (int)(e % n), i = (int)(e / n);
        const Ti* a = A + 2 * ((long)i + (long)j * in.ld);
        To* o = O + 2 * ((long)j + (long)i * out.ld);
        o[0] = sign * (To)a[0];
        o[1] = -sign * (To)a[1];
    } else {
        const int i = (int)(e % m), j = (int)(e / m);
        const Ti* a = A + 2 * ((long)i + (long)j * in.ld);
        To* o = O + 2 * ((long)i + (long)j * out.ld);
        o[0] = sign * (To)a[0];
        o[1] = sign * (To)a[1];
    }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_gamma(int n, BMat<T> S, BMatW<T> G, T scale, const int* active, int batch)
{
    const int p = blockIdx.y;
    if (active && !active[p]) return;
    const long e = (long)blockIdx.x * NT + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e % n), j = (int)(e / n);
    const T* __restrict__ A = bptr(S, p, 0);
    const T* a = A + 2 * ((long)i + (long)j * S.ld);
    const T* b = A + 2 * ((long)j + (long)i * S.ld);
    const T dr = a[0] - b[0];     // S_ij - conj(S_ji)
    const T di = a[1] + b[1];
    T* o = bptr(G, p, 0) + 2 * ((long)i + (long)j * G.ld);
    o[0] = -scale * di;           // i * (dr + i di) = -di + i dr
    o[1] = scale * dr;
}

template <typename T>
__global__ __launch_bounds__(NT) void k_shift_sub(int n, const T* z, BMat<T> A, BMat<T> B, BMatW<T> out,
                                                  const int* active, int batch)
{
    const int p = blockIdx.y;
    if (active && !active[p]) return;
    const long e = (long)blockIdx.x * NT + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e % n), j = (int)(e / n);
    const T* a = bptr(A, p, 0) + 2 * ((long)i + (long)j * A.ld);
    T vr = (i == j ? z[2 * p] : (T)0) - a[0];
    T vi = (i == j ? z[2 * p + 1] : (T)0) - a[1];
    if (B.base) {
        const T* b = bptr(B, p, 0) + 2 * ((long)i + (long)j * B.ld);
        vr = vr - b[0];
        vi = vi - b[1];
    }
    T* o = bptr(out, p, 0) + 2 * ((long)i + (long)j * out.ld);
    o[0] = vr;
    o[1] = vi;
}

template <typename T>
__global__ __launch_bounds__(NT) void k_dec_init(int n, BMat<T> H00, BMat<T> V, int alpha_dagger, BMatW<T> e1,
                                                 BMatW<T> es, BMatW<T> ab, int batch)
{
    const int p = blockIdx.y;
    const long e = (long)blockIdx.x * NT + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e % n), j = (int)(e / n);
    const T* h = bptr(H00, p, 0) + 2 * ((long)i + (long)j * H00.ld);
    const T* v = bptr(V, p, 0) + 2 * ((long)i + (long)j * V.ld);      // V(i,j)
    const T* vt = bptr(V, p, 0) + 2 * ((long)j + (long)i * V.ld);     // V(j,i)
    T* o1 = bptr(e1, p, 0) + 2 * ((long)i + (long)j * e1.ld);
    T* o2 = bptr(es, p, 0) + 2 * ((long)i + (long)j * es.ld);
    o1[0] = h[0]; o1[1] = h[1];
    o2[0] = h[0]; o2[1] = h[1];
    T* al = bptr(ab, p, 0) + 2 * ((long)i + (long)j * ab.ld);
    T* be = bptr(ab, p, 1) + 2 * ((long)i + (long)j * ab.ld);
    // alpha = V^+ and beta = V, or alpha = V and beta = V^+
    const T dr = vt[0], dim = -vt[1];   // (V^+)(i,j) = conj(V(j,i))
    if (alpha_dagger) {
        al[0] = dr; al[1] = dim;
        be[0] = v[0]; be[1] = v[1];
    } else {
        al[0] = v[0]; al[1] = v[1];
        be[0] = dr; be[1] = dim;
    }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_trace_prod(int n1, int n2, BMat<T> X, BMat<T> Y, double* out, int out_stride,
                                                   const int* active, int batch)
{
    const int p = blockIdx.x;
    if (active && !active[p]) return;
    const T* __restrict__ x = bptr(X, p, 0);
    const T* __restrict__ y = bptr(Y, p, 0);
    double acc = 0.0;
    // Re sum_{i,j} X(i,j) Y(j,i); thread over i (coalesced on X), loop j
    for (long e = threadIdx.x; e < (long)n1 * n2; e += NT) {
        const int i = (int)(e % n1), j = (int)(e / n1);
        const T* a = x + 2 * ((long)i + (long)j * X.ld);
        const T* b = y + 2 * ((long)j + (long)i * Y.ld);
        acc += (double)a[0] * (double)b[0] - (double)a[1] * (double)b[1];
    }
    __shared__ double red[NT / 64];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; ++w) s += red[w];
        out[(long)p * out_stride] = s;
    }
}

template <typename T>
// out[p] = coef sqrt(E[p]) Im Tr g[p] / pi; without E (null), coef Im Tr g[p] / pi
__global__ __launch_bounds__(NT) void k_trace_im(int n, BMat<T> g, const double* E, double coef, double* out, int batch)
{
    const int p = blockIdx.x;
    const T* __restrict__ x = bptr(g, p, 0);
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) acc += (double)x[2 * ((long)i + (long)i * g.ld) + 1];
    __shared__ double red[NT / 64];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; ++w) s += red[w];
        out[p] = coef * (E ? sqrt(E[p]) : 1.0) * s / 3.14159265358979323846;
    }
}

__global__ void k_make_shifts(const double* E, int batch, double eta, double eta0, int eta_mode, double f,
                              double* zl, double* zr, double* zd, float* zl32, float* zr32, float* zd32)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= batch) return;
    const double e = E[p];
    const double et = eta_mode == 0 ? e * eta + eta0 : eta;
    const double ed = f * et;
    if (zl) { zl[2 * p] = e; zl[2 * p + 1] = -et; }
    if (zr) { zr[2 * p] = e; zr[2 * p + 1] = et; }
    if (zd) { zd[2 * p] = e; zd[2 * p + 1] = ed; }
    if (zl32) { zl32[2 * p] = (float)e; zl32[2 * p + 1] = (float)-et; }
    if (zr32) { zr32[2 * p] = (float)e; zr32[2 * p + 1] = (float)et; }
    if (zd32) { zd32[2 * p] = (float)e; zd32[2 * p + 1] = (float)ed; }
}

__global__ void k_dec_step_update(int* active, unsigned long long* maxabs, int* iters, double conv, int* count,
                                  int* clear, int batch)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int still = 0;
    if (p < batch && active[p]) {
        const double m = __longlong_as_double((long long)maxabs[p]);
        if (iters) iters[p] += 1;
        still = (m > conv) ? 1 : 0;     // inf (NaN marker) stays active and hits the cap
        active[p] = still;
    }
    if (p < batch) maxabs[p] = 0ull;
    if (p == 0) *clear = 0;
    for (int off = 32; off > 0; off >>= 1) still += __shfl_xor(still, off);
    if ((threadIdx.x & 63) == 0 && still) atomicAdd(count, still);
}

__global__ void k_fill_int(int* p, int v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_finish_T(double* T, int n, int zero_all)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = T[i];
    if (v < 1.0e-60) v = 0.0;     // clamp of the reference (negf/lesser.f90:795-800)
    if (zero_all) v = 0.0;
    T[i] = v;
}

}  // namespace

template <typename T>
hipError_t launch_copy_op(int m, int n, BMat<T> in, BMatW<T> out, int dagger, T sign, const int* active, int batch,
                          hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_op<T>, dim3(blocks_for((long)m * n), batch), dim3(NT), 0, s, m, n, in, out, dagger,
                       sign, active, batch);
    return hipGetLastError();
}

template <typename Ti, typename To>
hipError_t launch_convert(int m, int n, BMat<Ti> in, BMatW<To> out, int dagger, To sign, int batch, hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_convert<Ti, To>), dim3(blocks_for((long)m * n), batch), dim3(NT), 0, s, m, n, in, out,
                       dagger, sign, batch);
    return hipGetLastError();
}
template hipError_t launch_convert<double, float>(int, int, BMat<double>, BMatW<float>, int, float, int, hipStream_t);
template hipError_t launch_convert<float, double>(int, int, BMat<float>, BMatW<double>, int, double, int, hipStream_t);

template <typename T>
hipError_t launch_gamma(int n, BMat<T> S, BMatW<T> G, T scale, const int* active, int batch, hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gamma<T>, dim3(blocks_for((long)n * n), batch), dim3(NT), 0, s, n, S, G, scale, active,
                       batch);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_shift_sub(int n, const T* z, BMat<T> A, BMat<T> B, BMatW<T> out, const int* active, int batch,
                            hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_shift_sub<T>, dim3(blocks_for((long)n * n), batch), dim3(NT), 0, s, n, z, A, B, out,
                       active, batch);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dec_init(int n, BMat<T> H00, BMat<T> V, int alpha_dagger, BMatW<T> e1, BMatW<T> es, BMatW<T> ab,
                           int batch, hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dec_init<T>, dim3(blocks_for((long)n * n), batch), dim3(NT), 0, s, n, H00, V, alpha_dagger,
                       e1, es, ab, batch);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_trace_prod(int n1, int n2, BMat<T> X, BMat<T> Y, double* out, int out_stride, const int* active,
                             int batch, hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_prod<T>, dim3(batch), dim3(NT), 0, s, n1, n2, X, Y, out, out_stride, active, batch);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_trace_im(int n, BMat<T> g, const double* E, double coef, double* out, int batch, hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_im<T>, dim3(batch), dim3(NT), 0, s, n, g, E, coef, out, batch);
    return hipGetLastError();
}

hipError_t launch_make_shifts(const double* E, int batch, double eta, double eta0, int eta_mode, double dev_factor,
                              double* zl, double* zr, double* zd, float* zl32, float* zr32, float* zd32,
                              hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_make_shifts, dim3(blocks_for(batch)), dim3(NT), 0, s, E, batch, eta, eta0, eta_mode,
                       dev_factor, zl, zr, zd, zl32, zr32, zd32);
    return hipGetLastError();
}

hipError_t launch_dec_step_update(int* active, unsigned long long* maxabs, int* iters, double conv, int* count,
                                  int* clear, int batch, hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dec_step_update, dim3(blocks_for(batch)), dim3(NT), 0, s, active, maxabs, iters, conv,
                       count, clear, batch);
    return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int v, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(n)), dim3(NT), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_finish_T(double* T, int n, int zero_all, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_finish_T, dim3(blocks_for(n)), dim3(NT), 0, s, T, n, zero_all);
    return hipGetLastError();
}

#define NT_INST(T)                                                                                             \
    template hipError_t launch_copy_op<T>(int, int, BMat<T>, BMatW<T>, int, T, const int*, int, hipStream_t); \
    template hipError_t launch_gamma<T>(int, BMat<T>, BMatW<T>, T, const int*, int, hipStream_t);             \
    template hipError_t launch_shift_sub<T>(int, const T*, BMat<T>, BMat<T>, BMatW<T>, const int*, int,        \
                                            hipStream_t);                                                      \
    template hipError_t launch_dec_init<T>(int, BMat<T>, BMat<T>, int, BMatW<T>, BMatW<T>, BMatW<T>, int,      \
                                           hipStream_t);                                                       \
    template hipError_t launch_trace_prod<T>(int, int, BMat<T>, BMat<T>, double*, int, const int*, int,        \
                                             hipStream_t);                                                     \
    template hipError_t launch_trace_im<T>(int, BMat<T>, const double*, double, double*, int, hipStream_t);
NT_INST(double)
NT_INST(float)
#undef NT_INST

}  // namespace nt
