// nt_unfold.hip -- Brillouin-zone unfolding of lead modes: the labels that
// split mode-resolved transmission into specular and non-specular parts.
//
// vel_mat_tdb (negf/velocity_mat.f90:2035-2076 and its three mirrors) gives
// every propagating mode i of a lead the weight, for each of the
// nb = n_bloch_x * n_bloch_y in-plane k-shifts j,
//     W(i, j) = sum_o | sum_ir exp(i(kc.r - kp_j.(r - r_uc(o)))) U(o + ir nuc, i) lambda_i |^2 / nb
// (o over the nuc orbitals of one transverse cell, ir over the nb cell images,
// r the orbital's position, kc = ksc . recivec, kp_j = move_in_ws2((ksc +
// shift_j) . recivec) folded into the unit cell's zone).  The mode's label is
// the first j of largest weight; its unfolded k is kp of that shift, with
// k_z in the third component.  The label pairs then split sum |t|^2 into
// equal-label (specular) and different-label (non-specular) parts
// (:3032-3099); write_specular / write_md (negf/write_data.f90:3-175) sum
// each incident mode's row over outgoing modes of matching in-plane k.
//
// All of it is O(n^2 nb) per point against the O(n^3) of the modes: one
// thread per mode, one workgroup per point for the sums.
#include "nt_kernels.h"

namespace nt {

namespace {

constexpr int NT_U = 64;

// k - n1 a1 - n2 a2 for the (n1, n2) in -2..2 of smallest |.|^2 (first in loop
// order: n1 outer, n2 inner) -- move_in_ws2, negf/util.f90:675-706
__device__ void ws2(const double* k, const double* a, double* out)
{
    double best = 0.0;
    int b1 = 0, b2 = 0;
    bool first = true;
    for (int i1 = -2; i1 <= 2; ++i1)
        for (int i2 = -2; i2 <= 2; ++i2) {
            double d = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double v = k[c] - (i1 * a[c] + i2 * a[3 + c]);
                d += v * v;
            }
            if (first || d < best) {
                best = d;
                b1 = i1;
                b2 = i2;
                first = false;
            }
        }
    for (int c = 0; c < 3; ++c) out[c] = k[c] - (b1 * a[c] + b2 * a[3 + c]);
}

// (ksc + shift_j) . recivec, shift_j = (j / nby, j % nby, 0) (config.f90:176-183)
__device__ void shifted_k(const double* ksc, int j, int nby, const double* rv, double* out)
{
    const double s0 = ksc[0] + (double)(j / nby), s1 = ksc[1] + (double)(j % nby), s2 = ksc[2];
    for (int c = 0; c < 3; ++c) out[c] = s0 * rv[c] + s1 * rv[3 + c] + s2 * rv[6 + c];
}

__global__ void __launch_bounds__(NT_U) k_unfold_weights(UnfoldDirArgs a)
{
    extern __shared__ double kp[];            // nb x 3 folded shifted k of this point
    const int p = blockIdx.x;
    if (p >= a.batch) return;
    const int n = a.n, nb = a.nbx * a.nby;
    const double* ksc = a.ksc + 3 * (a.kidx ? a.kidx[p] : 0);
    for (int j = threadIdx.x; j < nb; j += blockDim.x) {
        double k[3];
        shifted_k(ksc, j, a.nby, a.recivec, k);
        ws2(k, a.reci_uc, kp + 3 * j);
    }
    __syncthreads();
    const int i = blockIdx.y * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double kz = a.kz[(long)p * n + i];
    int* bz = a.bz + (long)p * n;
    double* kx = a.kxyz + (long)p * 3 * n;
    if (!(kz < 100.0)) {
        bz[i] = 0;
        kx[i] = kx[n + i] = kx[2 * n + i] = 100.0;
        return;
    }
    double kc[3];
    shifted_k(ksc, 0, a.nby, a.recivec, kc);      // shift 0 is (0, 0, 0): kc = ksc . recivec
    const double* U = a.U + 2 * ((long)p * a.ustride + (long)i * a.ldu);
    const double lr = a.w[2 * ((long)p * n + i)], li = a.w[2 * ((long)p * n + i) + 1];
    const double l2 = lr * lr + li * li;
    const int nuc = a.nuc;
    double best = -1.0;
    int jb = 0;
    for (int j = 0; j < nb; ++j) {
        const double* q = kp + 3 * j;
        double wsum = 0.0;
        for (int o = 0; o < nuc; ++o) {
            const double* ru = a.pos_uc + 3 * o;
            double sr = 0.0, si = 0.0;
            for (int ir = 0; ir < nb; ++ir) {
                const int r = o + ir * nuc;
                const double* x = a.pos + 3 * r;
                const double ph = kc[0] * x[0] + kc[1] * x[1] + kc[2] * x[2] -
                                  (q[0] * (x[0] - ru[0]) + q[1] * (x[1] - ru[1]) + q[2] * (x[2] - ru[2]));
                double s, c;
                sincos(ph, &s, &c);
                const double ur = U[2 * r], ui = U[2 * r + 1];
                sr += c * ur - s * ui;
                si += c * ui + s * ur;
            }
            wsum += (sr * sr + si * si) * l2 / (double)nb;
        }
        if (wsum > best) {                          // maxloc: the first maximum
            best = wsum;
            jb = j;
        }
    }
    bz[i] = jb + 1;
    double k[3], f[3];
    shifted_k(ksc, jb, a.nby, a.recivec, k);
    ws2(k, a.reci_uc, f);
    kx[i] = f[0];
    kx[n + i] = f[1];
    kx[2 * n + i] = kz;
}

__device__ double block_sum(double v, double* sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    return t;
}

// one workgroup per point: the eight split sums
__global__ void __launch_bounds__(256) k_unfold_split(UnfoldSplitArgs a)
{
    __shared__ double sh[4];
    const int p = blockIdx.x;
    if (p >= a.batch) return;
    for (int q = 0; q < 4; ++q) {
        const UnfoldChain& C = a.ch[q];
        const double* M = C.M + 2 * (long)p * C.mo * C.mi;
        const int* br = C.bz_row + (long)p * C.mo;
        const int* bc = C.bz_col + (long)p * C.mi;
        double s = 0.0, ns = 0.0;
        for (long e = threadIdx.x; e < (long)C.mo * C.mi; e += blockDim.x) {
            const int r = (int)(e % C.mo), c = (int)(e / C.mo);
            const int lr = br[r], lc = bc[c];
            if (lr <= 0 || lc <= 0) continue;
            const double v = M[2 * e] * M[2 * e] + M[2 * e + 1] * M[2 * e + 1];
            if (lr == lc)
                s += v;
            else
                ns += v;
        }
        s = block_sum(s, sh);
        ns = block_sum(ns, sh);
        if (threadIdx.x == 0) {
            a.split[8L * p + 2 * C.slot] = s;
            a.split[8L * p + 2 * C.slot + 1] = ns;
        }
    }
}

// one thread per incident mode: (total, specular) over outgoing propagating modes
__global__ void __launch_bounds__(NT_U) k_unfold_rows(UnfoldSplitArgs a)
{
    const int p = blockIdx.x, q = blockIdx.z;
    if (p >= a.batch || q >= 4) return;
    const UnfoldChain& C = a.ch[q];
    const int i = blockIdx.y * blockDim.x + threadIdx.x;
    if (!C.rows || i >= C.mi) return;
    const double* kb = C.k_col + (long)p * 3 * C.mi;
    const double* kf = C.k_row + (long)p * 3 * C.mo;
    double* out = C.rows + (long)p * 2 * C.mi;
    double t = 0.0, sp = 0.0;
    if (kb[2 * C.mi + i] < 99.9) {
        const double* M = C.M + 2 * ((long)p * C.mo * C.mi + (long)i * C.mo);
        const double bx = kb[i], by = kb[C.mi + i];
        for (int j = 0; j < C.mo; ++j) {
            if (!(kf[2 * C.mo + j] < 99.9)) continue;
            const double v = M[2 * j] * M[2 * j] + M[2 * j + 1] * M[2 * j + 1];
            t += v;
            const double dx = kf[j] - bx, dy = kf[C.mo + j] - by;
            if (dx * dx + dy * dy < 1.0e-5) sp += v;
        }
    }
    out[i] = t;
    out[C.mi + i] = sp;
}

}  // namespace

hipError_t launch_unfold_weights(const UnfoldDirArgs& a, hipStream_t s)
{
    if (a.batch <= 0 || a.n <= 0) return hipSuccess;
    const int nb = a.nbx * a.nby;
    hipLaunchKernelGGL(k_unfold_weights, dim3(a.batch, (a.n + NT_U - 1) / NT_U), dim3(NT_U),
                       sizeof(double) * 3 * nb, s, a);
    return hipGetLastError();
}

hipError_t launch_unfold_split(const UnfoldSplitArgs& a, hipStream_t s)
{
    if (a.batch <= 0) return hipSuccess;
    if (a.split) hipLaunchKernelGGL(k_unfold_split, dim3(a.batch), dim3(256), 0, s, a);
    int mi = 0;
    bool rows = false;
    for (const UnfoldChain& C : a.ch) {
        mi = std::max(mi, C.mi);
        rows = rows || C.rows;
    }
    if (rows) hipLaunchKernelGGL(k_unfold_rows, dim3(a.batch, (mi + NT_U - 1) / NT_U, 4), dim3(NT_U), 0, s, a);
    return hipGetLastError();
}

}  // namespace nt
