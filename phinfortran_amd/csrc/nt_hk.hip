// nt_hk.hip -- H(k) block assembly of the QE force-constant (tdb) path.
//
// gen_dyn_qe_tdb_k (negf/dyn.f90:1482-2080) builds every block as a sum over
// atom pairs (i, j) and in-plane supercell images of
//     w * Phi_ij(R) / sqrt(m_i m_j) * exp(i k . dr)
// (3x3 per pair).  The host tabulates the k-independent part once, as
// segments of entries (one segment per 3x3 sub-block, entries in the
// reference's summation order); this kernel applies the phases for all k at
// once, one thread per (k, sub-block element).
//
// gen_ham (tb/hamiltonian.f90:444-519) is the same sum with sub-blocks of
// norb_i x norb_j orbitals per atom pair (the 27 cell images within the
// cutoff as entries) followed by hermitian = (H + H^+) / 2
// (tb/util.f90:709-718): k_hk_phase_orb and k_hk_hermitize.
#include "nt_kernels.h"

namespace nt {

namespace {

__global__ void k_hk_zero(double* out, int rows, int cols, int ld, long kstride, int nk)
{
    const int k = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk || e >= (long)rows * cols) return;
    const int r = (int)(e % rows), c = (int)(e / rows);
    double* o = out + 2 * ((long)k * kstride + r + (long)c * ld);
    o[0] = 0.0;
    o[1] = 0.0;
}

__global__ void k_hk_phase(int nk, const double* kc, int nseg, const int* seg_off, const int* seg_rc,
                           const double* coef, const double* dr, double* out, int ld, long kstride)
{
    const int k = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk || e >= 9L * nseg) return;
    const int seg = (int)(e / 9), ab = (int)(e % 9), al = ab / 3, be = ab % 3;
    const double kx = kc[3 * k], ky = kc[3 * k + 1], kz = kc[3 * k + 2];
    double sr = 0.0, si = 0.0;
    for (int q = seg_off[seg]; q < seg_off[seg + 1]; ++q) {
        const double ph = dr[3 * q] * kx + dr[3 * q + 1] * ky + dr[3 * q + 2] * kz;
        double s, c;
        sincos(ph, &s, &c);
        const double f = coef[9L * q + ab];
        sr += f * c;
        si += f * s;
    }
    const int r = 3 * seg_rc[2 * seg] + al, col = 3 * seg_rc[2 * seg + 1] + be;
    double* o = out + 2 * ((long)k * kstride + r + (long)col * ld);
    o[0] = sr;
    o[1] = si;
}

// segment s: sub-block at orbital row r0, column c0 of ni x nj elements,
// seg_geom[4 s .. 4 s + 3] = (r0, c0, ni, nj); entry q holds its ni x nj
// coefficients row-major at coef + q * cstride
__global__ void k_hk_phase_orb(int nk, const double* kc, int nseg, const int* seg_off, const int* seg_geom,
                               int cstride, const double* coef, const double* dr, double* out, int ld, long kstride)
{
    const int k = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk || e >= (long)cstride * nseg) return;
    const int seg = (int)(e / cstride), ab = (int)(e % cstride);
    const int r0 = seg_geom[4 * seg], c0 = seg_geom[4 * seg + 1], ni = seg_geom[4 * seg + 2],
              nj = seg_geom[4 * seg + 3];
    if (ab >= ni * nj) return;
    const double kx = kc[3 * k], ky = kc[3 * k + 1], kz = kc[3 * k + 2];
    double sr = 0.0, si = 0.0;
    for (int q = seg_off[seg]; q < seg_off[seg + 1]; ++q) {
        const double ph = dr[3 * q] * kx + dr[3 * q + 1] * ky + dr[3 * q + 2] * kz;
        double s, c;
        sincos(ph, &s, &c);
        const double f = coef[(long)cstride * q + ab];
        sr += f * c;
        si += f * s;
    }
    double* o = out + 2 * ((long)k * kstride + r0 + ab / nj + (long)(c0 + ab % nj) * ld);
    o[0] = sr;
    o[1] = si;
}

// H <- (H + H^+) / 2 in place, n x n per k; thread (r, c), r <= c, owns both
// (r, c) and (c, r)
__global__ void k_hk_hermitize(double* out, int n, int ld, long kstride, int nk)
{
    const int k = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk || e >= (long)n * n) return;
    const int r = (int)(e % n), c = (int)(e / n);
    if (r > c) return;
    double* x = out + 2 * ((long)k * kstride + r + (long)c * ld);
    double* y = out + 2 * ((long)k * kstride + c + (long)r * ld);
    const double xr = x[0], xi = x[1], yr = y[0], yi = y[1];
    x[0] = 0.5 * (xr + yr);
    x[1] = 0.5 * (xi - yi);
    y[0] = 0.5 * (yr + xr);
    y[1] = 0.5 * (yi - xi);
}

}  // namespace

hipError_t launch_hk_assemble_orb(int nk, const double* kc, int nseg, const int* seg_off, const int* seg_geom,
                                  int cstride, const double* coef, const double* dr, int rows, int cols,
                                  double* out, int ld, long kstride, int hermitize, hipStream_t s)
{
    if (nk <= 0) return hipSuccess;
    const long nz = (long)rows * cols;
    hipLaunchKernelGGL(k_hk_zero, dim3((nz + 255) / 256, nk), dim3(256), 0, s, out, rows, cols, ld, kstride, nk);
    if (nseg > 0)
        hipLaunchKernelGGL(k_hk_phase_orb, dim3(((long)cstride * nseg + 255) / 256, nk), dim3(256), 0, s, nk, kc,
                           nseg, seg_off, seg_geom, cstride, coef, dr, out, ld, kstride);
    if (hermitize)
        hipLaunchKernelGGL(k_hk_hermitize, dim3((nz + 255) / 256, nk), dim3(256), 0, s, out, rows, ld, kstride, nk);
    return hipGetLastError();
}

hipError_t launch_hk_assemble(int nk, const double* kc, int nseg, const int* seg_off, const int* seg_rc,
                              const double* coef, const double* dr, int rows, int cols, double* out, int ld,
                              long kstride, hipStream_t s)
{
    if (nk <= 0) return hipSuccess;
    const long nz = (long)rows * cols;
    hipLaunchKernelGGL(k_hk_zero, dim3((nz + 255) / 256, nk), dim3(256), 0, s, out, rows, cols, ld, kstride, nk);
    if (nseg > 0)
        hipLaunchKernelGGL(k_hk_phase, dim3((9L * nseg + 255) / 256, nk), dim3(256), 0, s, nk, kc, nseg, seg_off,
                           seg_rc, coef, dr, out, ld, kstride);
    return hipGetLastError();
}

}  // namespace nt
