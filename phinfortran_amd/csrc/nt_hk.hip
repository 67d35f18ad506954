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
// (tb/util.f90:709-718): k_hk_orb, which writes every element once.
#include "nt_kernels.h"

#include <algorithm>

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
// coefficients row-major at coef + q * cstride.  The segments tile the block
// on a grid of row and column intervals; HkGrid maps an element to its
// segment: rblk / rloc give a row's interval and offset in it (-1: no
// segment covers the row), cblk / cloc the same for columns, and the
// segments of row interval b are sid[rptr[b] .. rptr[b+1]-1] with column
// intervals ckey[...] ascending.
// the segment holding element (r, c) and the element's offset in it; -1 if none
__device__ inline int hk_locate(const HkGrid& g, const int* seg_geom, int r, int c, long& ab)
{
    const int rb = g.rblk[r], cb = g.cblk[c];
    if (rb < 0 || cb < 0) return -1;
    int lo = g.rptr[rb], hi = g.rptr[rb + 1] - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1, key = g.ckey[mid];
        if (key == cb) {
            const int seg = g.sid[mid];
            ab = (long)g.rloc[r] * seg_geom[4 * seg + 3] + g.cloc[c];
            return seg;
        }
        if (key < cb) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

__device__ inline void hk_sum(int seg, long ab, double kx, double ky, double kz, const int* seg_off, int cstride,
                              const double* coef, const double* dr, double& sr, double& si)
{
    sr = 0.0;
    si = 0.0;
    if (seg < 0) return;
    for (int q = seg_off[seg]; q < seg_off[seg + 1]; ++q) {
        const double ph = dr[3 * q] * kx + dr[3 * q + 1] * ky + dr[3 * q + 2] * kz;
        double s, co;
        sincos(ph, &s, &co);
        const double v = coef[(long)cstride * q + ab];
        sr += v * co;
        si += v * s;
    }
}

// one thread per element (r, c), column-major order, for the k-points
// blockIdx.y, blockIdx.y + gridDim.y, ...: the element's segment (and, with
// hermitize, its mirror's) is found once; every element of every k is
// written once, (H(r, c) + conj H(c, r)) / 2 with hermitize
// A16: out is 16-byte aligned and every element is one 16-byte streaming
// (non-temporal) store; the output is written once and is larger than the caches
template <bool A16>
__global__ void k_hk_orb(int nk, const double* kc, const int* seg_off, const int* seg_geom, int cstride,
                         const double* coef, const double* dr, HkGrid g, double* out, int rows, int cols, int ld,
                         long kstride, int hermitize)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)rows * cols) return;
    const int r = (int)(e % rows), c = (int)(e / rows);
    long abx = 0, aby = 0;
    const int sx = hk_locate(g, seg_geom, r, c, abx);
    const int sy = hermitize ? hk_locate(g, seg_geom, c, r, aby) : -1;
    for (int k = blockIdx.y; k < nk; k += gridDim.y) {
        const double kx = kc[3 * k], ky = kc[3 * k + 1], kz = kc[3 * k + 2];
        double xr, xi;
        hk_sum(sx, abx, kx, ky, kz, seg_off, cstride, coef, dr, xr, xi);
        if (hermitize) {
            double yr, yi;
            hk_sum(sy, aby, kx, ky, kz, seg_off, cstride, coef, dr, yr, yi);
            xr = 0.5 * (xr + yr);
            xi = 0.5 * (xi - yi);
        }
        double* o = out + 2 * ((long)k * kstride + r + (long)c * ld);
        if constexpr (A16) {
            typedef double d2 __attribute__((ext_vector_type(2)));
            d2 v;
            v.x = xr;
            v.y = xi;
            __builtin_nontemporal_store(v, reinterpret_cast<d2*>(o));
        } else {
            o[0] = xr;
            o[1] = xi;
        }
    }
}

}  // namespace

hipError_t launch_hk_assemble_orb(int nk, const double* kc, const int* seg_off, const int* seg_geom, int cstride,
                                  const double* coef, const double* dr, const HkGrid& g, int rows, int cols,
                                  double* out, int ld, long kstride, int hermitize, hipStream_t s)
{
    if (nk <= 0) return hipSuccess;
    const long nz = (long)rows * cols;
    // k split over grid.y only as far as needed to fill the chip (about
    // 256 CUs x 2048 threads); large blocks loop over every k in one thread
    const long want = 256L * 2048;
    const int gy = (int)std::min<long>(nk, std::max<long>(1, (want + nz - 1) / nz));
    if (((uintptr_t)out & 15) == 0)
        hipLaunchKernelGGL(k_hk_orb<true>, dim3((unsigned)((nz + 255) / 256), gy), dim3(256), 0, s, nk, kc, seg_off,
                           seg_geom, cstride, coef, dr, g, out, rows, cols, ld, kstride, hermitize);
    else
        hipLaunchKernelGGL(k_hk_orb<false>, dim3((unsigned)((nz + 255) / 256), gy), dim3(256), 0, s, nk, kc, seg_off,
                           seg_geom, cstride, coef, dr, g, out, rows, cols, ld, kstride, hermitize);
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
