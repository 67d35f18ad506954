// nt_degen.hip -- the degenerate-subspace rotation of the mode-resolved solve
// (path_mode = 1; negf/velocity_mat.f90:1884-1978 and its three mirrors).
//
// For each group of degenerate eigenvalues of one mode set (found on the host
// with find_degenerate's rules), with lam the group's first eigenvalue:
//   H   = H00 + V / lam + V^+ lam, read from its upper triangle (ZHEEV "U")
//   B   = eigenvectors of H whose eigenvalue lies within 1e-4 of E (ascending),
//         divided by lam, one per group member; missing columns from the
//         eigenvectors next closest to E (the reference leaves them unset)
//   X   = argmin |A X - B|, A = the group's columns of U (normal equations:
//         A holds Gram-Schmidt-orthonormal propagating modes)
//   U(:, group) <- A X
// then every propagating column of U is normalised.  The eigenvectors of H
// come from the batched complex Schur kernel (Schur vectors of a Hermitian
// matrix are its orthonormal eigenvectors).
#include "nt_kernels.h"

namespace nt {

namespace {

constexpr int NT = 256;

struct C2 {
    double r, i;
};
__device__ inline C2 cmul(C2 a, C2 b) { return C2{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ inline C2 cdivz(C2 a, C2 b)
{
    const double d = b.r * b.r + b.i * b.i;
    return C2{(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}
__device__ inline C2 ldc(const double* M, long ld, int r, int c)
{
    const double* q = M + 2 * ((long)r + (long)c * ld);
    return C2{q[0], q[1]};
}

// H[g](i, j) from the upper triangle of H00 + V / lam + V^+ lam; grid (., G)
__global__ void k_deg_build(DegArgs a)
{
    const int g = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int m = a.m;
    if (g >= a.G || e >= (long)m * m) return;
    const DegGroup& G = a.g[g];
    const int i = (int)(e % m), j = (int)(e / m);
    const int r = i <= j ? i : j, c = i <= j ? j : i;
    const double* H0 = bptr(a.H00, G.p, 0);
    const double* V = bptr(a.V, G.p, 0);
    const C2 lam{G.lr, G.li};
    const C2 h0 = ldc(H0, a.H00.ld, r, c);
    const C2 v = ldc(V, a.V.ld, r, c);
    const C2 vt = ldc(V, a.V.ld, c, r);           // (V^+)(r, c) = conj(V(c, r))
    const C2 q = cdivz(v, lam);
    const C2 w = cmul(C2{vt.r, -vt.i}, lam);
    C2 h{h0.r + q.r + w.r, h0.i + q.i + w.i};
    if (i > j) h.i = -h.i;                          // Hermitian completion of the upper triangle
    if (i == j) h.i = 0.0;                          // the diagonal of a Hermitian matrix is real
    double* o = a.H + (long)g * a.hstride * 2 + 2 * e;
    o[0] = h.r;
    o[1] = h.i;
}

// one workgroup per group: select, least squares, rotate
__global__ __launch_bounds__(NT) void k_deg_solve(DegArgs a)
{
    const int g = blockIdx.x;
    if (g >= a.G) return;
    const DegGroup& G = a.g[g];
    const int m = a.m, nd = G.ndeg, t = threadIdx.x;
    __shared__ int sel[DEG_MAX];
    __shared__ double selv[DEG_MAX];
    __shared__ int nsel;
    __shared__ double Mr[DEG_MAX][DEG_MAX], Mi[DEG_MAX][DEG_MAX];
    __shared__ double Rr[DEG_MAX][DEG_MAX], Ri[DEG_MAX][DEG_MAX];
    __shared__ int bad;
    const double* T = a.H + (long)g * a.hstride * 2;
    const double* Z = a.Z + (long)g * a.hstride * 2;
    double* U = a.U + (long)G.p * a.ustride * 2;
    const double E = a.E[G.p];
    if (t == 0) {
        // the first nd eigenvalues within 1e-4 of E, ascending (ties: lower index)
        int k = 0;
        for (int i = 0; i < m; ++i) {
            const double ev = T[2 * ((long)i + (long)i * m)];
            if (!(fabs(E - ev) < 1.0e-4)) continue;
            int pos = k;
            while (pos > 0 && selv[pos - 1] > ev) --pos;
            if (pos >= nd) continue;
            const int last = k < nd ? k : nd - 1;
            for (int q = last; q > pos; --q) {
                selv[q] = selv[q - 1];
                sel[q] = sel[q - 1];
            }
            selv[pos] = ev;
            sel[pos] = i;
            if (k < nd) ++k;
        }
        // missing columns: the eigenvectors next closest to E (the reference
        // leaves them uninitialised); ties go to the lower index
        while (k < nd) {
            int bi = -1;
            double bd = 0.0;
            for (int i = 0; i < m; ++i) {
                bool used = false;
                for (int q = 0; q < k; ++q) used = used || sel[q] == i;
                if (used) continue;
                const double dist = fabs(E - T[2 * ((long)i + (long)i * m)]);
                if (bi < 0 || dist < bd) {
                    bi = i;
                    bd = dist;
                }
            }
            if (bi < 0) break;
            sel[k] = bi;
            selv[k] = T[2 * ((long)bi + (long)bi * m)];
            ++k;
        }
        nsel = k;
        bad = 0;
    }
    __syncthreads();
    const C2 lam{G.lr, G.li};
    // M = A^H A, R = A^H B
    if (t < nd * nd) {
        const int x = t % nd, y = t / nd;
        const int cx = G.idx[x], cy = G.idx[y];
        double mr = 0, mi = 0, rr = 0, ri = 0;
        const bool hasb = y < nsel;
        for (int r = 0; r < m; ++r) {
            const C2 ax = ldc(U, m, r, cx), ay = ldc(U, m, r, cy);
            mr += ax.r * ay.r + ax.i * ay.i;
            mi += ax.r * ay.i - ax.i * ay.r;
            if (hasb) {
                const C2 b = cdivz(ldc(Z, m, r, sel[y]), lam);
                rr += ax.r * b.r + ax.i * b.i;
                ri += ax.r * b.i - ax.i * b.r;
            }
        }
        Mr[x][y] = mr;
        Mi[x][y] = mi;
        Rr[x][y] = rr;
        Ri[x][y] = ri;
    }
    __syncthreads();
    if (t == 0) {
        // Gauss-Jordan elimination with partial pivoting (|re| + |im|): M X = R
        for (int k = 0; k < nd; ++k) {
            int pv = k;
            double best = fabs(Mr[k][k]) + fabs(Mi[k][k]);
            for (int r = k + 1; r < nd; ++r) {
                const double v = fabs(Mr[r][k]) + fabs(Mi[r][k]);
                if (v > best) {
                    best = v;
                    pv = r;
                }
            }
            if (!(best > 0.0)) {
                bad = -3;
                break;
            }
            if (pv != k)
                for (int c = 0; c < nd; ++c) {
                    double s;
                    s = Mr[k][c]; Mr[k][c] = Mr[pv][c]; Mr[pv][c] = s;
                    s = Mi[k][c]; Mi[k][c] = Mi[pv][c]; Mi[pv][c] = s;
                    s = Rr[k][c]; Rr[k][c] = Rr[pv][c]; Rr[pv][c] = s;
                    s = Ri[k][c]; Ri[k][c] = Ri[pv][c]; Ri[pv][c] = s;
                }
            // scale the pivot row, then clear column k in every other row
            const C2 inv = cdivz(C2{1.0, 0.0}, C2{Mr[k][k], Mi[k][k]});
            for (int c = 0; c < nd; ++c) {
                const C2 mv = cmul(C2{Mr[k][c], Mi[k][c]}, inv);
                Mr[k][c] = mv.r;
                Mi[k][c] = mv.i;
                const C2 rv = cmul(C2{Rr[k][c], Ri[k][c]}, inv);
                Rr[k][c] = rv.r;
                Ri[k][c] = rv.i;
            }
            for (int r = 0; r < nd; ++r) {
                if (r == k) continue;
                const C2 f{Mr[r][k], Mi[r][k]};
                for (int c = 0; c < nd; ++c) {
                    const C2 mk = cmul(f, C2{Mr[k][c], Mi[k][c]});
                    Mr[r][c] -= mk.r;
                    Mi[r][c] -= mk.i;
                    const C2 rk = cmul(f, C2{Rr[k][c], Ri[k][c]});
                    Rr[r][c] -= rk.r;
                    Ri[r][c] -= rk.i;
                }
            }
        }
        int st = a.gstatus[g];
        if (!st) st = bad;
        if (st) atomicCAS(a.status + G.p, 0, st);
    }
    __syncthreads();
    if (bad) return;
    // U(:, group) <- A X, row by row (each thread reads and writes its own rows)
    for (int r = t; r < m; r += NT) {
        C2 arow[DEG_MAX];
        for (int x = 0; x < nd; ++x) arow[x] = ldc(U, m, r, G.idx[x]);
        for (int y = 0; y < nd; ++y) {
            double vr = 0, vi = 0;
            for (int x = 0; x < nd; ++x) {
                const C2 p = cmul(arow[x], C2{Rr[x][y], Ri[x][y]});
                vr += p.r;
                vi += p.i;
            }
            double* o = U + 2 * ((long)r + (long)G.idx[y] * m);
            o[0] = vr;
            o[1] = vi;
        }
    }
}

// every propagating column (k_z < 99.9) divided by its norm; one WG per point
__global__ __launch_bounds__(NT) void k_deg_normalize(int m, double* U, long ustride, const double* kz, int P)
{
    const int p = blockIdx.x;
    if (p >= P) return;
    double* Up = U + (long)p * ustride * 2;
    __shared__ double red[NT / 64];
    __shared__ double tot;
    for (int c = 0; c < m; ++c) {
        if (!(kz[(long)p * m + c] < 99.9)) continue;
        double s = 0.0;
        for (int r = threadIdx.x; r < m; r += NT) {
            const C2 u = ldc(Up, m, r, c);
            s += u.r * u.r + u.i * u.i;
        }
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) tot = red[0] + red[1] + red[2] + red[3];
        __syncthreads();
        const double nrm = sqrt(tot);
        for (int r = threadIdx.x; r < m; r += NT) {
            double* o = Up + 2 * ((long)r + (long)c * m);
            o[0] /= nrm;
            o[1] /= nrm;
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_deg_build(const DegArgs& a, hipStream_t s)
{
    if (a.G <= 0) return hipSuccess;
    const long n = (long)a.m * a.m;
    hipLaunchKernelGGL(k_deg_build, dim3((unsigned)((n + NT - 1) / NT), a.G), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_deg_solve(const DegArgs& a, hipStream_t s)
{
    if (a.G <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_deg_solve, dim3(a.G), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_deg_normalize(int m, double* U, long ustride, const double* kz, int P, hipStream_t s)
{
    if (P <= 0 || m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_deg_normalize, dim3(P), dim3(NT), 0, s, m, U, ustride, kz, P);
    return hipGetLastError();
}

}  // namespace nt
