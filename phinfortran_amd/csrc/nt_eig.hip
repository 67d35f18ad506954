// nt_eig.hip -- batched complex Schur decomposition, eigenvectors and
// principal matrix square root (FP64), one workgroup of 256 threads per
// matrix, working matrices in global memory (L2-resident at these sizes).
//
// These replace the LAPACK calls of the mode-resolved transmission:
//   eigen  (negf/util.f90:226-262, ZGEEV)  -> k_schur + k_trevc + gemm + k_ev_normalize
//   sqrtm  (negf/util.f90:359-454, ZGEES + triangular square root)
//                                           -> k_schur + k_sqrt_tri + 2 gemm
//
// k_schur: Householder reduction to Hessenberg form (Z accumulates the
// reflectors), then single-shift complex QR sweeps with Wilkinson shifts and
// exceptional shifts every 10th sweep, deflating on |h(k,k-1)| <= ulp
// (|h(k-1,k-1)| + |h(k,k)|) (|re|+|im| norms).  The whole Schur form is kept
// (rotations applied to full rows and to Z), so A = Z T Z^H on exit.  Every
// loop is bounded: more than 30 max(10, n) sweeps for one eigenvalue, or
// 60 n in total, ends the kernel with status -3.
#include "nt_kernels.h"

namespace nt {

namespace {

constexpr int NT = 256;
constexpr double ULP = 2.220446049250313e-16;
constexpr double SAFMIN = 2.2250738585072014e-308;

struct C2 {
    double r, i;
};
__device__ __forceinline__ C2 cmk(double r, double i) { return C2{r, i}; }
__device__ __forceinline__ C2 cadd(C2 a, C2 b) { return C2{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ C2 csub(C2 a, C2 b) { return C2{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ C2 cmul(C2 a, C2 b) { return C2{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ C2 cconj(C2 a) { return C2{a.r, -a.i}; }
__device__ __forceinline__ C2 cscale(C2 a, double s) { return C2{a.r * s, a.i * s}; }
__device__ __forceinline__ double cabs1(C2 a) { return fabs(a.r) + fabs(a.i); }
__device__ __forceinline__ double cabs(C2 a) { return hypot(a.r, a.i); }
__device__ __forceinline__ C2 cdiv(C2 a, C2 b)
{
    // Smith's algorithm
    if (fabs(b.r) >= fabs(b.i)) {
        const double r = b.i / b.r, d = b.r + b.i * r;
        return C2{(a.r + a.i * r) / d, (a.i - a.r * r) / d};
    }
    const double r = b.r / b.i, d = b.i + b.r * r;
    return C2{(a.r * r + a.i) / d, (a.i * r - a.r) / d};
}
// principal square root (branch cut on the negative real axis, Re >= 0)
__device__ __forceinline__ C2 csqrt_p(C2 z)
{
    const double x = z.r, y = z.i;
    if (x == 0.0 && y == 0.0) return C2{0.0, 0.0};
    const double t = sqrt(0.5 * (hypot(x, y) + fabs(x)));
    if (x >= 0.0) return C2{t, y / (2.0 * t)};
    return C2{fabs(y) / (2.0 * t), copysign(t, y)};
}

__device__ __forceinline__ C2 ld(const double* M, int ldm, int r, int c)
{
    const double2 v = *reinterpret_cast<const double2*>(M + 2 * ((long)r + (long)c * ldm));
    return C2{v.x, v.y};
}
__device__ __forceinline__ void st(double* M, int ldm, int r, int c, C2 v)
{
    *reinterpret_cast<double2*>(M + 2 * ((long)r + (long)c * ldm)) = make_double2(v.r, v.i);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// block-wide sum of one double (all threads get the result)
__device__ double block_sum(double v, double* red)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// reduction scratch of k_modes_select: after prop[n], nrm2[n], coef[n]
__device__ __forceinline__ double* coef_red(unsigned char* base, int n)
{
    int* prop = reinterpret_cast<int*>(base);
    double* nrm2 = reinterpret_cast<double*>(prop + ((n + 1) & ~1));
    C2* coef = reinterpret_cast<C2*>(nrm2 + n);
    return reinterpret_cast<double*>(coef + n);
}

struct SchurShared {
    double red[4];
    int ired[4];
    C2 tau, beta, scale, mu, s;
    double c;
    int l, m, its, total, status, done;
};

}  // namespace

// ----------------------------------------------------------------- Schur
__global__ __launch_bounds__(NT) void k_schur(EigArgs a)
{
    const int p = blockIdx.x;
    if (p >= a.batch) return;
    const int n = a.n;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    __shared__ SchurShared S;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2* v = reinterpret_cast<C2*>(smem_raw);     // [n] Householder vector

    double* A = bptr(a.A, p, 0);
    double* Z = bptr(a.Z, p, 0);
    const int lda = a.A.ld, ldz = a.Z.ld;

    // Z = I
    for (long e = t; e < (long)n * n; e += NT) {
        const int r = (int)(e % n), c = (int)(e / n);
        st(Z, ldz, r, c, C2{r == c ? 1.0 : 0.0, 0.0});
    }
    if (t == 0) S.status = 0;
    __syncthreads();

    // ---- Hessenberg reduction
    for (int k = 0; k + 2 < n; ++k) {
        double xs = 0.0;
        for (int i = k + 2 + t; i < n; i += NT) {
            const C2 x = ld(A, lda, i, k);
            xs += x.r * x.r + x.i * x.i;
        }
        xs = block_sum(xs, S.red);
        if (t == 0) {
            const C2 alpha = ld(A, lda, k + 1, k);
            if (xs == 0.0 && alpha.i == 0.0) {
                S.tau = C2{0.0, 0.0};
            } else {
                const double nrm = sqrt(alpha.r * alpha.r + alpha.i * alpha.i + xs);
                const double beta = alpha.r >= 0.0 ? -nrm : nrm;
                S.tau = C2{(beta - alpha.r) / beta, -alpha.i / beta};
                S.beta = C2{beta, 0.0};
                S.scale = cdiv(C2{1.0, 0.0}, csub(alpha, C2{beta, 0.0}));
            }
        }
        __syncthreads();
        const C2 tau = S.tau;
        if (tau.r == 0.0 && tau.i == 0.0) continue;     // uniform: every thread reads the same value
        const int m = n - k - 1;                          // length of v
        const C2 scale = S.scale;
        for (int i = t; i < m; i += NT) v[i] = i == 0 ? C2{1.0, 0.0} : cmul(ld(A, lda, k + 1 + i, k), scale);
        __syncthreads();
        if (t == 0) st(A, lda, k + 1, k, S.beta);
        for (int i = k + 2 + t; i < n; i += NT) st(A, lda, i, k, C2{0.0, 0.0});
        // left: A[k+1:, j] -= conj(tau) v (v^H A[k+1:, j]), j > k; one wave per column
        const C2 ctau = cconj(tau);
        for (int j = k + 1 + wave; j < n; j += NT / 64) {
            double wr = 0.0, wi = 0.0;
            for (int i = lane; i < m; i += 64) {
                const C2 x = ld(A, lda, k + 1 + i, j), vi = v[i];
                wr += vi.r * x.r + vi.i * x.i;     // conj(v) x
                wi += vi.r * x.i - vi.i * x.r;
            }
            wr = wave_sum(wr);
            wi = wave_sum(wi);
            const C2 f = cmul(ctau, C2{wr, wi});
            for (int i = lane; i < m; i += 64) {
                const C2 x = ld(A, lda, k + 1 + i, j);
                st(A, lda, k + 1 + i, j, csub(x, cmul(v[i], f)));
            }
        }
        __syncthreads();
        // right: M[r, k+1:] -= tau (M[r, k+1:] v) v^H for M = A and Z
        for (int r = t; r < n; r += NT) {
            C2 y{0.0, 0.0}, yz{0.0, 0.0};
            for (int i = 0; i < m; ++i) {
                y = cadd(y, cmul(ld(A, lda, r, k + 1 + i), v[i]));
                yz = cadd(yz, cmul(ld(Z, ldz, r, k + 1 + i), v[i]));
            }
            y = cmul(tau, y);
            yz = cmul(tau, yz);
            for (int i = 0; i < m; ++i) {
                const C2 cv = cconj(v[i]);
                st(A, lda, r, k + 1 + i, csub(ld(A, lda, r, k + 1 + i), cmul(y, cv)));
                st(Z, ldz, r, k + 1 + i, csub(ld(Z, ldz, r, k + 1 + i), cmul(yz, cv)));
            }
        }
        __syncthreads();
    }

    // ---- shifted QR sweeps
    if (t == 0) {
        S.m = n - 1;
        S.its = 0;
        S.total = 0;
        S.done = 0;
    }
    __syncthreads();
    const int itmax = 30 * (n > 10 ? n : 10);
    const int totmax = 60 * n + 60;
    while (true) {
        const int m = S.m;
        if (m < 0) break;
        // deflation: largest k in (0, m] with a negligible subdiagonal
        int cand = 0;
        for (int k = 1 + t; k <= m; k += NT) {
            const double sub = cabs1(ld(A, lda, k, k - 1));
            const double tst = cabs1(ld(A, lda, k - 1, k - 1)) + cabs1(ld(A, lda, k, k));
            if (sub <= 100.0 * SAFMIN || sub <= ULP * tst) cand = k > cand ? k : cand;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(cand, off);
            cand = o > cand ? o : cand;
        }
        __syncthreads();
        if (lane == 0) S.ired[wave] = cand;
        __syncthreads();
        const int l = max(max(S.ired[0], S.ired[1]), max(S.ired[2], S.ired[3]));
        if (l > 0 && t == 0) st(A, lda, l, l - 1, C2{0.0, 0.0});
        if (l == m) {
            __syncthreads();
            if (t == 0) {
                S.m = m - 1;
                S.its = 0;
            }
            __syncthreads();
            continue;
        }
        // one sweep on the active window [l, m]
        if (t == 0) {
            S.its += 1;
            S.total += 1;
            if (S.its > itmax || S.total > totmax) S.status = -3;
            C2 mu;
            if (S.its % 10 == 0) {
                const C2 hm = ld(A, lda, m, m);
                mu = C2{hm.r + 0.75 * fabs(ld(A, lda, m, m - 1).r), hm.i};
            } else {
                const C2 aa = ld(A, lda, m - 1, m - 1), bb = ld(A, lda, m - 1, m), cc = ld(A, lda, m, m - 1),
                         dd = ld(A, lda, m, m);
                const C2 tr = cscale(csub(aa, dd), 0.5);
                const C2 disc = csqrt_p(cadd(cmul(tr, tr), cmul(bb, cc)));
                const C2 e1 = cadd(cadd(dd, tr), disc), e2 = csub(cadd(dd, tr), disc);
                mu = cabs(csub(e1, dd)) < cabs(csub(e2, dd)) ? e1 : e2;
            }
            S.mu = mu;
            S.l = l;
        }
        __syncthreads();
        if (S.status) break;
        for (int j = l; j < m; ++j) {
            if (t == 0) {
                C2 f, g;
                if (j == l) {
                    f = csub(ld(A, lda, l, l), S.mu);
                    g = ld(A, lda, l + 1, l);
                } else {
                    f = ld(A, lda, j, j - 1);
                    g = ld(A, lda, j + 1, j - 1);
                }
                double c;
                C2 s, r;
                if (g.r == 0.0 && g.i == 0.0) {
                    c = 1.0;
                    s = C2{0.0, 0.0};
                    r = f;
                } else if (f.r == 0.0 && f.i == 0.0) {
                    const double ag = cabs(g);
                    c = 0.0;
                    s = cscale(cconj(g), 1.0 / ag);
                    r = C2{ag, 0.0};
                } else {
                    const double af = cabs(f), ag = cabs(g), d = hypot(af, ag);
                    const C2 ph = cscale(f, 1.0 / af);
                    c = af / d;
                    s = cscale(cmul(ph, cconj(g)), 1.0 / d);
                    r = cscale(ph, d);
                }
                if (j > l) {
                    st(A, lda, j, j - 1, r);
                    st(A, lda, j + 1, j - 1, C2{0.0, 0.0});
                }
                S.c = c;
                S.s = s;
            }
            __syncthreads();
            const double c = S.c;
            const C2 s = S.s, sc = cconj(s);
            // rows j, j+1, columns j..n-1
            for (int col = j + t; col < n; col += NT) {
                const C2 x = ld(A, lda, j, col), y = ld(A, lda, j + 1, col);
                st(A, lda, j, col, cadd(cscale(x, c), cmul(s, y)));
                st(A, lda, j + 1, col, csub(cscale(y, c), cmul(sc, x)));
            }
            __syncthreads();
            // columns j, j+1: rows 0..min(j+2, m) of A, all rows of Z
            const int rmax = min(j + 2, m);
            for (int r = t; r < n; r += NT) {
                if (r <= rmax) {
                    const C2 x = ld(A, lda, r, j), y = ld(A, lda, r, j + 1);
                    st(A, lda, r, j, cadd(cscale(x, c), cmul(sc, y)));
                    st(A, lda, r, j + 1, csub(cscale(y, c), cmul(s, x)));
                }
                const C2 x = ld(Z, ldz, r, j), y = ld(Z, ldz, r, j + 1);
                st(Z, ldz, r, j, cadd(cscale(x, c), cmul(sc, y)));
                st(Z, ldz, r, j + 1, csub(cscale(y, c), cmul(s, x)));
            }
            __syncthreads();
        }
    }
    __syncthreads();
    // clean the strictly lower triangle (T is upper triangular)
    for (long e = t; e < (long)n * n; e += NT) {
        const int r = (int)(e % n), c = (int)(e / n);
        if (r > c) st(A, lda, r, c, C2{0.0, 0.0});
    }
    if (t == 0 && a.status) {
        const int v = S.status;
        if (v && a.status[p] == 0) a.status[p] = v;
    }
    if (a.w) {
        for (int i = t; i < n; i += NT) {
            const C2 d = ld(A, lda, i, i);
            a.w[2 * ((long)p * n + i)] = d.r;
            a.w[2 * ((long)p * n + i) + 1] = d.i;
        }
    }
}

// ------------------------------------------------------------ eigenvectors
// Upper triangular X with T X = X diag(T) (ZTREVC): column k solves
// (T[0:k,0:k] - T_kk) x = -T[0:k,k], x_k = 1, by back substitution; tiny
// pivots are replaced by max(ulp |T_kk|, n safmin / ulp).
__global__ __launch_bounds__(NT) void k_trevc(EigArgs a)
{
    const int p = blockIdx.y;
    const int k = blockIdx.x * NT + threadIdx.x;
    if (p >= a.batch) return;
    const int n = a.n;
    if (k >= n) return;
    const double* T = bptr(a.A, p, 0);
    double* X = bptr(a.X, p, 0);
    const int lt = a.A.ld, lx = a.X.ld;
    for (int i = k + 1; i < n; ++i) st(X, lx, i, k, C2{0.0, 0.0});
    st(X, lx, k, k, C2{1.0, 0.0});
    const C2 lam = ld(T, lt, k, k);
    const double smin = fmax(ULP * cabs1(lam), (double)n * SAFMIN / ULP);
    for (int i = k - 1; i >= 0; --i) {
        C2 acc = ld(T, lt, i, k);
        for (int j = i + 1; j < k; ++j) acc = cadd(acc, cmul(ld(T, lt, i, j), ld(X, lx, j, k)));
        C2 den = csub(ld(T, lt, i, i), lam);
        if (cabs1(den) < smin) den = C2{smin, 0.0};
        st(X, lx, i, k, cscale(cdiv(acc, den), -1.0));
    }
}

// ZGEEV's normalisation: unit 2-norm, then the component of largest modulus
// (first on ties) made real and positive.
__global__ __launch_bounds__(NT) void k_ev_normalize(EigArgs a)
{
    const int p = blockIdx.y;
    const int k = blockIdx.x;
    if (p >= a.batch || k >= a.n) return;
    const int n = a.n;
    double* V = bptr(a.V, p, 0);
    const int lv = a.V.ld;
    __shared__ double red[4];
    __shared__ double bestv[4];
    __shared__ int besti[4];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double s = 0.0, bv = -1.0;
    int bi = n;
    for (int i = t; i < n; i += NT) {
        const C2 x = ld(V, lv, i, k);
        const double q = x.r * x.r + x.i * x.i;
        s += q;
        if (q > bv) {
            bv = q;
            bi = i;
        }
    }
    s = block_sum(s, red);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
        }
    }
    if (lane == 0) {
        bestv[wave] = bv;
        besti[wave] = bi;
    }
    __syncthreads();
    double gv = bestv[0];
    int gi = besti[0];
    for (int w = 1; w < NT / 64; ++w)
        if (bestv[w] > gv || (bestv[w] == gv && besti[w] < gi)) {
            gv = bestv[w];
            gi = besti[w];
        }
    const double inv = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
    C2 ph{1.0, 0.0};
    if (gi < n && gv > 0.0) {
        const C2 x = ld(V, lv, gi, k);
        const double ax = sqrt(gv);
        ph = C2{x.r / ax, -x.i / ax};     // conj(x)/|x|
    }
    __syncthreads();
    for (int i = t; i < n; i += NT) {
        C2 x = cscale(cmul(ld(V, lv, i, k), ph), inv);
        if (i == gi) x.i = 0.0;
        st(V, lv, i, k, x);
    }
}

// ------------------------------------------------------------ square root
// Upper triangular U with U^2 = T (negf/util.f90:414-437): U_ii = sqrt(T_ii)
// (principal branch), then super-diagonal by super-diagonal
// U_ij = (T_ij - sum_{i<k<j} U_ik U_kj) / (U_ii + U_jj), 0 when the sum of
// the two roots is exactly zero.
__global__ __launch_bounds__(NT) void k_sqrt_tri(EigArgs a)
{
    const int p = blockIdx.x;
    if (p >= a.batch) return;
    const int n = a.n;
    const int t = threadIdx.x;
    const double* T = bptr(a.A, p, 0);
    double* U = bptr(a.X, p, 0);
    const int lt = a.A.ld, lu = a.X.ld;
    for (long e = t; e < (long)n * n; e += NT) {
        const int r = (int)(e % n), c = (int)(e / n);
        st(U, lu, r, c, r == c ? csqrt_p(ld(T, lt, r, r)) : C2{0.0, 0.0});
    }
    __syncthreads();
    for (int sd = 1; sd < n; ++sd) {
        for (int i = t; i + sd < n; i += NT) {
            const int j = i + sd;
            C2 acc = ld(T, lt, i, j);
            for (int k = i + 1; k < j; ++k) acc = csub(acc, cmul(ld(U, lu, i, k), ld(U, lu, k, j)));
            const C2 den = cadd(ld(U, lu, i, i), ld(U, lu, j, j));
            st(U, lu, i, j, (den.r != 0.0 || den.i != 0.0) ? cdiv(acc, den) : C2{0.0, 0.0});
        }
        __syncthreads();
    }
}

hipError_t launch_schur(const EigArgs& a, hipStream_t s)
{
    if (a.batch <= 0 || a.n <= 0) return hipSuccess;
    const size_t sm = sizeof(C2) * (size_t)a.n;
    hipLaunchKernelGGL(k_schur, dim3(a.batch), dim3(NT), sm, s, a);
    return hipGetLastError();
}

hipError_t launch_trevc(const EigArgs& a, hipStream_t s)
{
    if (a.batch <= 0 || a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trevc, dim3((a.n + NT - 1) / NT, a.batch), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ev_normalize(const EigArgs& a, hipStream_t s)
{
    if (a.batch <= 0 || a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ev_normalize, dim3(a.n, a.batch), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sqrt_tri(const EigArgs& a, hipStream_t s)
{
    if (a.batch <= 0 || a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sqrt_tri, dim3(a.batch), dim3(NT), 0, s, a);
    return hipGetLastError();
}

}  // namespace nt

namespace nt {

// ------------------------------------------------------ mode bookkeeping
// Propagating selection, k_z and Gram-Schmidt of one lead direction
// (negf/velocity_mat.f90:1855-1882 and its three mirrors), one workgroup per
// matrix.  Mode i propagates when ||lambda_i| - 1| < lambda_eta; its
// k_z = sign * log_fbz(lambda_i, a_z) (negf/util.f90:825-839), others get 100.
// The propagating columns of U are orthogonalised in index order with the
// reference's coefficient dot_product(U1_i, U2_j) / dot_product(U2_j, U2_j)
// (conjugate on the original column U1_i, negf/util.f90:800-823) and
// normalised.
__global__ __launch_bounds__(NT) void k_modes_select(ModesArgs a)
{
    const int p = blockIdx.x;
    if (p >= a.batch) return;
    const int n = a.n;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int* prop = reinterpret_cast<int*>(smem_raw);                        // [n] propagating indices
    double* nrm2 = reinterpret_cast<double*>(prop + ((n + 1) & ~1));     // [n] |U2_j|^2
    C2* coef = reinterpret_cast<C2*>(nrm2 + n);                           // [n]
    __shared__ int s_np;
    const double* w = a.w + 2 * (long)p * n;
    double* kz = a.kz + (long)p * n;
    double* U = bptr(a.U, p, 0);
    const int lu = a.U.ld;
    const double az = a.az, gz = 2.0 * 3.14159265358979323846 / az;
    if (t == 0) {
        int np = 0;
        for (int i = 0; i < n; ++i) {
            const double lr = w[2 * i], li = w[2 * i + 1];
            const double mod = hypot(lr, li);
            if (fabs(mod - 1.0) < a.lambda_eta) {
                prop[np++] = i;
                // Re(log(lambda) / (i az)) = arg(lambda) / az, folded into [-gz/2, gz/2]
                double k = fmod(atan2(li, lr) / az, gz);
                if (k < 0.0) k += gz;
                if (k > 0.5 * gz) k -= gz;
                kz[i] = a.sign * k;
            } else {
                kz[i] = 100.0;
            }
        }
        s_np = np;
        if (a.nprop) a.nprop[p] = np;
    }
    __syncthreads();
    const int np = s_np;
    for (int q = 0; q < np; ++q) {
        const int ci = prop[q];
        // coefficients against the finished columns j < q (one wave per j)
        for (int r = wave; r < q; r += NT / 64) {
            const int cj = prop[r];
            double cr = 0.0, cim = 0.0;
            for (int i = lane; i < n; i += 64) {
                const C2 x = ld(U, lu, i, ci), y = ld(U, lu, i, cj);
                cr += x.r * y.r + x.i * y.i;        // conj(x) y
                cim += x.r * y.i - x.i * y.r;
            }
            cr = wave_sum(cr);
            cim = wave_sum(cim);
            if (lane == 0) coef[r] = C2{cr / nrm2[r], cim / nrm2[r]};
        }
        __syncthreads();
        double s = 0.0;
        for (int i = t; i < n; i += NT) {
            C2 x = ld(U, lu, i, ci);
            for (int r = 0; r < q; ++r) x = csub(x, cmul(coef[r], ld(U, lu, i, prop[r])));
            st(U, lu, i, ci, x);
            s += x.r * x.r + x.i * x.i;
        }
        s = block_sum(s, coef_red(smem_raw, n));
        if (t == 0) nrm2[q] = s;
        __syncthreads();
    }
    for (int q = 0; q < np; ++q) {
        const double f = 1.0 / sqrt(nrm2[q]);
        for (int i = t; i < n; i += NT) st(U, lu, i, prop[q], cscale(ld(U, lu, i, prop[q]), f));
    }
}

// M[p] *= s[p] (complex per item)
__global__ void k_scale_items(double* M, long stride, int ld_, int m, int n, const double* s, int batch)
{
    const int p = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= batch || e >= (long)m * n) return;
    const int r = (int)(e % m), c = (int)(e / m);
    double* x = M + (long)p * stride * 2 + 2 * ((long)r + (long)c * ld_);
    const double xr = x[0], xi = x[1], sr = s[2 * p], si = s[2 * p + 1];
    x[0] = xr * sr - xi * si;
    x[1] = xr * si + xi * sr;
}

// per-mode sums of |S(j,i)|^2 over propagating rows j (kz_row < 99.9) for
// propagating columns i (kz_col < 99.9): col_sum[p][i] and total[p]
__global__ void k_mode_sums(const double* S, long stride, int ld_, int m, int n, const double* kz_row,
                            const double* kz_col, double* col_sum, double* total, int batch)
{
    const int p = blockIdx.x;
    if (p >= batch) return;
    const int t = threadIdx.x;
    __shared__ double red[4];
    double tot = 0.0;
    for (int i = t; i < n; i += NT) {
        double acc = 0.0;
        if (kz_col[(long)p * n + i] < 99.9)
            for (int j = 0; j < m; ++j)
                if (kz_row[(long)p * m + j] < 99.9) {
                    const double* x = S + (long)p * stride * 2 + 2 * ((long)j + (long)i * ld_);
                    acc += x[0] * x[0] + x[1] * x[1];
                }
        if (col_sum) col_sum[(long)p * n + i] = acc;
        tot += acc;
    }
    tot = block_sum(tot, red);
    if (t == 0 && total) total[p] = tot;
}

hipError_t launch_modes_select(const ModesArgs& a, hipStream_t s)
{
    if (a.batch <= 0 || a.n <= 0) return hipSuccess;
    const size_t sm = sizeof(int) * (size_t)((a.n + 1) & ~1) + sizeof(double) * a.n + sizeof(C2) * a.n +
                      sizeof(double) * 4;
    hipLaunchKernelGGL(k_modes_select, dim3(a.batch), dim3(NT), sm, s, a);
    return hipGetLastError();
}

hipError_t launch_scale_items(double* M, long stride, int ld_, int m, int n, const double* sc, int batch,
                              hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_items, dim3(((long)m * n + NT - 1) / NT, batch), dim3(NT), 0, s, M, stride, ld_, m, n,
                       sc, batch);
    return hipGetLastError();
}

hipError_t launch_mode_sums(const double* S, long stride, int ld_, int m, int n, const double* kz_row,
                            const double* kz_col, double* col_sum, double* total, int batch, hipStream_t s)
{
    if (batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mode_sums, dim3(batch), dim3(NT), 0, s, S, stride, ld_, m, n, kz_row, kz_col, col_sum, total,
                       batch);
    return hipGetLastError();
}

}  // namespace nt
