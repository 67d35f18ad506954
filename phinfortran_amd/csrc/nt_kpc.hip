// nt_kpc.hip -- primitive-cell unfolding of lead modes (find_k_pc).
//
// For every propagating mode of unfolded wave vector k (negf/dyn.f90:770-889):
//   * in_ws(reci_uc, k) == 2 (k on the boundary of the unit cell's zone): the
//     one candidate is k itself;
//   * else the candidates are k + n . reci_uc, n in [-l, l]^3 (l = max(n_bloch_x,
//     n_bloch_y)), that lie in or on the primitive cell's zone
//     (in_ws(recivec_pc, .) != 0), in loop order.
// Each candidate gets the primitive cell's dynamical matrix D(k) and D(k +- 1e-5
// e_x,y,z) (gen_dyn_pc_qe_k, the H(k) phase sums of nt_hk.hip over a host
// table), hermitian(D) = (D + D^H)/2, its eigenpairs (ascending, as ZHEEV
// returns them; cyclic complex Jacobi here) and the group velocities
// Re(z_j^H dD/dk z_j) / (2 sqrt|e_j|) (group_velocity_pc_qe, dyn.f90:599-659).
// A candidate matches when a band lies within 1e-2 of sqrt(E) (compare_evs,
// dyn.f90:662-768; the velocity of the last such band).  The mode's kp is the
// matching candidate of smallest |k| (first on ties, find_min_dist), its v the
// first matching candidate's velocity (velocity_mat.f90:2088-2089); zero when
// nothing matches.
//
// Tight binding (find_k_pc_l / _r, tb/hamiltonian.f90:666-916): the same
// candidates; per candidate the primitive cell's H(k) (gen_ham_pc_l,
// :918-989) and HF_d(k) = sum V r_d exp(i k.r) (gen_hf_pc_l, :1238-1313)
// from an element-wise table, hermitian(H) diagonalised, velocities
// Re(z^H [i antihermitian(HF_d)] z) (group_velocity_pc_l, :1316-1378); a band
// matches within 1e-4 of E (compare_evs_l, :1596-1686), within 1e-3 for the
// unit cell's own k on its zone boundary.
#include "nt_kernels.h"

namespace nt {

namespace {

constexpr int NT_K = 64;

// negf/util.f90:589-638
__device__ int in_ws(const double* a, const double* k)
{
    double m1 = 1e300, m2 = 1e300, d0 = 0.0;
    int idx = 0;
    for (int i1 = -2; i1 <= 2; ++i1)
        for (int i2 = -2; i2 <= 2; ++i2)
            for (int i3 = -2; i3 <= 2; ++i3) {
                double d = 0.0;
                for (int c = 0; c < 3; ++c) {
                    const double v = k[c] - (i1 * a[c] + i2 * a[3 + c] + i3 * a[6 + c]);
                    d += v * v;
                }
                if (idx == 62) d0 = d;
                if (d < m1) {
                    m2 = m1;
                    m1 = d;
                } else if (d < m2) {
                    m2 = d;
                }
                ++idx;
            }
    if (fabs(m2 - m1) < 1.0e-5) return 2;
    if (fabs(d0 - m1) < 1.0e-5) return 1;
    return 0;
}

__global__ void k_kpc_candidates(KpcArgs a, int m0, int M)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int m = m0 + i;
    const double k[3] = {a.k[3L * m], a.k[3L * m + 1], a.k[3L * m + 2]};
    double* out = a.cand + (long)i * a.maxc * 3;
    if (!(k[2] < 99.9)) {
        a.count[i] = -1;
        return;
    }
    if (in_ws(a.reci_uc, k) == 2) {
        a.count[i] = -2;           // direct: the one candidate is k
        out[0] = k[0];
        out[1] = k[1];
        out[2] = k[2];
        return;
    }
    int n = 0;
    for (int i1 = -a.l; i1 <= a.l; ++i1)
        for (int i2 = -a.l; i2 <= a.l; ++i2)
            for (int i3 = -a.l; i3 <= a.l; ++i3) {
                double q[3];
                for (int c = 0; c < 3; ++c)
                    q[c] = k[c] + (i1 * a.reci_uc[c] + i2 * a.reci_uc[3 + c] + i3 * a.reci_uc[6 + c]);
                if (in_ws(a.recivec_pc, q) != 0) {
                    out[3 * n] = q[0];
                    out[3 * n + 1] = q[1];
                    out[3 * n + 2] = q[2];
                    ++n;
                }
            }
    a.count[i] = n;
}

// compact candidate c (c0 <= c < c0 + nc) of chunk mode cm[c], slot cs[c]:
// its k into kc, the 7 k-vectors of its matrices into kl (local index)
__global__ void k_kpc_klist(KpcArgs a, const int* cm, const int* cs, int c0, int nc, double* kl, double* kc)
{
    const int cl = blockIdx.x * blockDim.x + threadIdx.x;
    if (cl >= nc) return;
    const int c = c0 + cl;
    const double* q = a.cand + ((long)cm[c] * a.maxc + cs[c]) * 3;
    kc[3L * c] = q[0];
    kc[3L * c + 1] = q[1];
    kc[3L * c + 2] = q[2];
    for (int s = 0; s < 7; ++s) {
        double* o = kl + (7L * cl + s) * 3;
        o[0] = q[0];
        o[1] = q[1];
        o[2] = q[2];
        if (s > 0) o[(s - 1) / 2] += (s % 2 ? -KPC_DEVI : KPC_DEVI);
    }
}

// Tight binding: thread per (element e = r + c n, candidate) sums the table
// entries of the element: H += V e^{i k.r}, HF_d += V r_d e^{i k.r}.
__global__ void k_kpc_tb_mats(KpcArgs a, const int* cm, const int* cs, int c0, int nc, const int* el_off,
                              const double* coef, const double* dr, double* D, double* kc)
{
    const int cl = blockIdx.y;
    const int n = a.n;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cl >= nc || e >= (long)n * n) return;
    const int c = c0 + cl;
    const double* q = a.cand + ((long)cm[c] * a.maxc + cs[c]) * 3;
    if (e == 0) {
        kc[3L * c] = q[0];
        kc[3L * c + 1] = q[1];
        kc[3L * c + 2] = q[2];
    }
    double hr = 0.0, hi = 0.0, fr[3] = {0.0, 0.0, 0.0}, fi[3] = {0.0, 0.0, 0.0};
    for (int j = el_off[e]; j < el_off[e + 1]; ++j) {
        const double* d = dr + 3L * j;
        double sn, cs_;
        sincos(d[0] * q[0] + d[1] * q[1] + d[2] * q[2], &sn, &cs_);
        const double v = coef[j];
        hr += v * cs_;
        hi += v * sn;
        for (int k = 0; k < 3; ++k) {
            fr[k] += v * d[k] * cs_;
            fi[k] += v * d[k] * sn;
        }
    }
    double* D0 = D + 2L * 7 * cl * n * n;
    D0[2 * e] = hr;
    D0[2 * e + 1] = hi;
    for (int k = 0; k < 3; ++k) {
        double* Dk = D0 + 2L * (1 + k) * n * n;
        Dk[2 * e] = fr[k];
        Dk[2 * e + 1] = fi[k];
    }
}

// hermitian(D)(r, c) of a column-major n x n complex matrix
__device__ __forceinline__ void herm(const double* D, int n, int r, int c, double& re, double& im)
{
    const double* x = D + 2 * ((long)r + (long)c * n);
    const double* y = D + 2 * ((long)c + (long)r * n);
    re = 0.5 * (x[0] + y[0]);
    im = 0.5 * (x[1] - y[1]);
}

// one workgroup per candidate: Jacobi eigenpairs of hermitian(D(k)), group
// velocities, band match
__global__ void __launch_bounds__(NT_K) k_kpc_eig(KpcArgs a, const int* cm, int m0, int c0, int nc,
                                                  const double* D, int* found, double* vout)
{
    extern __shared__ double sm[];
    const int cl = blockIdx.x;
    if (cl >= nc) return;
    const int c = c0 + cl;
    const int n = a.n, t = threadIdx.x;
    double* Ar = sm;                  // n x n, row-major [r * n + q]
    double* Ai = Ar + n * n;
    double* Vr = Ai + n * n;
    double* Vi = Vr + n * n;
    double* ev = Vi + n * n;          // n
    double* vel = ev + n;             // 3 n
    double* rot = vel + 3 * n;        // c, s, er, ei, skip, done
    int* perm = reinterpret_cast<int*>(rot + 8);
    const double* D0 = D + 2L * 7 * cl * n * n;
    for (int e = t; e < n * n; e += NT_K) {
        const int r = e / n, q = e % n;
        double re, im;
        herm(D0, n, r, q, re, im);
        Ar[e] = re;
        Ai[e] = im;
        Vr[e] = (r == q) ? 1.0 : 0.0;
        Vi[e] = 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < KPC_MAX_SWEEPS; ++sweep) {
        if (t == 0) {
            double off = 0.0, dia = 0.0;
            for (int p = 0; p < n; ++p) {
                dia += Ar[p * n + p] * Ar[p * n + p] + Ai[p * n + p] * Ai[p * n + p];
                for (int q = p + 1; q < n; ++q) off += Ar[p * n + q] * Ar[p * n + q] + Ai[p * n + q] * Ai[p * n + q];
            }
            rot[5] = (off <= 1.0e-30 * dia) ? 1.0 : 0.0;
        }
        __syncthreads();
        if (rot[5] != 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                if (t == 0) {
                    const double br = Ar[p * n + q], bi = Ai[p * n + q];
                    const double ab = sqrt(br * br + bi * bi);
                    if (ab == 0.0) {
                        rot[4] = 1.0;
                    } else {
                        const double al = Ar[p * n + p], ga = Ar[q * n + q];
                        const double th = (ga - al) / (2.0 * ab);
                        const double tt = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        const double cc = 1.0 / sqrt(tt * tt + 1.0);
                        rot[0] = cc;
                        rot[1] = tt * cc;
                        rot[2] = br / ab;
                        rot[3] = bi / ab;
                        rot[4] = 0.0;
                    }
                }
                __syncthreads();
                if (rot[4] == 0.0) {
                    const double cc = rot[0], ss = rot[1], er = rot[2], ei = rot[3];
                    // columns p, q of A and V times G = [[c, s], [-s e^-i phi, c e^-i phi]]
                    for (int r = t; r < n; r += NT_K) {
                        double xr = Ar[r * n + p], xi = Ai[r * n + p], yr = Ar[r * n + q], yi = Ai[r * n + q];
                        double gr = er * yr + ei * yi, gi = er * yi - ei * yr;     // e^-i phi y
                        Ar[r * n + p] = cc * xr - ss * gr;
                        Ai[r * n + p] = cc * xi - ss * gi;
                        Ar[r * n + q] = ss * xr + cc * gr;
                        Ai[r * n + q] = ss * xi + cc * gi;
                        xr = Vr[r * n + p];
                        xi = Vi[r * n + p];
                        yr = Vr[r * n + q];
                        yi = Vi[r * n + q];
                        gr = er * yr + ei * yi;
                        gi = er * yi - ei * yr;
                        Vr[r * n + p] = cc * xr - ss * gr;
                        Vi[r * n + p] = cc * xi - ss * gi;
                        Vr[r * n + q] = ss * xr + cc * gr;
                        Vi[r * n + q] = ss * xi + cc * gi;
                    }
                    __syncthreads();
                    // rows p, q times G^H
                    for (int r = t; r < n; r += NT_K) {
                        const double xr = Ar[p * n + r], xi = Ai[p * n + r], yr = Ar[q * n + r], yi = Ai[q * n + r];
                        const double gr = er * yr - ei * yi, gi = er * yi + ei * yr;   // e^i phi y
                        Ar[p * n + r] = cc * xr - ss * gr;
                        Ai[p * n + r] = cc * xi - ss * gi;
                        Ar[q * n + r] = ss * xr + cc * gr;
                        Ai[q * n + r] = ss * xi + cc * gi;
                    }
                    __syncthreads();
                    if (t == 0) {
                        Ar[p * n + q] = Ai[p * n + q] = Ar[q * n + p] = Ai[q * n + p] = 0.0;
                        Ai[p * n + p] = Ai[q * n + q] = 0.0;
                    }
                }
                __syncthreads();
            }
    }
    // ascending order, stable
    if (t == 0) {
        for (int j = 0; j < n; ++j) perm[j] = j;
        for (int j = 1; j < n; ++j) {
            const int pj = perm[j];
            const double v = Ar[pj * n + pj];
            int i = j - 1;
            while (i >= 0 && Ar[perm[i] * n + perm[i]] > v) {
                perm[i + 1] = perm[i];
                --i;
            }
            perm[i + 1] = pj;
        }
        for (int j = 0; j < n; ++j) ev[j] = Ar[perm[j] * n + perm[j]];
    }
    __syncthreads();
    // tight binding: Re(z^H M z), M = i antihermitian(HF_d):
    // M(r, q) = i (HF(r, q) - conj HF(q, r)) / 2
    if (a.tb) {
        for (int j = t; j < n; j += NT_K) {
            const int pj = perm[j];
            for (int d = 0; d < 3; ++d) {
                const double* F = D0 + 2L * (1 + d) * n * n;
                double acc = 0.0;
                for (int r = 0; r < n; ++r) {
                    double wr = 0.0, wi = 0.0;
                    for (int q = 0; q < n; ++q) {
                        const double* x = F + 2 * ((long)r + (long)q * n);
                        const double* y = F + 2 * ((long)q + (long)r * n);
                        const double mr = -0.5 * (x[1] + y[1]), mi = 0.5 * (x[0] - y[0]);
                        const double zr = Vr[q * n + pj], zi = Vi[q * n + pj];
                        wr += mr * zr - mi * zi;
                        wi += mr * zi + mi * zr;
                    }
                    const double zr = Vr[r * n + pj], zi = Vi[r * n + pj];
                    acc += zr * wr + zi * wi;
                }
                vel[3 * j + d] = acc;
            }
        }
        __syncthreads();
        if (t == 0) {
            const double tol = a.count[cm[c]] == -2 ? 1.0e-3 : 1.0e-4;
            const double E = a.E[m0 + cm[c]];
            int f = 0;
            double v0 = 0.0, v1 = 0.0, v2 = 0.0;
            for (int j = 0; j < n; ++j)
                if (fabs(E - ev[j]) < tol) {
                    f = 1;
                    v0 = vel[3 * j];
                    v1 = vel[3 * j + 1];
                    v2 = vel[3 * j + 2];
                }
            found[c] = f;
            vout[3L * c] = v0;
            vout[3L * c + 1] = v1;
            vout[3L * c + 2] = v2;
        }
        return;
    }
    // group velocities Re(z^H dD z) / (2 sqrt|e|), dD = (hermitian(D+) - hermitian(D-)) / (2 devi)
    for (int j = t; j < n; j += NT_K) {
        const int pj = perm[j];
        for (int d = 0; d < 3; ++d) {
            const double* Dm = D0 + 2L * (1 + 2 * d) * n * n;
            const double* Dp = D0 + 2L * (2 + 2 * d) * n * n;
            double acc = 0.0;
            for (int r = 0; r < n; ++r) {
                double wr = 0.0, wi = 0.0;       // (dD z)_r
                for (int q = 0; q < n; ++q) {
                    double pr, pi, mr, mi;
                    herm(Dp, n, r, q, pr, pi);
                    herm(Dm, n, r, q, mr, mi);
                    const double hr = (pr - mr) / (2.0 * KPC_DEVI), hi = (pi - mi) / (2.0 * KPC_DEVI);
                    const double zr = Vr[q * n + pj], zi = Vi[q * n + pj];
                    wr += hr * zr - hi * zi;
                    wi += hr * zi + hi * zr;
                }
                const double zr = Vr[r * n + pj], zi = Vi[r * n + pj];
                acc += zr * wr + zi * wi;        // Re(conj(z_r) w_r)
            }
            vel[3 * j + d] = acc / (2.0 * sqrt(fabs(ev[j])));
        }
    }
    __syncthreads();
    if (t == 0) {
        const double eq = sqrt(a.E[m0 + cm[c]]);
        int f = 0;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        for (int j = 0; j < n; ++j)
            if (fabs(eq - sqrt(fabs(ev[j]))) < 1.0e-2) {
                f = 1;
                v0 = vel[3 * j];
                v1 = vel[3 * j + 1];
                v2 = vel[3 * j + 2];
            }
        found[c] = f;
        vout[3L * c] = v0;
        vout[3L * c + 1] = v1;
        vout[3L * c + 2] = v2;
    }
}

// thread per mode: kp (smallest |k| of the matching candidates, first on
// ties) and v (the first matching candidate's), count of matches
__global__ void k_kpc_select(KpcArgs a, int m0, int M, const int* off, const int* found, const double* vc,
                             const double* kc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int m = m0 + i;
    double* kp = a.kp + 3L * m;
    double* v = a.v + 3L * m;
    const int cnt = a.count[i];
    if (cnt == -1) {
        kp[0] = kp[1] = kp[2] = 100.0;
        v[0] = v[1] = v[2] = 0.0;
        a.ncand[m] = 0;
        return;
    }
    const int o = off[i];
    if (cnt == -2) {
        kp[0] = kc[3L * o];
        kp[1] = kc[3L * o + 1];
        kp[2] = kc[3L * o + 2];
        for (int d = 0; d < 3; ++d) v[d] = found[o] ? vc[3L * o + d] : 0.0;
        a.ncand[m] = 1;
        return;
    }
    int nm = 0;
    double best = 0.0;
    kp[0] = kp[1] = kp[2] = 0.0;
    v[0] = v[1] = v[2] = 0.0;
    for (int s = 0; s < cnt; ++s) {
        const int c = o + s;
        if (!found[c]) continue;
        const double* q = kc + 3L * c;
        const double nq = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (nm == 0) {
            for (int d = 0; d < 3; ++d) v[d] = vc[3L * c + d];
        }
        if (nm == 0 || nq < best) {
            best = nq;
            kp[0] = q[0];
            kp[1] = q[1];
            kp[2] = q[2];
        }
        ++nm;
    }
    a.ncand[m] = nm;
}

}  // namespace

hipError_t launch_kpc_candidates(const KpcArgs& a, int m0, int M, hipStream_t s)
{
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kpc_candidates, dim3((M + 63) / 64), dim3(64), 0, s, a, m0, M);
    return hipGetLastError();
}

hipError_t launch_kpc_klist(const KpcArgs& a, const int* cm, const int* cs, int c0, int nc, double* kl, double* kc,
                            hipStream_t s)
{
    if (nc <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kpc_klist, dim3((nc + 63) / 64), dim3(64), 0, s, a, cm, cs, c0, nc, kl, kc);
    return hipGetLastError();
}

size_t kpc_eig_smem(int n) { return sizeof(double) * (4 * (size_t)n * n + 4 * n + 8) + sizeof(int) * n; }

hipError_t launch_kpc_eig(const KpcArgs& a, const int* cm, int m0, int c0, int nc, const double* D, int* found,
                          double* vout, hipStream_t s)
{
    if (nc <= 0) return hipSuccess;
    static const bool configured = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_kpc_eig),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kpc_eig_smem(KPC_MAX_N));
        return true;
    }();
    (void)configured;
    hipLaunchKernelGGL(k_kpc_eig, dim3(nc), dim3(NT_K), kpc_eig_smem(a.n), s, a, cm, m0, c0, nc, D, found, vout);
    return hipGetLastError();
}

hipError_t launch_kpc_tb_mats(const KpcArgs& a, const int* cm, const int* cs, int c0, int nc, const int* el_off,
                              const double* coef, const double* dr, double* D, double* kc, hipStream_t s)
{
    if (nc <= 0) return hipSuccess;
    const long ne = (long)a.n * a.n;
    hipLaunchKernelGGL(k_kpc_tb_mats, dim3((unsigned)((ne + 255) / 256), nc), dim3(256), 0, s, a, cm, cs, c0, nc,
                       el_off, coef, dr, D, kc);
    return hipGetLastError();
}

hipError_t launch_kpc_select(const KpcArgs& a, int m0, int M, const int* off, const int* found, const double* vc,
                             const double* kc, hipStream_t s)
{
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_kpc_select, dim3((M + 63) / 64), dim3(64), 0, s, a, m0, M, off, found, vc, kc);
    return hipGetLastError();
}

}  // namespace nt
