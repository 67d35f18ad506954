// nt_inv.hip -- batched complex matrix inverse (blocked Gauss-Jordan with
// partial row pivoting), one workgroup of 256 threads per matrix.
//
// Per panel of NB columns J:
//   1. the n x NB panel lives in registers, one row per thread (RPT rows for
//      n > 256); every column is pivoted (max |re|+|im| over the rows not yet
//      used as pivots, first index on ties, as LAPACK's IZAMAX) and swept
//      (pivot row scaled, all other rows eliminated).  Rows are exchanged by
//      ownership, so the row swap costs nothing.  Two barriers per column.
//   2. the swept panel W (rows J hold M_JJ^-1, others -M_OJ M_JJ^-1) is
//      written back and kept in LDS; the row exchanges of the panel are folded
//      into a gather permutation pi.
//   3. every other column strip of 16 is updated with MFMA:
//         new[:, q] = old[pi(:), q] (rows J zeroed) + W * old[pi(J), q];
//      software-pipelined (n <= 512): the next strip's pivot rows and row
//      gathers are loaded into registers while the current strip's MFMAs run.
//      Each wave owns MRT row tiles of 16 (MRT sized to n keeps two
//      workgroups per CU at n = 192).
// At the end, X[:, j] = W[:, sigma(j)] undoes the column order (sigma is the
// reverse composition of the row exchanges).
//
// LDS: W planes NB x npad (npad = n rounded up to 16), a 16-column strip of
// old pivot rows, the permutations and small exchange buffers.
#include "nt_kernels.h"

namespace nt {

namespace {

constexpr int NT = 256;
constexpr int QW = 16;

template <typename T>
struct Cx {
    T r, i;
};

template <typename T, int NB>
struct InvSmemFixed {
    T btr[NB][QW];
    T bti[NB][QW];
    T xr[2][2][NB];        // [buf][0 = pivot row, 1 = old row c][col]
    T xi[2][2][NB];
    double redv[2][4];
    int redi[2][4];
    int nan_flag;
    int info;
};

__device__ __forceinline__ int wave_of() { return threadIdx.x >> 6; }

template <typename T>
__device__ __forceinline__ T cabs1(T r, T i) { return fabs(r) + fabs(i); }

template <typename T, int NB, int RPT, int MRT>
__global__ __launch_bounds__(NT) void k_inv_gj(InvArgs<T> a)
{
    typedef typename acc4<T>::type A4;
    typedef typename vec2<T>::type V2;
    const int p = blockIdx.x;
    if (a.active && !a.active[p]) return;
    const int n = a.n;
    const int npad = (n + 15) & ~15;
    const int t = threadIdx.x;
    const int lane = lane_id();
    const int wave = wave_of();

    extern __shared__ __align__(16) unsigned char smem_raw[];
    InvSmemFixed<T, NB>& s = *reinterpret_cast<InvSmemFixed<T, NB>*>(smem_raw);
    T* wr = reinterpret_cast<T*>(smem_raw + sizeof(InvSmemFixed<T, NB>));   // [NB][npad]
    T* wi = wr + (long)NB * npad;
    int* pi = reinterpret_cast<int*>(wi + (long)NB * npad);                  // [n]
    int* ipiv = pi + n;                                                       // [n]
    int* sigma = ipiv + n;                                                    // [n]

    T* __restrict__ W = bptr(a.W, p, 0);
    const int ld = a.W.ld;

    if (t == 0) {
        s.nan_flag = 0;
        s.info = 0;
    }
    // Optional prologue: W = z I - src
    if (a.src.base) {
        const T* __restrict__ S = bptr(a.src, p, 0);
        const T zr = a.z[2 * p], zi = a.z[2 * p + 1];
        for (long e = t; e < (long)n * n; e += NT) {
            const int r = (int)(e % n), c = (int)(e / n);
            V2 v = *reinterpret_cast<const V2*>(S + 2 * ((long)r + (long)c * a.src.ld));
            V2 o;
            o.x = (r == c ? zr : (T)0) - v.x;
            o.y = (r == c ? zi : (T)0) - v.y;
            *reinterpret_cast<V2*>(W + 2 * ((long)r + (long)c * ld)) = o;
        }
    }
    // zero the padded rows of the W planes once
    for (int e = t; e < NB * (npad - n); e += NT) {
        const int c = e / (npad - n), r = n + e % (npad - n);
        wr[c * npad + r] = 0;
        wi[c * npad + r] = 0;
    }
    __syncthreads();

    Cx<T> x[RPT][NB];
    int step = 0;   // parity of the exchange buffers
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nbe = min(NB, n - j0);
        // ---- load panel rows into registers; pi restarts as the identity
        // (the previous panel's exchanges are already applied in memory)
        for (int r = t; r < n; r += NT) pi[r] = r;
#pragma unroll
        for (int sl = 0; sl < RPT; ++sl) {
            const int r = t + sl * NT;
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                Cx<T> v = {0, 0};
                if (r < n && c < nbe) {
                    V2 q = *reinterpret_cast<const V2*>(W + 2 * ((long)r + (long)(j0 + c) * ld));
                    v.r = q.x;
                    v.i = q.y;
                    if (q.x != q.x || q.y != q.y) s.nan_flag = 1;
                }
                x[sl][c] = v;
            }
        }
        __syncthreads();
        // ---- unblocked sweep of the panel
        for (int cl = 0; cl < nbe; ++cl) {
            const int c = j0 + cl;
            // local argmax over owned rows r >= c
            T bv = (T)-1;
            int bi = n;
#pragma unroll
            for (int sl = 0; sl < RPT; ++sl) {
                const int r = t + sl * NT;
                T v = (T)-1;
#pragma unroll
                for (int cc = 0; cc < NB; ++cc)
                    if (cc == cl) v = cabs1(x[sl][cc].r, x[sl][cc].i);
                if (r < n && r >= c && v > bv) {
                    bv = v;
                    bi = r;
                }
            }
            double dv = (double)bv;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(dv, off);
                const int oi = __shfl_xor(bi, off);
                if (ov > dv || (ov == dv && oi < bi)) {
                    dv = ov;
                    bi = oi;
                }
            }
            if (lane == 0) {
                s.redv[step][wave] = dv;
                s.redi[step][wave] = bi;
            }
            __syncthreads();
            double best = s.redv[step][0];
            int piv = s.redi[step][0];
#pragma unroll
            for (int w = 1; w < NT / 64; ++w) {
                const double ov = s.redv[step][w];
                const int oi = s.redi[step][w];
                if (ov > best || (ov == best && oi < piv)) {
                    best = ov;
                    piv = oi;
                }
            }
            if (piv >= n) piv = c;   // all candidates NaN: keep the diagonal
            // The pivot-row owner publishes the scaled row s_c' = pv_c' / pv_cl
            // (s_cl = 1 / pv_cl); the owner of row c publishes its old row.
#pragma unroll
            for (int sl = 0; sl < RPT; ++sl) {
                const int r = t + sl * NT;
                if (r == piv) {
                    T pr = 0, pim = 0;
#pragma unroll
                    for (int cc = 0; cc < NB; ++cc)
                        if (cc == cl) {
                            pr = x[sl][cc].r;
                            pim = x[sl][cc].i;
                        }
                    const T den = pr * pr + pim * pim;
                    T dr = 0, di = 0;
                    if (den != (T)0) {
                        dr = pr / den;
                        di = -pim / den;
                    } else if (s.info == 0) {
                        s.info = c + 1;
                    }
#pragma unroll
                    for (int cc = 0; cc < NB; ++cc) {
                        const T vr = x[sl][cc].r, vi = x[sl][cc].i;
                        s.xr[step][0][cc] = (cc == cl) ? dr : vr * dr - vi * di;
                        s.xi[step][0][cc] = (cc == cl) ? di : vr * di + vi * dr;
                    }
                }
                if (r == c) {
#pragma unroll
                    for (int cc = 0; cc < NB; ++cc) {
                        s.xr[step][1][cc] = x[sl][cc].r;
                        s.xi[step][1][cc] = x[sl][cc].i;
                    }
                }
            }
            if (t == 0) {
                ipiv[c] = piv;
                if (piv != c) {
                    const int tmp = pi[c];
                    pi[c] = pi[piv];
                    pi[piv] = tmp;
                }
            }
            __syncthreads();
#pragma unroll
            for (int sl = 0; sl < RPT; ++sl) {
                const int r = t + sl * NT;
                if (r >= n) continue;
                if (r == c) {
#pragma unroll
                    for (int cc = 0; cc < NB; ++cc) {
                        x[sl][cc].r = s.xr[step][0][cc];
                        x[sl][cc].i = s.xi[step][0][cc];
                    }
                } else {
                    if (r == piv) {
#pragma unroll
                        for (int cc = 0; cc < NB; ++cc) {
                            x[sl][cc].r = s.xr[step][1][cc];
                            x[sl][cc].i = s.xi[step][1][cc];
                        }
                    }
                    T fr = 0, fi = 0;
#pragma unroll
                    for (int cc = 0; cc < NB; ++cc)
                        if (cc == cl) {
                            fr = x[sl][cc].r;
                            fi = x[sl][cc].i;
                        }
#pragma unroll
                    for (int cc = 0; cc < NB; ++cc) {
                        const T sr = s.xr[step][0][cc], si = s.xi[step][0][cc];
                        const T br = (cc == cl) ? (T)0 : x[sl][cc].r;
                        const T bim = (cc == cl) ? (T)0 : x[sl][cc].i;
                        x[sl][cc].r = br - (fr * sr - fi * si);
                        x[sl][cc].i = bim - (fr * si + fi * sr);
                    }
                }
            }
            step ^= 1;
        }
        // ---- write panel back (global + LDS W planes)
#pragma unroll
        for (int sl = 0; sl < RPT; ++sl) {
            const int r = t + sl * NT;
            if (r < n) {
#pragma unroll
                for (int cc = 0; cc < NB; ++cc) {
                    if (cc < nbe) {
                        V2 q;
                        q.x = x[sl][cc].r;
                        q.y = x[sl][cc].i;
                        *reinterpret_cast<V2*>(W + 2 * ((long)r + (long)(j0 + cc) * ld)) = q;
                    }
                    wr[cc * npad + r] = (cc < nbe) ? x[sl][cc].r : (T)0;
                    wi[cc * npad + r] = (cc < nbe) ? x[sl][cc].i : (T)0;
                }
            }
        }
        // ---- trailing update of all other columns, strip by strip
        const int nrt = npad / 16;
        constexpr int BPT = (NB * QW + NT - 1) / NT;
        auto next_strip = [&](int q) {      // next strip not inside J, or n
            for (q += QW; q < n; q += QW)
                if (!(q >= j0 && q + QW <= j0 + nbe)) return q;
            return n;
        };
        T bpr[BPT], bpi[BPT];               // old pivot rows of the strip (zero beyond nbe / n)
        T cr[MRT][4], ci[MRT][4];           // accumulator init: old[pi(row)][col], rows outside J
        auto load_strip = [&](int qs) {
#pragma unroll
            for (int v = 0; v < BPT; ++v) {
                const int e = t + v * NT;
                const int k = e / QW, qq = e % QW;
                const int col = qs + qq;
                T vr = 0, vi = 0;
                if (e < NB * QW && k < nbe && col < n && !(col >= j0 && col < j0 + nbe)) {
                    const int row = pi[j0 + k];
                    V2 q = *reinterpret_cast<const V2*>(W + 2 * ((long)row + (long)col * ld));
                    vr = q.x;
                    vi = q.y;
                }
                bpr[v] = vr;
                bpi[v] = vi;
            }
#pragma unroll
            for (int u = 0; u < MRT; ++u) {
                const int rt = wave + 4 * u;
                const int row = rt * 16 + (lane & 15);
#pragma unroll
                for (int r4 = 0; r4 < 4; ++r4) {
                    const int col = qs + mfma_out_row<T>(lane, r4);
                    T vr = 0, vi = 0;
                    if (rt < nrt && row < n && col < n && !(row >= j0 && row < j0 + nbe)) {
                        const int src = pi[row];
                        V2 q = *reinterpret_cast<const V2*>(W + 2 * ((long)src + (long)col * ld));
                        vr = q.x;
                        vi = q.y;
                    }
                    cr[u][r4] = vr;
                    ci[u][r4] = vi;
                }
            }
        };
        // (with RPT > 2 the prefetch registers would spill: load in place)
        constexpr bool PIPE = RPT <= 2;
        int q0 = next_strip(-QW);
        if (PIPE && q0 < n) load_strip(q0);
        while (q0 < n) {
            if (!PIPE) load_strip(q0);
            __syncthreads();                // the previous strip's readers of btr are done
#pragma unroll
            for (int v = 0; v < BPT; ++v) {
                const int e = t + v * NT;
                if (e < NB * QW) {
                    s.btr[e / QW][e % QW] = bpr[v];
                    s.bti[e / QW][e % QW] = bpi[v];
                }
            }
            A4 accR[MRT], accI[MRT];
#pragma unroll
            for (int u = 0; u < MRT; ++u)
#pragma unroll
                for (int r4 = 0; r4 < 4; ++r4) {
                    accR[u][r4] = cr[u][r4];
                    accI[u][r4] = ci[u][r4];
                }
            // strips are disjoint column ranges: the next strip's loads never
            // touch the columns this one stores
            const int qn = next_strip(q0);
            if (PIPE && qn < n) load_strip(qn);
            __syncthreads();
#pragma unroll
            for (int u = 0; u < MRT; ++u) {
                const int rt = wave + 4 * u;
                if (rt >= nrt) continue;
                const int row = rt * 16 + (lane & 15);
#pragma unroll
                for (int kk = 0; kk < NB; kk += 4) {
                    const int k = kk + (lane >> 4);
                    const T ar = s.btr[k][lane & 15], ai = s.bti[k][lane & 15];
                    const T br = wr[k * npad + row], bim = wi[k * npad + row];
                    accR[u] = mfma16x16x4(ar, br, accR[u]);
                    accR[u] = mfma16x16x4(-ai, bim, accR[u]);
                    accI[u] = mfma16x16x4(ai, br, accI[u]);
                    accI[u] = mfma16x16x4(ar, bim, accI[u]);
                }
#pragma unroll
                for (int r4 = 0; r4 < 4; ++r4) {
                    const int col = q0 + mfma_out_row<T>(lane, r4);
                    if (row < n && col < n && !(col >= j0 && col < j0 + nbe)) {
                        V2 q;
                        q.x = accR[u][r4];
                        q.y = accI[u][r4];
                        *reinterpret_cast<V2*>(W + 2 * ((long)row + (long)col * ld)) = q;
                    }
                }
            }
            q0 = qn;
        }
        __syncthreads();
    }
    // ---- column order: sigma = reverse composition of the row exchanges
    if (t == 0) {
        for (int j = 0; j < n; ++j) sigma[j] = j;
        for (int c = n - 1; c >= 0; --c) {
            const int q = ipiv[c];
            if (q != c) {
                const int tmp = sigma[c];
                sigma[c] = sigma[q];
                sigma[q] = tmp;
            }
        }
        // merge into the caller's status: the first error of an item wins
        const int v = s.nan_flag ? -1 : s.info;
        if (a.info && v != 0 && a.info[p] == 0) a.info[p] = v;
    }
    __syncthreads();
    T* __restrict__ X = bptr(a.X, p, 0);
    for (long e = t; e < (long)n * n; e += NT) {
        const int r = (int)(e % n), j = (int)(e / n);
        V2 q = *reinterpret_cast<const V2*>(W + 2 * ((long)r + (long)sigma[j] * ld));
        *reinterpret_cast<V2*>(X + 2 * ((long)r + (long)j * a.X.ld)) = q;
    }
}

template <typename T, int NB>
size_t inv_smem_bytes(int n)
{
    const int npad = (n + 15) & ~15;
    return sizeof(InvSmemFixed<T, NB>) + 2 * sizeof(T) * (size_t)NB * npad + 3 * sizeof(int) * (size_t)n;
}

template <typename T, int NB, int RPT, int MRT>
hipError_t launch_inv_t(const InvArgs<T>& a, hipStream_t s)
{
    const size_t sm = inv_smem_bytes<T, NB>(a.n);
    // once per instantiation (a function-local static is initialised once,
    // also when several host threads launch concurrently)
    static const bool configured = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_inv_gj<T, NB, RPT, MRT>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        return true;
    }();
    (void)configured;
    hipLaunchKernelGGL((k_inv_gj<T, NB, RPT, MRT>), dim3(a.batch), dim3(NT), sm, s, a);
    return hipGetLastError();
}

// Panel width NB is chosen so the LDS copy of the panel stays under ~100 KB;
// RPT (rows per thread) = ceil(n / 256) must not exceed MAXRPT for that NB.
// MRT (row tiles of 16 per wave in the trailing update) = ceil(n / 64).
template <typename T, int NB, int MAXRPT>
hipError_t launch_inv_nb(const InvArgs<T>& a, hipStream_t s)
{
    const int rpt = (a.n + NT - 1) / NT;
    const int mrt = (a.n + 63) / 64;
    if (rpt <= 1) {
        if (mrt <= 1) return launch_inv_t<T, NB, 1, 1>(a, s);
        if (mrt == 2) return launch_inv_t<T, NB, 1, 2>(a, s);
        if (mrt == 3) return launch_inv_t<T, NB, 1, 3>(a, s);
        return launch_inv_t<T, NB, 1, 4>(a, s);
    }
    if constexpr (MAXRPT >= 2) if (rpt == 2) return launch_inv_t<T, NB, 2, 8>(a, s);
    if constexpr (MAXRPT >= 3) if (rpt == 3) return launch_inv_t<T, NB, 3, 12>(a, s);
    if constexpr (MAXRPT >= 4) if (rpt == 4) return launch_inv_t<T, NB, 4, 16>(a, s);
    return hipErrorInvalidValue;
}


// ------------------------------------------------------------------------
// Blocked path: the same Gauss-Jordan steps with the trailing update of every
// panel done by the batched GEMM (nt_gemm.hip) instead of inside the
// workgroup.  Per panel J (NB columns from j0):
//   k_inv_panel  sweeps the n x NB panel in registers as above, writes the
//                swept panel W to the panel columns and to Wb (n x NB), applies
//                the panel's row exchanges to every other column (explicitly,
//                in pivot order), copies the pivot rows of the other columns
//                to Pb (NB x n, zero in the panel columns) and zeroes them;
//   GEMM         C = C + Wb Pb  (in place), i.e. new[:, q] = old[:, q] with
//                rows J zeroed + W old[J, q] for every column outside J.
// k_inv_final undoes the column order (sigma) into X.  Wb, Pb and the pivot
// record live in X, which is only written by k_inv_final: per item, complex
// elements [0, n NB) hold Wb, [n NB, 2 n NB) Pb, then n ints ipiv and two
// status words (NaN seen, first zero pivot).

template <typename T>
struct InvScratch {
    __device__ static T* wb(T* X, int n, int nb) { return X; }
    __device__ static T* pb(T* X, int n, int nb) { return X + 2L * n * nb; }
    __device__ static int* ipiv(T* X, int n, int nb) { return reinterpret_cast<int*>(X + 4L * n * nb); }
};

template <typename T>
__global__ __launch_bounds__(NT) void k_inv_prologue(InvArgs<T> a)
{
    const int p = blockIdx.x;
    if (a.active && !a.active[p]) return;
    const int n = a.n;
    const long e = (long)blockIdx.y * blockDim.x + threadIdx.x;
    if (e >= (long)n * n) return;
    const int r = (int)(e % n), c = (int)(e / n);
    const T* __restrict__ S = bptr(a.src, p, 0);
    T* __restrict__ W = bptr(a.W, p, 0);
    const T zr = a.z[2 * p], zi = a.z[2 * p + 1];
    typedef typename vec2<T>::type V2;
    V2 v = *reinterpret_cast<const V2*>(S + 2 * ((long)r + (long)c * a.src.ld));
    V2 o;
    o.x = (r == c ? zr : (T)0) - v.x;
    o.y = (r == c ? zi : (T)0) - v.y;
    *reinterpret_cast<V2*>(W + 2 * ((long)r + (long)c * a.W.ld)) = o;
}

template <typename T, int NB, int RPT>
__global__ __launch_bounds__(NT) void k_inv_panel(InvArgs<T> a, int j0)
{
    typedef typename vec2<T>::type V2;
    const int p = blockIdx.x;
    if (a.active && !a.active[p]) return;
    const int n = a.n;
    const int nbe = min(NB, n - j0);
    const int t = threadIdx.x;
    const int lane = lane_id();
    const int wave = wave_of();
    __shared__ InvSmemFixed<T, NB> s;
    __shared__ int piv_s[NB];

    T* __restrict__ C = bptr(a.W, p, 0);
    const int ld = a.W.ld;
    T* __restrict__ X = bptr(a.X, p, 0);
    T* __restrict__ Wb = InvScratch<T>::wb(X, n, NB);
    T* __restrict__ Pb = InvScratch<T>::pb(X, n, NB);
    int* __restrict__ ipiv = InvScratch<T>::ipiv(X, n, NB);
    int* __restrict__ stat = ipiv + n;     // [0] NaN seen, [1] first zero pivot (1-based)

    if (t == 0) {
        s.nan_flag = j0 == 0 ? 0 : stat[0];
        s.info = j0 == 0 ? 0 : stat[1];
    }
    __syncthreads();
    Cx<T> x[RPT][NB];
#pragma unroll
    for (int sl = 0; sl < RPT; ++sl) {
        const int r = t + sl * NT;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            Cx<T> v = {0, 0};
            if (r < n && c < nbe) {
                V2 q = *reinterpret_cast<const V2*>(C + 2 * ((long)r + (long)(j0 + c) * ld));
                v.r = q.x;
                v.i = q.y;
                if (q.x != q.x || q.y != q.y) s.nan_flag = 1;
            }
            x[sl][c] = v;
        }
    }
    __syncthreads();
    int step = 0;
    for (int cl = 0; cl < nbe; ++cl) {
        const int c = j0 + cl;
        T bv = (T)-1;
        int bi = n;
#pragma unroll
        for (int sl = 0; sl < RPT; ++sl) {
            const int r = t + sl * NT;
            T v = (T)-1;
#pragma unroll
            for (int cc = 0; cc < NB; ++cc)
                if (cc == cl) v = cabs1(x[sl][cc].r, x[sl][cc].i);
            if (r < n && r >= c && v > bv) {
                bv = v;
                bi = r;
            }
        }
        double dv = (double)bv;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(dv, off);
            const int oi = __shfl_xor(bi, off);
            if (ov > dv || (ov == dv && oi < bi)) {
                dv = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            s.redv[step][wave] = dv;
            s.redi[step][wave] = bi;
        }
        __syncthreads();
        double best = s.redv[step][0];
        int piv = s.redi[step][0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) {
            const double ov = s.redv[step][w];
            const int oi = s.redi[step][w];
            if (ov > best || (ov == best && oi < piv)) {
                best = ov;
                piv = oi;
            }
        }
        if (piv >= n) piv = c;
#pragma unroll
        for (int sl = 0; sl < RPT; ++sl) {
            const int r = t + sl * NT;
            if (r == piv) {
                T pr = 0, pim = 0;
#pragma unroll
                for (int cc = 0; cc < NB; ++cc)
                    if (cc == cl) {
                        pr = x[sl][cc].r;
                        pim = x[sl][cc].i;
                    }
                const T den = pr * pr + pim * pim;
                T dr = 0, di = 0;
                if (den != (T)0) {
                    dr = pr / den;
                    di = -pim / den;
                } else if (s.info == 0) {
                    s.info = c + 1;
                }
#pragma unroll
                for (int cc = 0; cc < NB; ++cc) {
                    const T vr = x[sl][cc].r, vi = x[sl][cc].i;
                    s.xr[step][0][cc] = (cc == cl) ? dr : vr * dr - vi * di;
                    s.xi[step][0][cc] = (cc == cl) ? di : vr * di + vi * dr;
                }
            }
            if (r == c) {
#pragma unroll
                for (int cc = 0; cc < NB; ++cc) {
                    s.xr[step][1][cc] = x[sl][cc].r;
                    s.xi[step][1][cc] = x[sl][cc].i;
                }
            }
        }
        if (t == 0) piv_s[cl] = piv;
        __syncthreads();
#pragma unroll
        for (int sl = 0; sl < RPT; ++sl) {
            const int r = t + sl * NT;
            if (r >= n) continue;
            if (r == c) {
#pragma unroll
                for (int cc = 0; cc < NB; ++cc) {
                    x[sl][cc].r = s.xr[step][0][cc];
                    x[sl][cc].i = s.xi[step][0][cc];
                }
            } else {
                if (r == piv) {
#pragma unroll
                    for (int cc = 0; cc < NB; ++cc) {
                        x[sl][cc].r = s.xr[step][1][cc];
                        x[sl][cc].i = s.xi[step][1][cc];
                    }
                }
                T fr = 0, fi = 0;
#pragma unroll
                for (int cc = 0; cc < NB; ++cc)
                    if (cc == cl) {
                        fr = x[sl][cc].r;
                        fi = x[sl][cc].i;
                    }
#pragma unroll
                for (int cc = 0; cc < NB; ++cc) {
                    const T sr = s.xr[step][0][cc], si = s.xi[step][0][cc];
                    const T br = (cc == cl) ? (T)0 : x[sl][cc].r;
                    const T bim = (cc == cl) ? (T)0 : x[sl][cc].i;
                    x[sl][cc].r = br - (fr * sr - fi * si);
                    x[sl][cc].i = bim - (fr * si + fi * sr);
                }
            }
        }
        step ^= 1;
    }
    // swept panel -> panel columns of C and Wb (zero columns beyond nbe)
#pragma unroll
    for (int sl = 0; sl < RPT; ++sl) {
        const int r = t + sl * NT;
        if (r < n) {
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) {
                V2 q;
                q.x = cc < nbe ? x[sl][cc].r : (T)0;
                q.y = cc < nbe ? x[sl][cc].i : (T)0;
                if (cc < nbe) *reinterpret_cast<V2*>(C + 2 * ((long)r + (long)(j0 + cc) * ld)) = q;
                *reinterpret_cast<V2*>(Wb + 2 * ((long)r + (long)cc * n)) = q;
            }
        }
    }
    __threadfence_block();
    // row exchanges of the panel, in pivot order, on every other column
    for (int cl = 0; cl < nbe; ++cl) {
        const int c = j0 + cl, q = piv_s[cl];
        if (q != c) {
            for (int col = t; col < n; col += NT) {
                if (col >= j0 && col < j0 + nbe) continue;
                V2* a0 = reinterpret_cast<V2*>(C + 2 * ((long)c + (long)col * ld));
                V2* a1 = reinterpret_cast<V2*>(C + 2 * ((long)q + (long)col * ld));
                const V2 u = *a0, v = *a1;
                *a0 = v;
                *a1 = u;
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    // pivot rows of the other columns -> Pb (NB x n), then zeroed in C
    for (long e = t; e < (long)NB * n; e += NT) {
        const int k = (int)(e % NB), col = (int)(e / NB);
        V2 q;
        q.x = 0;
        q.y = 0;
        if (k < nbe && !(col >= j0 && col < j0 + nbe)) {
            V2* src = reinterpret_cast<V2*>(C + 2 * ((long)(j0 + k) + (long)col * ld));
            q = *src;
            V2 z;
            z.x = 0;
            z.y = 0;
            *src = z;
        }
        *reinterpret_cast<V2*>(Pb + 2 * e) = q;
    }
    if (t < nbe) ipiv[j0 + t] = piv_s[t];
    if (t == 0) {
        stat[0] = s.nan_flag;
        stat[1] = s.info;
    }
}

template <typename T, int NB>
__global__ __launch_bounds__(NT) void k_inv_final(InvArgs<T> a)
{
    typedef typename vec2<T>::type V2;
    const int p = blockIdx.x;
    if (a.active && !a.active[p]) return;
    const int n = a.n;
    __shared__ int sigma[1024];
    T* __restrict__ C = bptr(a.W, p, 0);
    T* __restrict__ X = bptr(a.X, p, 0);
    const int* ipiv = InvScratch<T>::ipiv(X, n, NB);
    if (threadIdx.x == 0) {
        for (int j = 0; j < n; ++j) sigma[j] = j;
        for (int c = n - 1; c >= 0; --c) {
            const int q = ipiv[c];
            if (q != c) {
                const int tmp = sigma[c];
                sigma[c] = sigma[q];
                sigma[q] = tmp;
            }
        }
        const int* stat = ipiv + n;
        const int v = stat[0] ? -1 : stat[1];
        if (a.info && v != 0 && a.info[p] == 0) a.info[p] = v;
    }
    __syncthreads();
    for (long e = threadIdx.x; e < (long)n * n; e += NT) {
        const int r = (int)(e % n), j = (int)(e / n);
        V2 q = *reinterpret_cast<const V2*>(C + 2 * ((long)r + (long)sigma[j] * a.W.ld));
        *reinterpret_cast<V2*>(X + 2 * ((long)r + (long)j * a.X.ld)) = q;
    }
}

template <typename T, int NB, int RPT>
hipError_t launch_inv_blocked_t(const InvArgs<T>& a, hipStream_t s)
{
    const int n = a.n;
    if (a.src.base) {
        hipLaunchKernelGGL(k_inv_prologue<T>, dim3(a.batch, ((long)n * n + NT - 1) / NT), dim3(NT), 0, s, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    GemmArgs<T> g;
    g.M = n;
    g.N = n;
    g.K = NB;
    g.opA = OP_N;
    g.opB = OP_N;
    g.A = BMat<T>{a.X.base, a.X.stride, 0, nullptr, n, 0};
    g.B = BMat<T>{a.X.base + 2L * n * NB, a.X.stride, 0, nullptr, NB, 0};
    g.Cin = BMat<T>{a.W.base, a.W.stride, 0, a.W.idx, a.W.ld, a.W.idx_n};
    g.C = a.W;
    g.C2 = BMatW<T>{nullptr, 0, 0, nullptr, 0, 0};
    g.alpha_re = 1;
    g.alpha_im = 0;
    g.beta_re = 1;
    g.beta_im = 0;
    g.diag = nullptr;
    g.active = a.active;
    g.maxabs = nullptr;
    g.batch = a.batch;
    g.groups = 1;
    g.m3 = 0;
    for (int j0 = 0; j0 < n; j0 += NB) {
        hipLaunchKernelGGL((k_inv_panel<T, NB, RPT>), dim3(a.batch), dim3(NT), 0, s, a, j0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = launch_gemm<T>(g, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_inv_final<T, NB>), dim3(a.batch), dim3(NT), 0, s, a);
    return hipGetLastError();
}

}  // namespace

int inv_max_n() { return 1024; }

// Largest n each panel width supports (LDS copy of the panel <= ~130 KB and
// rows per thread within the kernel's instantiation); width INV_BLOCKED (1)
// is the panel-kernel + batched-GEMM path, for 64 <= n <= 1024.
template <typename T>
int inv_panel_widths(int n, int* out)
{
    const int lim32 = sizeof(T) == 8 ? 192 : 256, lim16 = sizeof(T) == 8 ? 384 : 512;
    int k = 0;
    if (n <= lim32) out[k++] = 32;
    if (n <= 256) out[k++] = 24;
    if (n <= 256) out[k++] = 20;
    if (n <= lim16) out[k++] = 16;
    if (n <= inv_max_n()) out[k++] = 8;
    if (n >= 64 && n <= inv_max_n()) out[k++] = INV_BLOCKED;
    return k;
}
template int inv_panel_widths<double>(int, int*);
template int inv_panel_widths<float>(int, int*);

// The blocked path keeps its scratch in X: X must be compact (ld = n, item
// stride >= n^2, no item map) and n large enough for 2 n NB complex + n + 2
// ints.  Panel width 32 for n <= 256 (one row per thread), 16 up to 512, else 8.
template <typename T>
bool inv_blocked_ok(const InvArgs<T>& a)
{
    const int nb = a.n <= 256 ? 32 : (a.n <= 512 ? 16 : 8);
    const double need = 2.0 * a.n * nb * 2 * sizeof(T) + sizeof(int) * (a.n + 2.0);
    return a.X.ld == a.n && !a.X.idx && a.X.stride >= (long)a.n * a.n && a.n >= 64 && a.n <= 1024 &&
           need <= 2.0 * sizeof(T) * a.n * a.n;
}

template <typename T>
hipError_t launch_inv_blocked(const InvArgs<T>& a, hipStream_t s)
{
    const int rpt = (a.n + NT - 1) / NT;
    if (rpt == 1) return launch_inv_blocked_t<T, 32, 1>(a, s);
    if (rpt == 2) return launch_inv_blocked_t<T, 16, 2>(a, s);
    if (rpt == 3) return launch_inv_blocked_t<T, 8, 3>(a, s);
    return launch_inv_blocked_t<T, 8, 4>(a, s);
}

template <typename T>
hipError_t launch_inv_any(const InvArgs<T>& a, hipStream_t s)
{
    if (a.batch <= 0 || a.n <= 0) return hipSuccess;
    if (a.n > inv_max_n()) return hipErrorInvalidValue;
    int w[INV_MAX_WIDTHS];
    const int k = inv_panel_widths<T>(a.n, w);
    // default: 16 columns where it fits (measured fastest at b = 192: two
    // workgroups per CU against one with 32), else the widest that fits; the
    // panel + GEMM path from n = 320 (1.6x the 16-column kernel at n = 384)
    int nb = w[0];
    for (int i = 0; i < k; ++i)
        if (w[i] == 16) nb = 16;
    if (a.n >= 320 && inv_blocked_ok(a)) nb = INV_BLOCKED;
    for (int i = 0; i < k; ++i)
        if (a.nb == w[i]) nb = a.nb;
    if (nb == INV_BLOCKED) {
        if (inv_blocked_ok(a)) return launch_inv_blocked<T>(a, s);
        nb = w[0] == INV_BLOCKED ? 8 : w[0];    // scratch does not fit in X: one-workgroup path
    }
    if (nb == 32) return launch_inv_nb<T, 32, 1>(a, s);
    if (nb == 24) return launch_inv_nb<T, 24, 1>(a, s);
    if (nb == 20) return launch_inv_nb<T, 20, 1>(a, s);
    if (nb == 16) return launch_inv_nb<T, 16, 2>(a, s);
    return launch_inv_nb<T, 8, 4>(a, s);
}

template <>
hipError_t launch_inv<double>(const InvArgs<double>& a, hipStream_t s)
{
    return launch_inv_any<double>(a, s);
}

template <>
hipError_t launch_inv<float>(const InvArgs<float>& a, hipStream_t s)
{
    return launch_inv_any<float>(a, s);
}

}  // namespace nt
