// nt_gemm.hip -- batched complex GEMM on CDNA4 MFMA (FP64 and FP32).
//
// One workgroup computes a TB x TB complex tile of one batch item: TB = 64
// with 4 wave64 (the default), or TB = 32 with one wave64 (less padding when
// the matrix size is not a multiple of 64: 24, 96, 216, ...; chosen per size
// by the autotuner).  Operand tiles of depth 8 are staged through two LDS
// buffers as separate real and imaginary planes (one barrier per stage); each
// wave owns a 32x32 quadrant as 2x2 MFMA 16x16 tiles and issues four real
// MFMAs per complex 16x16x4 step.
//
// The product is formed transposed (C^T = B^T A^T): the MFMA output layout
// then maps 16 consecutive lanes to 16 consecutive rows of the column-major C,
// so every store is a 256-byte (FP64) contiguous run.
#include "nt_kernels.h"

namespace nt {

namespace {

constexpr int BK = 8;         // operand depth per LDS stage
constexpr int NXCD = 8;       // accelerator complex dies of the MI355X

// threads of a TB x TB tile's workgroup and operand elements each stages per stage
template <int TB> struct Tile {
    static constexpr int NT = TB == 64 ? 256 : 64;
    static constexpr int EPT = TB * BK / NT;
};

// two stages of operand tiles: the MFMAs of one read the buffer the other's
// loads are not writing, so each K-stage needs a single barrier
template <typename T, int TB>
struct GemmSmem {
    T ar[2][BK][TB];
    T ai[2][BK][TB];
    T br[2][BK][TB];
    T bi[2][BK][TB];
};

// Element j of thread t.  op N (A) / op C (B): row r = t % TB (contiguous in
// memory), depth k = t / TB + (NT / TB) j.  op C (A) / op N (B): depth
// k = t % BK (contiguous), row r = t / BK + (NT / BK) j.
template <typename T, int TB>
__device__ __forceinline__ void load_a_regs(const T* __restrict__ A, int lda, int opA, int M, int K,
                                            int m0, int k0, T (&r)[Tile<TB>::EPT], T (&i)[Tile<TB>::EPT])
{
    constexpr int NT = Tile<TB>::NT, EPT = Tile<TB>::EPT;
    const int t = threadIdx.x;
    typedef typename vec2<T>::type V2;
    if (opA == OP_N) {
        const int m = t % TB, kq = t / TB;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int k = kq + (NT / TB) * j;
            const int gm = m0 + m, gk = k0 + k;
            V2 v = {0, 0};
            if (gm < M && gk < K) v = *reinterpret_cast<const V2*>(A + 2 * ((long)gm + (long)gk * lda));
            r[j] = v.x;
            i[j] = v.y;
        }
    } else {
        const int k = t % BK, mq = t / BK;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int m = mq + (NT / BK) * j;
            const int gm = m0 + m, gk = k0 + k;
            V2 v = {0, 0};
            if (gm < M && gk < K) v = *reinterpret_cast<const V2*>(A + 2 * ((long)gk + (long)gm * lda));
            r[j] = v.x;
            i[j] = -v.y;
        }
    }
}

template <typename T, int TB>
__device__ __forceinline__ void store_a_smem(GemmSmem<T, TB>& s, int buf, int opA, const T (&r)[Tile<TB>::EPT],
                                             const T (&i)[Tile<TB>::EPT])
{
    constexpr int NT = Tile<TB>::NT, EPT = Tile<TB>::EPT;
    const int t = threadIdx.x;
    if (opA == OP_N) {
        const int m = t % TB, kq = t / TB;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            s.ar[buf][kq + (NT / TB) * j][m] = r[j];
            s.ai[buf][kq + (NT / TB) * j][m] = i[j];
        }
    } else {
        const int k = t % BK, mq = t / BK;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            s.ar[buf][k][mq + (NT / BK) * j] = r[j];
            s.ai[buf][k][mq + (NT / BK) * j] = i[j];
        }
    }
}

template <typename T, int TB>
__device__ __forceinline__ void load_b_regs(const T* __restrict__ B, int ldb, int opB, int K, int N,
                                            int k0, int n0, T (&r)[Tile<TB>::EPT], T (&i)[Tile<TB>::EPT])
{
    constexpr int NT = Tile<TB>::NT, EPT = Tile<TB>::EPT;
    const int t = threadIdx.x;
    typedef typename vec2<T>::type V2;
    if (opB == OP_N) {
        const int k = t % BK, nq = t / BK;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int n = nq + (NT / BK) * j;
            const int gk = k0 + k, gn = n0 + n;
            V2 v = {0, 0};
            if (gk < K && gn < N) v = *reinterpret_cast<const V2*>(B + 2 * ((long)gk + (long)gn * ldb));
            r[j] = v.x;
            i[j] = v.y;
        }
    } else {
        const int n = t % TB, kq = t / TB;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int k = kq + (NT / TB) * j;
            const int gk = k0 + k, gn = n0 + n;
            V2 v = {0, 0};
            if (gk < K && gn < N) v = *reinterpret_cast<const V2*>(B + 2 * ((long)gn + (long)gk * ldb));
            r[j] = v.x;
            i[j] = -v.y;
        }
    }
}

template <typename T, int TB>
__device__ __forceinline__ void store_b_smem(GemmSmem<T, TB>& s, int buf, int opB, const T (&r)[Tile<TB>::EPT],
                                             const T (&i)[Tile<TB>::EPT])
{
    constexpr int NT = Tile<TB>::NT, EPT = Tile<TB>::EPT;
    const int t = threadIdx.x;
    if (opB == OP_N) {
        const int k = t % BK, nq = t / BK;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            s.br[buf][k][nq + (NT / BK) * j] = r[j];
            s.bi[buf][k][nq + (NT / BK) * j] = i[j];
        }
    } else {
        const int n = t % TB, kq = t / TB;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            s.br[buf][kq + (NT / TB) * j][n] = r[j];
            s.bi[buf][kq + (NT / TB) * j][n] = i[j];
        }
    }
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* addr, double v)
{
    atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

// M3: complex products by three real products (P1 = Ar Br, P2 = Ai Bi,
// P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2, Im = P3 - P1 - P2) instead of four.
// Occupancy: three (M3: two) 256-thread workgroups per CU, the same number of
// waves per SIMD with one-wave workgroups.
template <typename T, bool M3, int TB>
__global__ __launch_bounds__(Tile<TB>::NT, (M3 ? 2 : 3) * (256 / Tile<TB>::NT)) void k_gemm(GemmArgs<T> a)
{
    constexpr int EPT = Tile<TB>::EPT;
    constexpr int BM = TB, BN = TB;
    typedef typename acc4<T>::type A4;
    const int tiles_m = (a.M + BM - 1) / BM;
    const int tiles_n = (a.N + BN - 1) / BN;
    const int tiles = tiles_m * tiles_n;
    // Workgroups are dispatched round-robin over the 8 XCDs, each with its own
    // L2.  From 8 points on, renumber so that point p (all its groups and
    // tiles) runs on XCD p mod 8, its blocks consecutive in that XCD's
    // sequence: the tiles of one item then share that XCD's L2 (its A and B
    // panels are fetched from HBM once), and points masked out by
    // convergence, which come in contiguous ranges of energy, leave the
    // remaining work spread over all XCDs.  The grid then covers the point
    // count rounded up to 8; the blocks of the padding points exit.
    long bid = blockIdx.x;
    if (a.batch >= NXCD) {
        const long per_point = (long)tiles * a.groups;
        const long x = bid % NXCD, l = bid / NXCD;
        const long pp = x + NXCD * (l / per_point);
        if (pp >= a.batch) return;
        bid = pp * per_point + l % per_point;
    }
    const int tile = (int)(bid % tiles);
    const long item = bid / tiles;
    const int p = (int)(item / a.groups);
    const int g = (int)(item % a.groups);
    if (a.active && !a.active[p]) return;
    const int m0 = (tile % tiles_m) * BM;
    const int n0 = (tile / tiles_m) * BN;

    __shared__ GemmSmem<T, TB> s;
    const T* __restrict__ A = bptr(a.A, p, g);
    const T* __restrict__ B = bptr(a.B, p, g);

    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;       // 0 when TB = 32
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;

    A4 accR[2][2], accI[2][2], accS[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            accR[x][y] = A4{0, 0, 0, 0};
            accI[x][y] = A4{0, 0, 0, 0};
            accS[x][y] = A4{0, 0, 0, 0};
        }

    T ra[EPT], ia[EPT], rb[EPT], ib[EPT];
    const int nst = (a.K + BK - 1) / BK;
    load_a_regs<T, TB>(A, a.A.ld, a.opA, a.M, a.K, m0, 0, ra, ia);
    load_b_regs<T, TB>(B, a.B.ld, a.opB, a.K, a.N, 0, n0, rb, ib);
    store_a_smem<T, TB>(s, 0, a.opA, ra, ia);
    store_b_smem<T, TB>(s, 0, a.opB, rb, ib);
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int cur = st & 1;
        const bool more = st + 1 < nst;
        if (more) {
            load_a_regs<T, TB>(A, a.A.ld, a.opA, a.M, a.K, m0, (st + 1) * BK, ra, ia);
            load_b_regs<T, TB>(B, a.B.ld, a.opB, a.K, a.N, (st + 1) * BK, n0, rb, ib);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int k = kk + lk;
            T xr[2], xi[2], yr[2], yi[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                xr[x] = s.ar[cur][k][wm + 16 * x + lr];
                xi[x] = s.ai[cur][k][wm + 16 * x + lr];
            }
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                yr[y] = s.br[cur][k][wn + 16 * y + lr];
                yi[y] = s.bi[cur][k][wn + 16 * y + lr];
            }
            if constexpr (M3) {
                T xs[2], ys[2];
#pragma unroll
                for (int x = 0; x < 2; ++x) xs[x] = xr[x] + xi[x];
#pragma unroll
                for (int y = 0; y < 2; ++y) ys[y] = yr[y] + yi[y];
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) {
                        // accR = P1, accI = P2, accS = P3
                        accR[x][y] = mfma16x16x4(yr[y], xr[x], accR[x][y]);
                        accI[x][y] = mfma16x16x4(yi[y], xi[x], accI[x][y]);
                        accS[x][y] = mfma16x16x4(ys[y], xs[x], accS[x][y]);
                    }
            } else {
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) {
                        // (C^T) = B^T A^T: MFMA "A" operand is B, "B" operand is A.
                        accR[x][y] = mfma16x16x4(yr[y], xr[x], accR[x][y]);
                        accR[x][y] = mfma16x16x4(-yi[y], xi[x], accR[x][y]);
                        accI[x][y] = mfma16x16x4(yi[y], xr[x], accI[x][y]);
                        accI[x][y] = mfma16x16x4(yr[y], xi[x], accI[x][y]);
                    }
            }
        }
        // the other buffer was last read in the previous stage, before its barrier
        if (more) {
            store_a_smem<T, TB>(s, cur ^ 1, a.opA, ra, ia);
            store_b_smem<T, TB>(s, cur ^ 1, a.opB, rb, ib);
        }
        __syncthreads();
    }

    // Epilogue.  Lane holds D[n = mfma_out_row(lane, r)][m = lr] of each sub-tile.
    T* __restrict__ C = bptr(a.C, p, g);
    const T* __restrict__ Cin = a.Cin.base ? bptr(a.Cin, p, g) : nullptr;
    T* __restrict__ C2 = a.C2.base ? bptr(a.C2, p, g) : nullptr;
    T dr = 0, di = 0;
    if (a.diag) {
        dr = a.diag[2 * p];
        di = a.diag[2 * p + 1];
    }
    double vmax = 0.0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * x + lr;
                const int n = n0 + wn + 16 * y + mfma_out_row<T>(lane, r);
                if (m < a.M && n < a.N) {
                    T pr, pi;
                    if constexpr (M3) {
                        const T p1 = accR[x][y][r], p2 = accI[x][y][r], p3 = accS[x][y][r];
                        pr = p1 - p2;
                        pi = p3 - p1 - p2;
                    } else {
                        pr = accR[x][y][r];
                        pi = accI[x][y][r];
                    }
                    T vr = a.alpha_re * pr - a.alpha_im * pi;
                    T vi = a.alpha_re * pi + a.alpha_im * pr;
                    T cr = vr, ci = vi;
                    if (Cin) {
                        const T* c = Cin + 2 * ((long)m + (long)n * a.Cin.ld);
                        const T qr = c[0], qi = c[1];
                        cr += a.beta_re * qr - a.beta_im * qi;
                        ci += a.beta_re * qi + a.beta_im * qr;
                    }
                    if (m == n) {
                        cr += dr;
                        ci += di;
                    }
                    T* o = C + 2 * ((long)m + (long)n * a.C.ld);
                    o[0] = cr;
                    o[1] = ci;
                    if (C2) {
                        T* o2 = C2 + 2 * ((long)m + (long)n * a.C2.ld);
                        o2[0] += vr;
                        o2[1] += vi;
                    }
                    if (a.maxabs) {
                        const double ab = sqrt((double)cr * (double)cr + (double)ci * (double)ci);
                        vmax = (ab > vmax || ab != ab) ? ab : vmax;
                    }
                }
            }
    if (a.maxabs && g == 0) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(vmax, off);
            vmax = (o > vmax || o != o) ? o : vmax;
        }
        if (lane == 0) {
            // NaN is encoded as +inf so that it always wins and fails the convergence test.
            atomic_max_nonneg(&a.maxabs[p], vmax != vmax ? __longlong_as_double(0x7ff0000000000000LL) : vmax);
        }
    }
}

}  // namespace

template <typename T>
hipError_t launch_gemm(const GemmArgs<T>& a, hipStream_t s)
{
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0 || a.groups <= 0) return hipSuccess;
    if (a.tile != 0 && a.tile != 32 && a.tile != 64) return hipErrorInvalidValue;
    const int tb = a.tile == 32 ? 32 : 64;
    const long tiles = (long)((a.M + tb - 1) / tb) * ((a.N + tb - 1) / tb);
    // from 8 points on the kernel interleaves points over the XCDs and the
    // grid covers the point count rounded up to a multiple of 8
    const long points = a.batch >= NXCD ? ((long)a.batch + NXCD - 1) / NXCD * NXCD : a.batch;
    const long blocks = tiles * points * a.groups;
    if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
    if (tb == 32) {
        if (a.m3)
            hipLaunchKernelGGL((k_gemm<T, true, 32>), dim3((unsigned)blocks), dim3(Tile<32>::NT), 0, s, a);
        else
            hipLaunchKernelGGL((k_gemm<T, false, 32>), dim3((unsigned)blocks), dim3(Tile<32>::NT), 0, s, a);
    } else {
        if (a.m3)
            hipLaunchKernelGGL((k_gemm<T, true, 64>), dim3((unsigned)blocks), dim3(Tile<64>::NT), 0, s, a);
        else
            hipLaunchKernelGGL((k_gemm<T, false, 64>), dim3((unsigned)blocks), dim3(Tile<64>::NT), 0, s, a);
    }
    return hipGetLastError();
}

template hipError_t launch_gemm<double>(const GemmArgs<double>&, hipStream_t);
template hipError_t launch_gemm<float>(const GemmArgs<float>&, hipStream_t);

}  // namespace nt
