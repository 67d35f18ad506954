// nt_mixed.hip -- mixed-precision Caroli solve with FP64 iterative refinement.
//
// Each sweep direction of get_G_full_tdb (negf/lesser.f90:555-800) inverts a
// block-tridiagonal system A, of which T needs one block column:
//   layers s = 0..S in sweep order (device layers, then the closing lead
//   layer), A_ss = z - Dm_s, A_s,s+1 = -C_s, A_s+1,s = -C_s^+,
//   Dm_0 = H_first + V^+ g V (the opening lead folded in, FP64),
//   Dm_S = H_lead + Sigma_lead (closing layer), X = A^-1 e_S (n_S columns).
// The engine
//   1. factorises A in complex FP32 on MFMA: G_s = (z - Dm_s - C_s-1^+ P_s-1)^-1,
//      P_s = G_s C_s (the left-connected blocks of the FP64 sweep, in FP32);
//   2. forms X in FP32 (X_S = G_S, X_s = P_s X_s+1) and stores it in FP64;
//   3. refines: R = e_S - A X in FP64 (three FP64 GEMMs per layer), solves
//      A dX = R with the FP32 factors (forward y_s = G_s (R_s + C_s-1^+ y_s-1),
//      backward dX_s = y_s + P_s dX_s+1) and adds dX to X in FP64;
//   4. measures delta = max|dX| / max|X| per point.  Refinement contracts the
//      error by a factor of the order of the initial error, so after one step
//      the remaining error is about delta^2; points with delta > 1e-4 get a
//      second step, after which the estimate is delta_2^2 / delta_1.  A point
//      whose estimate exceeds 1e-8 (or whose FP32 factorisation failed) is
//      flagged, and the caller recomputes it on the FP64 sweep.
// G_0,N+1 (or G_N+1,0) = g V X_0 and the Caroli traces are then FP64.
#include "nt_host.h"

namespace {

constexpr int NTE = 256;
constexpr double MIXED_SECOND_STEP = 1.0e-4;   // delta above which a second refinement step runs
constexpr double MIXED_ACCEPT = 1.0e-8;        // largest accepted estimate of the remaining relative error

inline unsigned nb1(long n) { return (unsigned)std::max<long>(1, (n + NTE - 1) / NTE); }

__device__ __forceinline__ void amax_nonneg(unsigned long long* a, double v)
{
    // NaN is recorded as +inf so that it always fails the acceptance test
    if (v != v) v = __longlong_as_double(0x7ff0000000000000LL);
    atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// out = A + B (n x n), grid.y = item
__global__ __launch_bounds__(NTE) void k_add2(int n, BMat<double> A, BMat<double> B, BMatW<double> out)
{
    const int p = blockIdx.y;
    const long e = (long)blockIdx.x * NTE + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e % n), j = (int)(e / n);
    const double* a = bptr(A, p, 0) + 2 * (i + (long)j * A.ld);
    const double* b = bptr(B, p, 0) + 2 * (i + (long)j * B.ld);
    double* o = bptr(out, p, 0) + 2 * (i + (long)j * out.ld);
    o[0] = a[0] + b[0];
    o[1] = a[1] + b[1];
}

// r32 = R - z X + (unit ? I : 0)   (n x m; R, X FP64; r32 FP32)
__global__ __launch_bounds__(NTE) void k_resid(int n, int m, BMat<double> R, BMat<double> X, const double* z,
                                               int unit, BMatW<float> out, const int* active)
{
    const int p = blockIdx.y;
    if (active && !active[p]) return;
    const long e = (long)blockIdx.x * NTE + threadIdx.x;
    if (e >= (long)n * m) return;
    const int i = (int)(e % n), j = (int)(e / n);
    const double* r = bptr(R, p, 0) + 2 * (i + (long)j * R.ld);
    const double* x = bptr(X, p, 0) + 2 * (i + (long)j * X.ld);
    const double zr = z[2 * p], zi = z[2 * p + 1];
    double vr = r[0] - (zr * x[0] - zi * x[1]);
    double vi = r[1] - (zr * x[1] + zi * x[0]);
    if (unit && i == j) vr += 1.0;
    float* o = bptr(out, p, 0) + 2 * (i + (long)j * out.ld);
    o[0] = (float)vr;
    o[1] = (float)vi;
}

// X += dX (FP32 correction into the FP64 solution); per point max |dX| and
// max |X| after the update
__global__ __launch_bounds__(NTE) void k_xupdate(int n, int m, BMat<float> D, BMatW<double> X, const int* active,
                                                 unsigned long long* dmax, unsigned long long* xmax)
{
    const int p = blockIdx.y;
    if (active && !active[p]) return;
    const long e = (long)blockIdx.x * NTE + threadIdx.x;
    double dv = 0.0, xv = 0.0;
    if (e < (long)n * m) {
        const int i = (int)(e % n), j = (int)(e / n);
        const float* d = bptr(D, p, 0) + 2 * (i + (long)j * D.ld);
        double* x = bptr(X, p, 0) + 2 * (i + (long)j * X.ld);
        const double dr = d[0], di = d[1];
        const double xr = x[0] + dr, xi = x[1] + di;
        x[0] = xr;
        x[1] = xi;
        dv = sqrt(dr * dr + di * di);
        xv = sqrt(xr * xr + xi * xi);
        if (dv != dv) dv = __longlong_as_double(0x7ff0000000000000LL);
        if (xv != xv) xv = __longlong_as_double(0x7ff0000000000000LL);
    }
    for (int off = 32; off > 0; off >>= 1) {
        dv = fmax(dv, __shfl_xor(dv, off));
        xv = fmax(xv, __shfl_xor(xv, off));
    }
    if ((threadIdx.x & 63) == 0) {
        amax_nonneg(&dmax[p], dv);
        amax_nonneg(&xmax[p], xv);
    }
}

// Refinement bookkeeping after step `step` (1 or 2).  delta = dmax / xmax.
// Step 1: record delta_1; active = (delta_1 > MIXED_SECOND_STEP); points that
// stay inactive are accepted.  Step 2: accept when delta_2^2 / delta_1 <=
// MIXED_ACCEPT and delta_2 < delta_1.  A failed FP32 factorisation (st32) is
// never accepted.  Counters: [0] points, [1] second steps, [2] rejected.
__global__ void k_mixed_decide(int step, int P, unsigned long long* dmax, unsigned long long* xmax, double* delta1,
                               const int* st32, int* active, int* accept, int* counters)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (step == 2 && !active[p]) return;
    const double dm = __longlong_as_double((long long)dmax[p]);
    const double xm = __longlong_as_double((long long)xmax[p]);
    double d = xm > 0.0 ? dm / xm : (dm > 0.0 ? __longlong_as_double(0x7ff0000000000000LL) : 0.0);
    if (d != d) d = __longlong_as_double(0x7ff0000000000000LL);
    dmax[p] = 0ull;
    xmax[p] = 0ull;
    const bool bad = st32[p] != 0;
    if (step == 1) {
        delta1[p] = d;
        const bool again = !bad && d > MIXED_SECOND_STEP && d < 1.0;
        active[p] = again ? 1 : 0;
        accept[p] = (!bad && d <= MIXED_SECOND_STEP) ? 1 : 0;
        atomicAdd(&counters[0], 1);
        if (again) atomicAdd(&counters[1], 1);
        if (!again && !accept[p]) atomicAdd(&counters[2], 1);
    } else {
        const double d1 = delta1[p];
        const bool ok = !bad && d < d1 && d * d <= MIXED_ACCEPT * d1;
        accept[p] = ok ? 1 : 0;
        if (!ok) atomicAdd(&counters[2], 1);
    }
}

__global__ void k_and_int(const int* a, int* b, int P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) b[p] = (a[p] && b[p]) ? 1 : 0;
}

// One sweep direction's chain: sizes n_s (s = 0..S), Dm_s, couplings op(C_s).
template <typename T>
struct Chain {
    std::vector<int> n;
    std::vector<BMat<T>> Dm;     // S + 1
    std::vector<BMat<T>> C;      // S
    std::vector<int> opC;        // op giving C_s (n_s x n_s+1) from C[s]
};

inline int flip(int op) { return op == OP_N ? OP_C : OP_N; }

struct DirBufs {
    float *G, *Pm, *Y, *W, *Tt, *Dm0, *DmS;   // FP32 slots
    double *X, *R, *Dm0d, *DmSd, *T1;         // FP64 slots
    unsigned long long *dmax, *xmax;
    double* delta1;
    int *act, *st32;
};

// slots of one direction: FP32 3S + 6, FP64 S + 5; per-point scalars
DirBufs carve(Arena& a, int S, long mm, int P)
{
    DirBufs d;
    const size_t f = 8 * (size_t)mm * P, g = 16 * (size_t)mm * P;
    d.G = a.ptr<float>(a.take(f * (S + 1)));
    d.Pm = a.ptr<float>(a.take(f * std::max(S, 1)));
    d.Y = a.ptr<float>(a.take(f * (S + 1)));
    d.W = a.ptr<float>(a.take(f));
    d.Tt = a.ptr<float>(a.take(f));
    d.Dm0 = a.ptr<float>(a.take(f));
    d.DmS = a.ptr<float>(a.take(f));
    d.X = a.ptr<double>(a.take(g * (S + 1)));
    d.R = a.ptr<double>(a.take(g));
    d.Dm0d = a.ptr<double>(a.take(g));
    d.DmSd = a.ptr<double>(a.take(g));
    d.T1 = a.ptr<double>(a.take(g));
    d.dmax = a.ptr<unsigned long long>(a.take(8 * (size_t)P));
    d.xmax = a.ptr<unsigned long long>(a.take(8 * (size_t)P));
    d.delta1 = a.ptr<double>(a.take(8 * (size_t)P));
    d.act = a.ptr<int>(a.take(4 * (size_t)P));
    d.st32 = a.ptr<int>(a.take(4 * (size_t)P));
    return d;
}

// Factor, initial column, refinement of one direction; accept[p] set, the
// refined column in d.X (slot s of point p at X + 2 (s P + p) mm).
void refine_direction(nt_ctx* c, const Chain<double>& c64, const Chain<float>& c32, const double* z, const float* z32,
                      int P, long mm, const DirBufs& d, int* accept, int* counters)
{
    const int S = (int)c64.n.size() - 1;
    const int m = c64.n[S];
    hipStream_t s = c->stream;
    auto F = [&](float* base, int sl, int ld) { return bw<float>(base + 2 * (long)sl * mm * P, mm, ld); };
    auto Fr = [&](float* base, int sl, int ld) { return bm<float>(base + 2 * (long)sl * mm * P, mm, ld); };
    auto D = [&](double* base, int sl, int ld) { return bw<double>(base + 2 * (long)sl * mm * P, mm, ld); };
    auto Dr = [&](double* base, int sl, int ld) { return bm<double>(base + 2 * (long)sl * mm * P, mm, ld); };
    const auto& n = c64.n;
    zero_dev(c, d.st32, sizeof(int) * P, s);
    zero_dev(c, d.dmax, sizeof(unsigned long long) * P, s);
    zero_dev(c, d.xmax, sizeof(unsigned long long) * P, s);

    // ---- 1. FP32 factorisation
    for (int l = 0; l <= S; ++l) {
        const int nl_ = n[l];
        if (l == 0) {
            shift_sub<float>(c, nl_, z32, c32.Dm[0], none<float>(), F(d.W, 0, nl_), P);
        } else {
            // W = z - Dm_l - C_l-1^+ P_l-1
            gemm<float>(c, nl_, nl_, n[l - 1], flip(c32.opC[l - 1]), c32.C[l - 1], OP_N, Fr(d.Pm, l - 1, n[l - 1]),
                        F(d.W, 0, nl_), -1.0, c32.Dm[l], -1.0, z32, P);
        }
        inverse<float>(c, nl_, F(d.W, 0, nl_), none<float>(), nullptr, F(d.G, l, nl_), d.st32, P);
        if (l < S)
            gemm<float>(c, nl_, n[l + 1], nl_, OP_N, Fr(d.G, l, nl_), c32.opC[l], c32.C[l], F(d.Pm, l, nl_), 1.0,
                        none<float>(), 0.0, nullptr, P);
    }
    // ---- 2. initial column in FP32, stored in FP64
    convert_op<float, double>(c, n[S], m, Fr(d.G, S, n[S]), D(d.X, S, n[S]), 0, 1.0, P);
    for (int l = S - 1; l >= 0; --l) {
        const float* nxt = l + 1 == S ? d.G + 2 * (long)S * mm * P : d.Y + 2 * (long)(l + 1) * mm * P;
        gemm<float>(c, n[l], m, n[l + 1], OP_N, Fr(d.Pm, l, n[l]), OP_N, bm<float>(nxt, mm, n[l + 1]),
                    F(d.Y, l, n[l]), 1.0, none<float>(), 0.0, nullptr, P);
        convert_op<float, double>(c, n[l], m, Fr(d.Y, l, n[l]), D(d.X, l, n[l]), 0, 1.0, P);
    }
    // ---- 3./4. refinement steps
    for (int step = 1; step <= 2; ++step) {
        // the second step runs masked (points with active = 0 leave every
        // launch at once), so no host synchronisation is needed; its flops
        // are not counted in the profile
        const int* act = step == 1 ? nullptr : d.act;
        const double ni = step == 1 ? -1.0 : 0.0;
        for (int l = 0; l <= S; ++l) {
            const int nl_ = n[l];
            // R = Dm_l X_l + C_l X_l+1 + C_l-1^+ X_l-1
            gemm<double>(c, nl_, m, nl_, OP_N, c64.Dm[l], OP_N, Dr(d.X, l, nl_), D(d.R, 0, nl_), 1.0,
                         none<double>(), 0.0, nullptr, P, ni, act);
            if (l < S)
                gemm<double>(c, nl_, m, n[l + 1], c64.opC[l], c64.C[l], OP_N, Dr(d.X, l + 1, n[l + 1]),
                             D(d.R, 0, nl_), 1.0, Dr(d.R, 0, nl_), 1.0, nullptr, P, ni, act);
            if (l > 0)
                gemm<double>(c, nl_, m, n[l - 1], flip(c64.opC[l - 1]), c64.C[l - 1], OP_N,
                             Dr(d.X, l - 1, n[l - 1]), D(d.R, 0, nl_), 1.0, Dr(d.R, 0, nl_), 1.0, nullptr, P, ni,
                             act);
            // r = e_S - A X in FP32
            audit<double>(c, Dr(d.R, 0, nl_), nl_, m, P, 1, "resid R");
            audit<double>(c, Dr(d.X, l, nl_), nl_, m, P, 1, "resid X");
            audit<float>(c, F(d.Tt, 0, nl_), nl_, m, P, 1, "resid out");
            check_range(c, z, sizeof(double) * 2 * P, "resid z");
            launch(c, NT_K_ELEM, 0, "resid", [&] {
                hipLaunchKernelGGL(k_resid, dim3(nb1((long)nl_ * m), P), dim3(NTE), 0, s, nl_, m, Dr(d.R, 0, nl_),
                                   Dr(d.X, l, nl_), z, l == S ? 1 : 0, F(d.Tt, 0, nl_), act);
                return hipGetLastError();
            });
            // forward: y_l = G_l (r_l + C_l-1^+ y_l-1)
            if (l > 0)
                gemm<float>(c, nl_, m, n[l - 1], flip(c32.opC[l - 1]), c32.C[l - 1], OP_N, Fr(d.Y, l - 1, n[l - 1]),
                            F(d.Tt, 0, nl_), 1.0, Fr(d.Tt, 0, nl_), 1.0, nullptr, P, ni, act);
            gemm<float>(c, nl_, m, nl_, OP_N, Fr(d.G, l, nl_), OP_N, Fr(d.Tt, 0, nl_), F(d.Y, l, nl_), 1.0,
                        none<float>(), 0.0, nullptr, P, ni, act);
        }
        // backward: dX_l = y_l + P_l dX_l+1 (in place), X += dX
        for (int l = S; l >= 0; --l) {
            if (l < S)
                gemm<float>(c, n[l], m, n[l + 1], OP_N, Fr(d.Pm, l, n[l]), OP_N, Fr(d.Y, l + 1, n[l + 1]),
                            F(d.Y, l, n[l]), 1.0, Fr(d.Y, l, n[l]), 1.0, nullptr, P, ni, act);
            audit<float>(c, Fr(d.Y, l, n[l]), n[l], m, P, 1, "xupdate dX");
            audit<double>(c, D(d.X, l, n[l]), n[l], m, P, 1, "xupdate X");
            launch(c, NT_K_ELEM, 0, "xupdate", [&] {
                hipLaunchKernelGGL(k_xupdate, dim3(nb1((long)n[l] * m), P), dim3(NTE), 0, s, n[l], m,
                                   Fr(d.Y, l, n[l]), D(d.X, l, n[l]), act, d.dmax, d.xmax);
                return hipGetLastError();
            });
        }
        check_range(c, d.dmax, sizeof(unsigned long long) * P, "decide dmax");
        check_range(c, d.xmax, sizeof(unsigned long long) * P, "decide xmax");
        check_range(c, d.delta1, sizeof(double) * P, "decide delta1");
        check_range(c, d.act, sizeof(int) * P, "decide active");
        check_range(c, d.st32, sizeof(int) * P, "decide st32");
        check_range(c, accept, sizeof(int) * P, "decide accept");
        check_range(c, counters, sizeof(int) * 3, "decide counters");
        hipLaunchKernelGGL(k_mixed_decide, dim3(nb1(P)), dim3(NTE), 0, s, step, P, d.dmax, d.xmax, d.delta1,
                           d.st32, d.act, accept, counters);
        hipcheck(c, hipGetLastError(), "mixed decide");
    }
}

// every buffer of mixed_caroli, carved from off (a dry run gives the size)
struct MixedBufs {
    DirBufs dl, dr;
    double *go, *tx, *ty, *go2;
    int* acc2;
};

MixedBufs carve_all(Arena& sa, int L, long mm, int P)
{
    MixedBufs b;
    b.dl = carve(sa, L, mm, P);
    b.dr = carve(sa, L, mm, P);
    b.go = sa.ptr<double>(sa.take(16 * (size_t)mm * P));    // G_LR when not an output
    b.tx = sa.ptr<double>(sa.take(16 * (size_t)mm * P));
    b.ty = sa.ptr<double>(sa.take(16 * (size_t)mm * P));
    b.go2 = sa.ptr<double>(sa.take(16 * (size_t)mm * P));   // G_RL when not an output
    b.acc2 = sa.ptr<int>(sa.take(4 * (size_t)P));
    return b;
}

}  // namespace

namespace nth {

size_t mixed_caroli_scratch(int nl, int nr, int b, int L, int P)
{
    const int bmax = std::max(std::max(nl, nr), b);
    Arena a;
    (void)carve_all(a, L, (long)bmax * bmax, P);
    return a.off;
}

// Mixed-precision T_LR / T_RL of P points (see the file comment).  in64 /
// in32 describe the same system in FP64 and FP32 (lead quantities retarded,
// per point); zd / zd32 the device shifts.  accept[p] (device) = 1 for points
// whose refined column passed the error estimate, 0 for points the caller
// must recompute in FP64 (their outputs here are not to be used).
// counters (device, 6 ints): per direction {points, second steps, rejected}.
void mixed_caroli(nt_ctx* c, int nl, int nr, const RgfInput<double>& in, const RgfInput<float>& i32, const double* zd,
                  const float* zd32, int P, RgfOutput<double> out, Arena& ar, size_t off, int* accept,
                  int* counters)
{
    const int L = in.L;
    int bmax = std::max(nl, nr);
    for (int b : in.bsz) bmax = std::max(bmax, b);
    const long mm = (long)bmax * bmax;
    Arena sa = ar;
    sa.off = off;
    const MixedBufs mb = carve_all(sa, L, mm, P);
    const DirBufs& dl = mb.dl;
    const DirBufs& dr = mb.dr;
    double *go = mb.go, *tx = mb.tx, *ty = mb.ty, *go2 = mb.go2;
    int* acc2 = mb.acc2;
    auto Wd = [&](double* p, int ld) { return bw<double>(p, mm, ld); };
    auto Rd = [&](double* p, int ld) { return bm<double>(p, mm, ld); };
    auto Wf = [&](float* p, int ld) { return bw<float>(p, mm, ld); };
    auto Rf = [&](float* p, int ld) { return bm<float>(p, mm, ld); };
    const int b0 = in.bsz[0], bN = in.bsz[L - 1];

    // ---- chains.  Left to right: Dm_0 = Hd_0 + Hld^+ gl Hld, couplings Hdv_i,
    // closing coupling Vc = Hrd_fwd or Hrd^+, Dm_L = Hamr + sr.
    Chain<double> f64;
    Chain<float> f32;
    {
        gemm<double>(c, b0, nl, nl, OP_C, in.Hld, OP_N, in.gl, Wd(dl.T1, b0), 1.0, none<double>(), 0.0, nullptr, P);
        gemm<double>(c, b0, b0, nl, OP_N, Rd(dl.T1, b0), OP_N, in.Hld, Wd(dl.Dm0d, b0), 1.0, in.Hd[0], 1.0, nullptr,
                     P);
        audit<double>(c, Wd(dl.DmSd, nr), nr, nr, P, 1, "DmS");
        audit<double>(c, in.Hamr, nr, nr, P, 1, "Hamr");
        audit<double>(c, in.sr, nr, nr, P, 1, "sr");
        launch(c, NT_K_ELEM, 0, "add2", [&] {
            hipLaunchKernelGGL(k_add2, dim3(nb1((long)nr * nr), P), dim3(NTE), 0, c->stream, nr, in.Hamr, in.sr,
                               Wd(dl.DmSd, nr));
            return hipGetLastError();
        });
        convert_op<double, float>(c, b0, b0, Rd(dl.Dm0d, b0), Wf(dl.Dm0, b0), 0, 1.0f, P);
        convert_op<double, float>(c, nr, nr, Rd(dl.DmSd, nr), Wf(dl.DmS, nr), 0, 1.0f, P);
        const bool fwd = in.Hrd_fwd.base != nullptr;
        if (fwd != (i32.Hrd_fwd.base != nullptr)) {
            c->err = "mixed solve: FP32 and FP64 closing couplings differ";
            throw Fail{NT_ERR_ARG};
        }
        for (int i = 0; i < L; ++i) {
            f64.n.push_back(in.bsz[i]);
            f32.n.push_back(in.bsz[i]);
            f64.Dm.push_back(i == 0 ? Rd(dl.Dm0d, b0) : in.Hd[i]);
            f32.Dm.push_back(i == 0 ? Rf(dl.Dm0, b0) : i32.Hd[i]);
            if (i + 1 < L) {
                f64.C.push_back(in.Hdv[i]);
                f32.C.push_back(i32.Hdv[i]);
                f64.opC.push_back(OP_N);
                f32.opC.push_back(OP_N);
            }
        }
        f64.C.push_back(fwd ? in.Hrd_fwd : in.Hrd);
        f32.C.push_back(fwd ? i32.Hrd_fwd : i32.Hrd);
        f64.opC.push_back(fwd ? OP_N : OP_C);
        f32.opC.push_back(fwd ? OP_N : OP_C);
        f64.n.push_back(nr);
        f32.n.push_back(nr);
        f64.Dm.push_back(Rd(dl.DmSd, nr));
        f32.Dm.push_back(Rf(dl.DmS, nr));
    }
    // Right to left: Dm_0 = Hd_L-1 + Hrd^+ gr Hrd, couplings V_i,i-1 =
    // Hdv_back_i-1 or Hdv_i-1^+, closing coupling Hld^+, Dm_L = Haml + sl.
    Chain<double> b64;
    Chain<float> b32;
    {
        gemm<double>(c, bN, nr, nr, OP_C, in.Hrd, OP_N, in.gr, Wd(dr.T1, bN), 1.0, none<double>(), 0.0, nullptr, P);
        gemm<double>(c, bN, bN, nr, OP_N, Rd(dr.T1, bN), OP_N, in.Hrd, Wd(dr.Dm0d, bN), 1.0, in.Hd[L - 1], 1.0,
                     nullptr, P);
        audit<double>(c, Wd(dr.DmSd, nl), nl, nl, P, 1, "DmS");
        audit<double>(c, in.Haml, nl, nl, P, 1, "Haml");
        audit<double>(c, in.sl, nl, nl, P, 1, "sl");
        launch(c, NT_K_ELEM, 0, "add2", [&] {
            hipLaunchKernelGGL(k_add2, dim3(nb1((long)nl * nl), P), dim3(NTE), 0, c->stream, nl, in.Haml, in.sl,
                               Wd(dr.DmSd, nl));
            return hipGetLastError();
        });
        convert_op<double, float>(c, bN, bN, Rd(dr.Dm0d, bN), Wf(dr.Dm0, bN), 0, 1.0f, P);
        convert_op<double, float>(c, nl, nl, Rd(dr.DmSd, nl), Wf(dr.DmS, nl), 0, 1.0f, P);
        for (int i = L - 1; i >= 0; --i) {
            b64.n.push_back(in.bsz[i]);
            b32.n.push_back(in.bsz[i]);
            b64.Dm.push_back(i == L - 1 ? Rd(dr.Dm0d, bN) : in.Hd[i]);
            b32.Dm.push_back(i == L - 1 ? Rf(dr.Dm0, bN) : i32.Hd[i]);
            if (i >= 1) {
                const bool back = in.Hdv_back.size() > (size_t)(i - 1) && in.Hdv_back[i - 1].base;
                if (back != (i32.Hdv_back.size() > (size_t)(i - 1) && i32.Hdv_back[i - 1].base != nullptr)) {
                    c->err = "mixed solve: FP32 and FP64 couplings differ";
                    throw Fail{NT_ERR_ARG};
                }
                b64.C.push_back(back ? in.Hdv_back[i - 1] : in.Hdv[i - 1]);
                b32.C.push_back(back ? i32.Hdv_back[i - 1] : i32.Hdv[i - 1]);
                b64.opC.push_back(back ? OP_N : OP_C);
                b32.opC.push_back(back ? OP_N : OP_C);
            }
        }
        b64.C.push_back(in.Hld);
        b32.C.push_back(i32.Hld);
        b64.opC.push_back(OP_C);
        b32.opC.push_back(OP_C);
        b64.n.push_back(nl);
        b32.n.push_back(nl);
        b64.Dm.push_back(Rd(dr.DmSd, nl));
        b32.Dm.push_back(Rf(dr.DmS, nl));
    }
    zero_dev(c, counters, sizeof(int) * 6, c->stream);

    // ---- both directions, on two streams when enabled
    hipStream_t main_stream = c->stream;
    const bool two = c->overlap && c->aux;
    if (two) {
        hipcheck(c, hipEventRecord(c->fork, main_stream), "event fork");
        hipcheck(c, hipStreamWaitEvent(c->aux, c->fork, 0), "wait fork");
    }
    refine_direction(c, f64, f32, zd, zd32, P, mm, dl, accept, counters);
    // G_LR = gl Hld X_0, T_LR = Re Tr[(G_LR gmr)(G_LR^+ gml)], G_R = X_N+1
    BMatW<double> gdlr = out.G_D_LR.base ? out.G_D_LR : Wd(go, nl);
    gemm<double>(c, nl, nr, b0, OP_N, in.Hld, OP_N, bm<double>(dl.X, mm, b0), Wd(dl.T1, nl), 1.0, none<double>(), 0.0,
                 nullptr, P);
    gemm<double>(c, nl, nr, nl, OP_N, in.gl, OP_N, Rd(dl.T1, nl), gdlr, 1.0, none<double>(), 0.0, nullptr, P);
    gemm<double>(c, nl, nr, nr, OP_N, ro(gdlr), OP_N, in.gmr, Wd(tx, nl), 1.0, none<double>(), 0.0, nullptr, P);
    gemm<double>(c, nr, nl, nl, OP_C, ro(gdlr), OP_N, in.gml, Wd(ty, nr), 1.0, none<double>(), 0.0, nullptr, P);
    trace_prod<double>(c, nl, nr, Rd(tx, nl), Rd(ty, nr), out.T_LR, P);
    if (out.G_R.base)
        copy_op<double>(c, nr, nr, bm<double>(dl.X + 2 * (long)L * mm * P, mm, nr), out.G_R, 0, 1.0, P);

    if (two) {
        c->stream = c->aux;
    }
    refine_direction(c, b64, b32, zd, zd32, P, mm, dr, acc2, counters + 3);
    BMatW<double> gdrl = out.G_D_RL.base ? out.G_D_RL : Wd(go2, nr);
    double* tx2 = dr.R;            // free after the refinement
    double* ty2 = dr.T1;
    gemm<double>(c, nr, nl, bN, OP_N, in.Hrd, OP_N, bm<double>(dr.X, mm, bN), Wd(dr.Dm0d, nr), 1.0, none<double>(),
                 0.0, nullptr, P);
    gemm<double>(c, nr, nl, nr, OP_N, in.gr, OP_N, Rd(dr.Dm0d, nr), gdrl, 1.0, none<double>(), 0.0, nullptr, P);
    gemm<double>(c, nr, nl, nl, OP_N, ro(gdrl), OP_N, in.gml, Wd(tx2, nr), 1.0, none<double>(), 0.0, nullptr, P);
    gemm<double>(c, nl, nr, nr, OP_C, ro(gdrl), OP_N, in.gmr, Wd(ty2, nl), 1.0, none<double>(), 0.0, nullptr, P);
    trace_prod<double>(c, nr, nl, Rd(tx2, nr), Rd(ty2, nl), out.T_RL, P);
    if (out.G_L.base)
        copy_op<double>(c, nl, nl, bm<double>(dr.X + 2 * (long)L * mm * P, mm, nl), out.G_L, 0, 1.0, P);
    if (two) {
        hipcheck(c, hipEventRecord(c->join, c->aux), "event join");
        c->stream = main_stream;
        hipcheck(c, hipStreamWaitEvent(main_stream, c->join, 0), "wait join");
    }
    // a point is accepted when both directions are
    check_range(c, acc2, sizeof(int) * P, "accept 2");
    check_range(c, accept, sizeof(int) * P, "accept");
    launch(c, NT_K_ELEM, 0, "accept", [&] {
        hipLaunchKernelGGL(k_and_int, dim3(nb1(P)), dim3(NTE), 0, c->stream, acc2, accept, P);
        return hipGetLastError();
    });
}

}  // namespace nth
