// nt_host.h -- host-side machinery shared by the engine's translation units:
// the context, error handling, launch bookkeeping (profiling events), the
// workspace arena, batched-matrix descriptors, the bounds audit every launch
// runs on the host, and the typed wrappers of the GEMM, inverse and
// elementwise kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/negf_transport.h"
#include "nt_kernels.h"

using namespace nt;

// smallest part of a batched solve split over worker contexts
constexpr int NT_SPLIT_MIN_POINTS = 256;

struct nt_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t aux = nullptr;     // second stream: the right-to-left sweep runs beside the left-to-right one
    hipEvent_t fork = nullptr, join = nullptr;
    bool overlap = true;
    bool gemm_m3 = false;          // FP64 GEMMs with three real products per complex product
    hipStream_t stream = nullptr;
    std::string err;
    int batch_limit = 512;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    size_t ws_cap = 0;             // > 0: chunks of points shrink until their workspace fits
    bool ws_poison = false;        // fill the workspace with NaN bit patterns before each use
    int* host_flag = nullptr;      // pinned, two slots (the decimation's pipelined checks)
    hipEvent_t chk[2] = {nullptr, nullptr};
    // profiling
    bool prof = false;
    struct Rec {
        hipEvent_t a, b;
        int kind;
        double flops;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    long long launches[NT_K_COUNT] = {};
    double ms[NT_K_COUNT] = {};
    double flops[NT_K_COUNT] = {};
    // flops known only after the launches (decimation, from the per-item
    // iteration counts), added to flops[] when the profile is harvested
    double flops_pending[NT_K_COUNT] = {};
    double last_mean_iters = 0.0;
    // decimation: the host reads the number of unconverged items every
    // dec_check iterations and stops at zero; 0 = never (max_iter iterations,
    // converged items masked: no host synchronisation inside a solve)
    int dec_check = 4;
    // mixed precision since the last profile reset: points, second refinement
    // steps (both directions), points recomputed in FP64
    long long mixed_counts[3] = {0, 0, 0};
    // adjoint: keep G^L of every layer (0), every n-th (n > 0) or about
    // sqrt(L - lo)-th (< 0) from the window to the end; the rest recomputed
    int adj_ckpt = 0;
    // T_RL (and G_D_RL, G_L) from the left-to-right sweep and a corner
    // correction instead of the right-to-left sweep
    bool one_sweep = true;
    // device address ranges a launch may touch: the workspace and the caller's
    // buffers of the current call (sizes derived from the call's arguments)
    std::vector<std::pair<uintptr_t, uintptr_t>> regions;
    // unitary block DFT U = P1 (1 (x) P2) of circular decimation, per lead size
    struct CircU {
        int n, nbx, nby;
        double* d;
    };
    std::vector<CircU> circ_u;
    // autotuned inverse panel width per (element bytes, n); absent: size rule
    std::map<std::pair<int, int>, int> inv_nb;
    // autotuned GEMM output tile (64 or 32) per (element bytes, max(M, N)); absent: 64
    std::map<std::pair<int, int>, int> gemm_tile;
    // degenerate-group descriptors and statuses of the path_mode = 1 rotation
    void* deg_buf = nullptr;
    size_t deg_bytes = 0;
    // uneven-device calls: worker contexts (own stream and workspace, one host
    // thread each) that solve groups of devices concurrently
    int multi_workers = 4;
    // batched solves of at least point_split * NT_SPLIT_MIN_POINTS points
    // run as point_split parts side by side on worker contexts (1: one stream)
    int point_split = 2;
    std::vector<nt_ctx*> subs;
};

namespace nth {

// ------------------------------------------------------------- helpers
struct Fail {
    int code;
};

inline void hipcheck(nt_ctx* c, hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        c->err = std::string(what) + ": " + hipGetErrorString(e);
        throw Fail{e == hipErrorOutOfMemory ? NT_ERR_NOMEM : NT_ERR_HIP};
    }
}

inline hipEvent_t take_event(nt_ctx* c)
{
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hipcheck(c, hipEventCreate(&e), "hipEventCreate");
    return e;
}

template <class F>
void launch(nt_ctx* c, int kind, double flops, const char* what, F&& f)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (c->prof) {
        a = take_event(c);
        hipcheck(c, hipEventRecord(a, c->stream), "hipEventRecord");
    }
    hipcheck(c, f(), what);
    if (c->prof) {
        b = take_event(c);
        hipcheck(c, hipEventRecord(b, c->stream), "hipEventRecord");
        c->recs.push_back({a, b, kind, flops});
    }
}

inline void harvest_profile(nt_ctx* c)
{
    if (!c->recs.empty()) hipcheck(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (auto& r : c->recs) {
        float ms = 0.f;
        hipcheck(c, hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
        c->launches[r.kind] += 1;
        c->ms[r.kind] += ms;
        c->flops[r.kind] += r.flops;
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->recs.clear();
    for (int k = 0; k < NT_K_COUNT; ++k) {
        c->flops[k] += c->flops_pending[k];
        c->flops_pending[k] = 0.0;
    }
}

// Zero a device buffer of 4-byte words with a fill kernel (a memset node in a
// captured graph is not reliably replayed; a kernel node is).
inline void zero_dev(nt_ctx* c, void* p, size_t bytes, hipStream_t s)
{
    if (!p || !bytes) return;
    if (bytes % 4 || reinterpret_cast<uintptr_t>(p) % 4) {
        c->err = "zero_dev: buffer not made of 4-byte words";
        throw Fail{NT_ERR_ARG};
    }
    hipcheck(c, launch_fill_int(static_cast<int*>(p), 0, (int)(bytes / 4), s), "fill");
}

// Bump allocator over the context workspace.
struct Arena {
    size_t off = 0;
    char* base = nullptr;
    size_t take(size_t bytes)
    {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    }
    template <typename P>
    P* ptr(size_t o) const
    {
        return reinterpret_cast<P*>(base + o);
    }
};



template <typename T>
BMat<T> bm(const T* base, long stride, int ld, const int* idx = nullptr, long gstride = 0, int idx_n = 0)
{
    return BMat<T>{base, stride, gstride, idx, ld, idx_n};
}
template <typename T>
BMatW<T> bw(T* base, long stride, int ld, const int* idx = nullptr, long gstride = 0, int idx_n = 0)
{
    return BMatW<T>{base, stride, gstride, idx, ld, idx_n};
}
template <typename T>
BMat<T> ro(const BMatW<T>& w)
{
    return BMat<T>{w.base, w.stride, w.gstride, w.idx, w.ld, w.idx_n};
}
template <typename T>
BMat<T> none()
{
    return BMat<T>{nullptr, 0, 0, nullptr, 0, 0};
}
template <typename T>
BMatW<T> nonew()
{
    return BMatW<T>{nullptr, 0, 0, nullptr, 0, 0};
}

// ---------------------------------------------------------- bounds audit
// Every launch checks, on the host, that each operand's whole address range
// (all items, groups, rows and columns) lies inside one registered region.
// An addressing mistake then becomes NT_ERR_ARG instead of a GPU fault.
inline void reg_region(nt_ctx* c, const void* p, size_t bytes)
{
    if (p && bytes) c->regions.push_back({(uintptr_t)p, (uintptr_t)p + bytes});
}

inline void unreg_region(nt_ctx* c, const void* p)
{
    for (size_t i = 0; i < c->regions.size(); ++i)
        if (c->regions[i].first == (uintptr_t)p) {
            c->regions.erase(c->regions.begin() + i);
            return;
        }
}

inline void check_range(nt_ctx* c, const void* p, size_t bytes, const char* what)
{
    if (!p || !bytes) return;
    const uintptr_t b = (uintptr_t)p, e = b + bytes;
    for (auto& r : c->regions)
        if (b >= r.first && e <= r.second) return;
    char buf[256];
    snprintf(buf, sizeof(buf), "bounds audit: %s touches [%p, +%zu) outside every registered buffer", what, p,
             bytes);
    c->err = buf;
    throw Fail{NT_ERR_ARG};
}

// all-ones words (a NaN in every double, -1 in every int) over [p, p + bytes)
inline void poison_dev(nt_ctx* c, void* p, size_t bytes, hipStream_t s)
{
    int* w = static_cast<int*>(p);
    for (size_t n = bytes / 4; n > 0;) {
        const int k = (int)std::min<size_t>(n, size_t(1) << 30);
        hipcheck(c, launch_fill_int(w, -1, k, s), "poison");
        w += k;
        n -= k;
    }
}

inline void ensure_ws(nt_ctx* c, size_t bytes)
{
    if (c->ws_bytes < bytes) {
        if (c->ws) {
            hipcheck(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            unreg_region(c, c->ws);
            hipcheck(c, hipFree(c->ws), "hipFree");
            c->ws = nullptr;
            c->ws_bytes = 0;
        }
        size_t want = bytes + bytes / 8;
        hipcheck(c, hipMalloc(&c->ws, want), "hipMalloc workspace");
        c->ws_bytes = want;
        reg_region(c, c->ws, c->ws_bytes);
    }
    if (c->ws_poison) poison_dev(c, c->ws, bytes, c->stream);
}

template <typename T, typename M>
void audit(nt_ctx* c, const M& m, int rows, int cols, int batch, int groups, const char* what)
{
    if (!m.base || rows <= 0 || cols <= 0 || batch <= 0) return;
    if (m.stride < 0 || m.gstride < 0 || m.ld < rows || (m.idx && m.idx_n <= 0)) {
        c->err = std::string("bounds audit: bad descriptor for ") + what;
        throw Fail{NT_ERR_ARG};
    }
    const long items = m.idx ? m.idx_n : batch;
    const long last = (items - 1) * m.stride + (long)(groups - 1) * m.gstride + (rows - 1) + (long)(cols - 1) * m.ld;
    check_range(c, m.base, (size_t)(last + 1) * 2 * sizeof(T), what);
}

// C = alpha op(A) op(B) + beta Cin + diag
template <typename T>
void gemm(nt_ctx* c, int M, int N, int K, int opA, BMat<T> A, int opB, BMat<T> B, BMatW<T> C, double alpha,
          BMat<T> Cin, double beta, const T* diag, int batch, double nitems = -1, const int* active = nullptr,
          int groups = 1, BMatW<T> C2 = BMatW<T>{nullptr, 0, 0, nullptr, 0},
          unsigned long long* maxabs = nullptr)
{
    GemmArgs<T> a;
    a.M = M;
    a.N = N;
    a.K = K;
    a.opA = opA;
    a.opB = opB;
    a.A = A;
    a.B = B;
    a.Cin = Cin;
    a.C = C;
    a.C2 = C2;
    a.alpha_re = (T)alpha;
    a.alpha_im = 0;
    a.beta_re = (T)beta;
    a.beta_im = 0;
    a.diag = diag;
    a.active = active;
    a.maxabs = maxabs;
    a.batch = batch;
    a.groups = groups;
    a.m3 = (sizeof(T) == 8 && c->gemm_m3) ? 1 : 0;
    if (!c->gemm_tile.empty()) {
        auto tuned = c->gemm_tile.find({(int)sizeof(T), std::max(M, N)});
        a.tile = tuned == c->gemm_tile.end() ? 0 : tuned->second;
    }
    const double items = (nitems < 0 ? (double)batch : nitems) * groups;
    audit<T>(c, A, opA == OP_N ? M : K, opA == OP_N ? K : M, batch, groups, "gemm A");
    audit<T>(c, B, opB == OP_N ? K : N, opB == OP_N ? N : K, batch, groups, "gemm B");
    audit<T>(c, C, M, N, batch, groups, "gemm C");
    audit<T>(c, Cin, M, N, batch, groups, "gemm Cin");
    audit<T>(c, C2, M, N, batch, groups, "gemm C2");
    check_range(c, diag, sizeof(T) * 2 * batch, "gemm diag");
    check_range(c, active, sizeof(int) * batch, "gemm active");
    check_range(c, maxabs, sizeof(unsigned long long) * batch, "gemm maxabs");
    launch(c, NT_K_GEMM, 8.0 * M * N * K * items, "gemm", [&] { return launch_gemm<T>(a, c->stream); });
}

template <typename T>
void inverse(nt_ctx* c, int n, BMatW<T> W, BMat<T> src, const T* z, BMatW<T> X, int* info, int batch,
             double nitems = -1, const int* active = nullptr)
{
    InvArgs<T> a;
    a.n = n;
    a.W = W;
    a.src = src;
    a.z = z;
    a.X = X;
    a.active = active;
    a.info = info;
    a.batch = batch;
    auto tuned = c->inv_nb.find({(int)sizeof(T), n});
    a.nb = tuned == c->inv_nb.end() ? 0 : tuned->second;
    const double items = nitems < 0 ? (double)batch : nitems;
    audit<T>(c, W, n, n, batch, 1, "inverse W");
    audit<T>(c, src, n, n, batch, 1, "inverse src");
    audit<T>(c, X, n, n, batch, 1, "inverse X");
    if (src.base) check_range(c, z, sizeof(T) * 2 * batch, "inverse z");
    check_range(c, info, sizeof(int) * batch, "inverse info");
    check_range(c, active, sizeof(int) * batch, "inverse active");
    if (W.base == X.base) {
        c->err = "inverse: work and output alias";
        throw Fail{NT_ERR_ARG};
    }
    launch(c, NT_K_INV, 8.0 * n * (double)n * n * items, "inverse", [&] { return launch_inv<T>(a, c->stream); });
}

template <typename T>
void copy_op(nt_ctx* c, int m, int n, BMat<T> in, BMatW<T> out, int dagger, T sign, int batch,
             const int* active = nullptr)
{
    audit<T>(c, in, m, n, batch, 1, "copy in");
    audit<T>(c, out, dagger ? n : m, dagger ? m : n, batch, 1, "copy out");
    check_range(c, active, sizeof(int) * batch, "copy active");
    launch(c, NT_K_ELEM, 0, "copy",
           [&] { return launch_copy_op<T>(m, n, in, out, dagger, sign, active, batch, c->stream); });
}

template <typename Ti, typename To>
void convert_op(nt_ctx* c, int m, int n, BMat<Ti> in, BMatW<To> out, int dagger, To sign, int batch)
{
    audit<Ti>(c, in, m, n, batch, 1, "convert in");
    audit<To>(c, out, dagger ? n : m, dagger ? m : n, batch, 1, "convert out");
    launch(c, NT_K_ELEM, 0, "convert",
           [&] { return launch_convert<Ti, To>(m, n, in, out, dagger, sign, batch, c->stream); });
}

template <typename T>
void gamma_op(nt_ctx* c, int n, BMat<T> S, BMatW<T> G, int batch)
{
    audit<T>(c, S, n, n, batch, 1, "gamma in");
    audit<T>(c, G, n, n, batch, 1, "gamma out");
    launch(c, NT_K_ELEM, 0, "gamma", [&] { return launch_gamma<T>(n, S, G, (T)1, nullptr, batch, c->stream); });
}

template <typename T>
void shift_sub(nt_ctx* c, int n, const T* z, BMat<T> A, BMat<T> B, BMatW<T> out, int batch,
               const int* active = nullptr)
{
    check_range(c, z, sizeof(T) * 2 * batch, "shift z");
    audit<T>(c, A, n, n, batch, 1, "shift A");
    audit<T>(c, B, n, n, batch, 1, "shift B");
    audit<T>(c, out, n, n, batch, 1, "shift out");
    check_range(c, active, sizeof(int) * batch, "shift active");
    launch(c, NT_K_ELEM, 0, "shift_sub",
           [&] { return launch_shift_sub<T>(n, z, A, B, out, active, batch, c->stream); });
}

template <typename T>
void trace_prod(nt_ctx* c, int n1, int n2, BMat<T> X, BMat<T> Y, double* out, int batch,
                const int* active = nullptr)
{
    audit<T>(c, X, n1, n2, batch, 1, "trace X");
    audit<T>(c, Y, n2, n1, batch, 1, "trace Y");
    check_range(c, out, sizeof(double) * batch, "trace out");
    check_range(c, active, sizeof(int) * batch, "trace active");
    launch(c, NT_K_ELEM, 0, "trace",
           [&] { return launch_trace_prod<T>(n1, n2, X, Y, out, 1, active, batch, c->stream); });
}

// ------------------------------------------------------------------ RGF
template <typename T>
struct RgfInput {
    int L;
    std::vector<int> bsz;
    std::vector<BMat<T>> Hd, Hdv, Hdv_back;   // Hdv_back[i].base may be null
    BMat<T> Hld, Hrd, Hrd_fwd, Haml, Hamr;     // Hrd_fwd.base may be null
    // retarded lead quantities (per point)
    BMat<T> gl, gr, gml, gmr, sl, sr;
    // Buttiker probes: per-layer device shifts z - sigma_w(i) (L x P complex,
    // layer-major; null: every device layer uses the common shift), and L slots
    // that receive the left-connected G of layers 0..L-2 and the full G of
    // layer L-1 (null: not kept)
    const T* zl = nullptr;
    T* keep = nullptr;
    // optional per-point mask: only points with act[p] != 0 are computed
    const int* act = nullptr;
    // retarded lead shift z_lead = E + i eta(E) of each point (complex at
    // zlead + zlead_stride p): with it the right-to-left result can come from
    // the left-to-right sweep (one_sweep_correction)
    const T* zlead = nullptr;
    int zlead_stride = 2;
};

template <typename T>
struct RgfOutput {
    double* T_LR;
    double* T_RL;
    int* status;
    BMatW<T> G_D_LR, G_D_RL, G_L, G_R;   // may be null (then scratch)
};

// Mixed-precision Caroli solve with FP64 iterative refinement (nt_mixed.hip).
size_t mixed_caroli_scratch(int nl, int nr, int b, int L, int P);
void mixed_caroli(nt_ctx* c, int nl, int nr, const RgfInput<double>& in, const RgfInput<float>& i32,
                  const double* zd, const float* zd32, int P, RgfOutput<double> out, Arena& ar, size_t off,
                  int* accept, int* counters);


// device address and size of the allocation-free solve's static scratch of
// this device (null before nt_create has looked it up on that device)
void* scratch_region(int device, size_t* bytes);

inline int check_params(const nt_params* p)
{
    if (!p || p->max_iter < 1 || !(p->convergence > 0.0) || (p->eta_mode != 0 && p->eta_mode != 1)) return NT_ERR_ARG;
    if (p->precision != NT_PREC_FP64 && p->precision != NT_PREC_MIXED) return NT_ERR_ARG;
    if (p->circular && (p->n_bloch_x < 1 || p->n_bloch_y < 1)) return NT_ERR_ARG;
    if (p->dos_mode < NT_DOS_FROM_ETA || p->dos_mode > NT_DOS_TB) return NT_ERR_ARG;
    return 0;
}

// circular decimation needs both lead sizes divisible by the number of copies
inline int check_circular(const nt_params* p, int nl, int nr)
{
    if (!p->circular) return 0;
    const int nb = p->n_bloch_x * p->n_bloch_y;
    if (nl % nb || nr % nb || (nl / p->n_bloch_y) % p->n_bloch_x) return NT_ERR_ARG;
    return 0;
}

inline unsigned nblk(long n, int t = 256) { return (unsigned)std::max<long>(1, (n + t - 1) / t); }

inline int first_status_host(const int* st, int n)
{
    for (int i = 0; i < n; ++i)
        if (st[i]) return st[i];
    return 0;
}

template <class F>
inline int guarded(nt_ctx* c, F&& f)
{
    if (!c) return NT_ERR_ARG;
    try {
        hipcheck(c, hipSetDevice(c->device), "hipSetDevice");
        c->regions.clear();
        reg_region(c, c->ws, c->ws_bytes);
        size_t sb = 0;
        if (void* sp = scratch_region(c->device, &sb)) reg_region(c, sp, sb);
        int r = f();
        if (c->prof) harvest_profile(c);
        return r;
    } catch (const Fail& e) {
        return e.code;
    } catch (const std::bad_alloc&) {
        c->err = "host allocation failed";
        return NT_ERR_NOMEM;
    }
}

// device buffer RAII for the host-pointer entry points
struct DevBuf {
    nt_ctx* c;
    void* p = nullptr;
    DevBuf(nt_ctx* c_, size_t bytes) : c(c_)
    {
        if (bytes) hipcheck(c, hipMalloc(&p, bytes), "hipMalloc");
        reg_region(c, p, bytes);
    }
    ~DevBuf()
    {
        if (p) {
            unreg_region(c, p);
            (void)hipFree(p);
        }
    }
    template <typename P>
    P* as()
    {
        return static_cast<P*>(p);
    }
};

inline void* up(nt_ctx* c, DevBuf& d, const void* h, size_t bytes)
{
    if (h && bytes) hipcheck(c, hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, c->stream), "H2D");
    return d.p;
}
inline void down(nt_ctx* c, void* h, const DevBuf& d, size_t bytes)
{
    if (h && bytes) hipcheck(c, hipMemcpyAsync(h, d.p, bytes, hipMemcpyDeviceToHost, c->stream), "D2H");
}

}  // namespace nth

using namespace nth;
