// nt_engine.hip -- orchestration of the batched (omega,k)-solve and the C ABI.
//
// One context owns a HIP stream and a device workspace.  A solve is a chain
// of batched launches over P points at once:
//   decimation   Lopez-Sancho iterations for all lead problems, FP64, with a
//                per-item convergence mask (negf/surface.f90:482-497)
//   lead post    g = (z - es)^-1, Sigma = V^+ g V, Gamma = i(Sigma - Sigma^+),
//                adv -> ret conversion (negf/lesser.f90:141-154)
//   RGF          left-to-right and right-to-left block sweeps
//                (negf/lesser.f90:555-784, negf/device.f90:330-525)
//   Caroli       T = Re Tr[G Gamma G^+ Gamma]  (negf/lesser.f90:787-807)
// Shared operands (H(k) blocks) are addressed with stride 0 or a k-index map,
// so one H(k) copy serves every energy of that k-point.
#include <atomic>
#include <mutex>
#include <thread>
#include <type_traits>

#include "nt_host.h"
#include "nt_solve.h"

// Static device scratch of the allocation-free batched solve
// (nt_solve_tdb_batched_noalloc_dev).  Each GPU the library runs on holds its
// own copy until the process ends.  Host code never uses the array's name as
// an address: nt_create looks up each device's copy with hipGetSymbolAddress
// and every use goes through that pointer (scratch_ptr).
constexpr size_t NT_SCRATCH_BYTES = 256ull << 20;
static __device__ __align__(256) char scratch[NT_SCRATCH_BYTES];

namespace {

constexpr int NT_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[NT_MAX_DEVICES] = {};

// device address of this context's device's scratch (set by nt_create)
void* scratch_ptr(nt_ctx* c)
{
    void* p = c->device >= 0 && c->device < NT_MAX_DEVICES ? g_scratch_dev[c->device] : nullptr;
    if (!p) {
        c->err = "static scratch of this device is not set up";
        throw Fail{NT_ERR_HIP};
    }
    return p;
}

}  // namespace

namespace nth {
void* scratch_region(int device, size_t* bytes)
{
    *bytes = NT_SCRATCH_BYTES;
    return device >= 0 && device < NT_MAX_DEVICES ? g_scratch_dev[device] : nullptr;
}
}  // namespace nth

namespace {

// ------------------------------------------------------- engine kernels
__global__ void k_lead_setup(const double* E, const int* kidx, int nk, int P, int two_sided, double eta, double eta0,
                             int eta_mode, double f, int* idx_dec, double* z_dec, double* z_pm, double* z_dev)
{
    // Combined lead batch: item q = 2p + side (side 0 = left, 1 = right).
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double e = E[p];
    const double et = eta_mode == NT_ETA_RELATIVE ? e * eta + eta0 : eta;
    const int k = kidx ? kidx[p] : 0;
    if (two_sided) {
        idx_dec[2 * p] = k;
        idx_dec[2 * p + 1] = nk + k;
    }
    // left advanced E - i eta, right retarded E + i eta (surface.f90:479, :540)
    z_dec[4 * p + 0] = e;
    z_dec[4 * p + 1] = -et;
    z_dec[4 * p + 2] = e;
    z_dec[4 * p + 3] = et;
    // left reflected E + i eta, right reflected E - i eta (surface.f90:510, :571)
    if (z_pm) {
        z_pm[4 * p + 0] = e;
        z_pm[4 * p + 1] = et;
        z_pm[4 * p + 2] = e;
        z_pm[4 * p + 3] = -et;
    }
    z_dev[2 * p] = e;
    z_dev[2 * p + 1] = f * et;
}

// Leads of different sizes (nl != nr) decimate in two batches, one per lead,
// with contiguous per-side shifts; the interleaved [left, right] shifts and
// k indices of k_lead_setup are split for them (side s of point p), and the
// per-side statuses and iteration counts merged back afterwards.
__global__ void k_lead_split(const double* z, const double* zpm, const int* idx2, int P, double* zs, double* zpms,
                             int* idx1)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    for (int side = 0; side < 2; ++side) {
        zs[2 * (side * P + p) + 0] = z[4 * p + 2 * side + 0];
        zs[2 * (side * P + p) + 1] = z[4 * p + 2 * side + 1];
        if (zpm) {
            zpms[2 * (side * P + p) + 0] = zpm[4 * p + 2 * side + 0];
            zpms[2 * (side * P + p) + 1] = zpm[4 * p + 2 * side + 1];
        }
    }
    idx1[p] = idx2[2 * p];
}

__global__ void k_lead_merge(const int* st_s, const int* it_s, int P, int* st2, int* it2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    for (int side = 0; side < 2; ++side) {
        st2[2 * p + side] = st_s[side * P + p];
        it2[2 * p + side] = it_s[side * P + p];
    }
}

// The blocks of the one-sweep corner system for block elimination:
//   M11 = I + d G_00 (nl x nl), M22 = I - d G_cc (nr x nr),
//   dGc0 = d G_c0 (nr x nl), dG0c = d G_0c (nl x nr), d = z_dev - z_lead.
// blockIdx.z selects the block, blockIdx.y the point.
__global__ void k_corner_blocks(int nl, int nr, BMat<double> G00, BMat<double> G0c, BMat<double> Gc0,
                                BMat<double> Gcc, const double* zd, const double* zl, int zls, BMatW<double> M11,
                                BMatW<double> M22, BMatW<double> dGc0, BMatW<double> dG0c, const int* active)
{
    const int p = blockIdx.y, which = blockIdx.z;
    if (active && !active[p]) return;
    const int rows = (which == 0 || which == 3) ? nl : nr;
    const int cols = (which == 0 || which == 2) ? nl : nr;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)rows * cols) return;
    const int i = (int)(e % rows), j = (int)(e / rows);
    const BMat<double>& G = which == 0 ? G00 : which == 1 ? Gcc : which == 2 ? Gc0 : G0c;
    const BMatW<double>& O = which == 0 ? M11 : which == 1 ? M22 : which == 2 ? dGc0 : dG0c;
    const double* g = bptr(G, p, 0) + 2 * (i + (long)j * G.ld);
    const double gr = g[0], gi = g[1];
    // d = i (Im z_dev - Im z_lead) (same real part E); M22 carries -d
    const double dr = zd[2 * p] - zl[(long)zls * p], di = zd[2 * p + 1] - zl[(long)zls * p + 1];
    const double sg = which == 1 ? -1.0 : 1.0;
    const double one = (which < 2 && i == j) ? 1.0 : 0.0;
    double* o = bptr(O, p, 0) + 2 * (i + (long)j * O.ld);
    o[0] = one + sg * (gr * dr - gi * di);
    o[1] = sg * (gr * di + gi * dr);
}

// fb[p] = !accept[p]; *count += number of rejected points
__global__ void k_reject_mask(const int* accept, int* fb, int* count, int P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    fb[p] = accept[p] ? 0 : 1;
    if (!accept[p]) atomicAdd(count, 1);
}

__global__ void k_dec_finish(const int* active, int* status, int Q)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < Q && active[q] && status[q] == 0) status[q] = NT_ERR_NOT_CONVERGED;
}

// point status = first nonzero of (dec left, dec right, rgf); iters = mean
__global__ void k_point_status(const int* st_dec, const int* it_dec, const int* st_rgf, int P, int* status, int* iters,
                               const int* first_mask, double* T_LR, double* T_RL)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int s = st_dec[2 * p];
    if (!s) s = st_dec[2 * p + 1];
    if (!s) s = st_rgf[p];
    if (status) status[p] = s;
    if (iters) iters[p] = (it_dec[2 * p] + it_dec[2 * p + 1] + 1) / 2;
    double tl = T_LR[p], tr = T_RL[p];
    if (tl < 1.0e-60) tl = 0.0;       // negf/lesser.f90:795-800
    if (tr < 1.0e-60) tr = 0.0;
    if (first_mask && first_mask[p]) {   // negf/lesser.f90:802-807
        tl = 0.0;
        tr = 0.0;
    }
    T_LR[p] = tl;
    T_RL[p] = tr;
}

__global__ void k_copy_lead_stack(const double2* a, const double2* b, double2* out, long n_a)
{
    // out = [a (n_a elems), b (n_a elems)]
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_a) {
        out[i] = a[i];
        out[n_a + i] = b[i];
    }
}


// ------------------------------------------------------------ decimation
// Q items of size n.  H00 / V addressed per item; alpha = V^+ (alpha_dagger)
// or V.  Outputs g (and optionally Sigma = alpha0 g beta0, Gamma).
template <typename T>
struct DecimationOut {
    BMatW<T> g, sigma, gamma;
};

template <typename T>
void decimate(nt_ctx* c, int n, int Q, BMat<T> H00, BMat<T> V, int alpha_dagger, const T* z, double conv,
              int max_iter, DecimationOut<T> out, int* status, int* iters, Arena& ar, size_t scratch_off)
{
    const long nn = (long)n * n;
    const size_t mb = sizeof(T) * 2 * nn;
    Arena sa = ar;
    sa.off = scratch_off;
    T* e1 = sa.ptr<T>(sa.take(mb * Q));
    T* es = sa.ptr<T>(sa.take(mb * Q));
    T* ab0 = sa.ptr<T>(sa.take(2 * mb * Q));
    T* ab1 = sa.ptr<T>(sa.take(2 * mb * Q));
    T* ive = sa.ptr<T>(sa.take(mb * Q));
    T* t12 = sa.ptr<T>(sa.take(2 * mb * Q));
    T* W = sa.ptr<T>(sa.take(mb * Q));
    int* active = sa.ptr<int>(sa.take(sizeof(int) * Q));
    unsigned long long* maxabs = sa.ptr<unsigned long long>(sa.take(sizeof(unsigned long long) * Q));
    int* nact = sa.ptr<int>(sa.take(sizeof(int) * 2));    // two-slot count of unconverged items

    hipStream_t s = c->stream;
    audit<T>(c, H00, n, n, Q, 1, "dec H00");
    audit<T>(c, V, n, n, Q, 1, "dec V");
    audit<T>(c, bw(ab0, 2 * nn, n, nullptr, nn), n, n, Q, 2, "dec ab");
    launch(c, NT_K_ELEM, 0, "dec_init",
           [&] { return launch_dec_init<T>(n, H00, V, alpha_dagger, bw(e1, nn, n), bw(es, nn, n),
                                           bw(ab0, 2 * nn, n, nullptr, nn), Q, s); });
    check_range(c, z, sizeof(T) * 2 * Q, "dec z");
    check_range(c, status, sizeof(int) * Q, "dec status");
    check_range(c, iters, sizeof(int) * Q, "dec iters");
    zero_dev(c, iters, sizeof(int) * Q, s);
    zero_dev(c, maxabs, sizeof(unsigned long long) * Q, s);
    zero_dev(c, nact, sizeof(int) * 2, s);
    launch(c, NT_K_ELEM, 0, "fill", [&] { return launch_fill_int(active, 1, Q, s); });
    // initial maxdiff = max|alpha| (surface.f90:484): the GEMM epilogue's
    // maxabs on a K = 0 product with Cin = alpha0 (a copy), then one update
    // (update 0: slot 0) that does not count as an iteration
    gemm<T>(c, n, n, 0, OP_N, none<T>(), OP_N, none<T>(), bw(W, nn, n), 1.0, bm<T>(ab0, 2 * nn, n), 1.0, nullptr, Q,
            0.0, nullptr, 1, nonew<T>(), maxabs);
    launch(c, NT_K_ELEM, 0, "dec_update",
           [&] { return launch_dec_step_update(active, maxabs, nullptr, conv, nact, nact + 1, Q, s); });

    // Iteration it runs its launches on the items still active (converged
    // items are masked inside every kernel, so an extra iteration changes
    // nothing) and ends with update it + 1.  Every dec_check iterations the
    // count of update it is copied to the host behind an event, and the host
    // waits for the copy of the previous check (dec_check iterations older)
    // and stops if that count was zero: the stream never runs dry while the
    // host decides, at the price of up to 2 dec_check - 1 masked iterations
    // after the last item has converged.  The flops of the profile follow
    // from the per-item iteration counts at the end.
    T* abc = ab0;
    T* abn = ab1;
    const int every = std::max(0, c->dec_check);
    int pend = -1;     // host slot of the outstanding check, or none
    for (int it = 0; it < max_iter; ++it) {
        if (every > 0 && it % every == 0) {
            const int slot = (it / every) & 1;
            hipcheck(c, hipMemcpyAsync(c->host_flag + slot, nact + (it & 1), sizeof(int), hipMemcpyDeviceToHost, s),
                     "memcpy");
            hipcheck(c, hipEventRecord(c->chk[slot], s), "event check");
            if (pend >= 0) {
                hipcheck(c, hipEventSynchronize(c->chk[pend]), "event sync");
                if (c->host_flag[pend] == 0) break;
            }
            pend = slot;
        }
        const double nf = 0.0;    // profiled flops: added from the iteration counts below
        // ive = (z - e1)^-1
        inverse<T>(c, n, bw(W, nn, n), bm<T>(e1, nn, n), z, bw(ive, nn, n), status, Q, nf, active);
        // [t1, t2] = [alpha, beta] ive
        gemm<T>(c, n, n, n, OP_N, bm<T>(abc, 2 * nn, n, nullptr, nn), OP_N, bm<T>(ive, nn, n),
                bw(t12, 2 * nn, n, nullptr, nn), 1.0, none<T>(), 0.0, nullptr, Q, nf, active, 2);
        // e1 += t1 beta ; es += t1 beta
        gemm<T>(c, n, n, n, OP_N, bm<T>(t12, 2 * nn, n), OP_N, bm<T>(abc + 2 * nn, 2 * nn, n), bw(e1, nn, n), 1.0,
                bm<T>(e1, nn, n), 1.0, nullptr, Q, nf, active, 1, bw(es, nn, n));
        // e1 += t2 alpha
        gemm<T>(c, n, n, n, OP_N, bm<T>(t12 + 2 * nn, 2 * nn, n), OP_N, bm<T>(abc, 2 * nn, n), bw(e1, nn, n), 1.0,
                bm<T>(e1, nn, n), 1.0, nullptr, Q, nf, active);
        // [alpha', beta'] = [t1 alpha, t2 beta], maxabs over alpha'
        gemm<T>(c, n, n, n, OP_N, bm<T>(t12, 2 * nn, n, nullptr, nn), OP_N, bm<T>(abc, 2 * nn, n, nullptr, nn),
                bw(abn, 2 * nn, n, nullptr, nn), 1.0, none<T>(), 0.0, nullptr, Q, nf, active, 2,
                nonew<T>(), maxabs);
        const int u = it + 1;
        launch(c, NT_K_ELEM, 0, "dec_update", [&] {
            return launch_dec_step_update(active, maxabs, iters, conv, nact + (u & 1), nact + ((u + 1) & 1), Q, s);
        });
        std::swap(abc, abn);
    }
    if (c->prof) {
        // every active item-iteration: one inverse and six n^3 products
        std::vector<int> h(Q);
        hipcheck(c, hipMemcpyAsync(h.data(), iters, sizeof(int) * Q, hipMemcpyDeviceToHost, s), "memcpy");
        hipcheck(c, hipStreamSynchronize(s), "sync");
        double tot = 0.0;
        for (int q = 0; q < Q; ++q) tot += h[q];
        const double f = 8.0 * n * (double)n * n * tot;
        c->flops_pending[NT_K_INV] += f;
        c->flops_pending[NT_K_GEMM] += 6.0 * f;
    }
    hipLaunchKernelGGL(k_dec_finish, dim3(nblk(Q)), dim3(256), 0, s, active, status, Q);
    hipcheck(c, hipGetLastError(), "dec_finish");
    // g = (z - es)^-1 for every item (surface.f90:497)
    inverse<T>(c, n, bw(W, nn, n), bm<T>(es, nn, n), z, out.g, status, Q);
    if (out.sigma.base) {
        // Sigma = alpha0 g beta0 = V^+ g V (alpha_dagger) ; t1 = V^+ g
        gemm<T>(c, n, n, n, alpha_dagger ? OP_C : OP_N, V, OP_N, ro(out.g), bw(t12, 2 * nn, n), 1.0, none<T>(), 0.0,
                nullptr, Q);
        gemm<T>(c, n, n, n, OP_N, bm<T>(t12, 2 * nn, n), alpha_dagger ? OP_N : OP_C, V, out.sigma, 1.0, none<T>(),
                0.0, nullptr, Q);
        if (out.gamma.base) gamma_op<T>(c, n, ro(out.sigma), out.gamma, Q);
    }
}

// ------------------------------------------------ circular decimation
// Leads that are nby x nbx transverse copies of a smaller cell have block-
// circulant H00 and V.  circular_g (negf/circular.f90:43-114) block-diagonalises
// them with the nested block DFT of their first block row (util.f90:732-760),
// decimates each nac x nac block on its own (nac = n / (nby nbx)) and returns
// g = P1 (1 (x) P2) blockdiag(g_s) (1 (x) P2)^+ P1^+ with P of set_pmatrix
// (circular.f90:9-40).  The small problems form one larger decimation batch.

// out[m * NB + i * nbx + l] = sum_{j,q} M_m[0:nac, j na + q nac + :] e^{i thy i j / nby} e^{i thx l q / nbx}
__global__ void k_circ_forward(int n, int nby, int nbx, double thy, double thx, const double2* M, int dag, double2* out,
                               int nm)
{
    const int NB = nby * nbx, na = n / nby, nac = na / nbx;
    const int o = blockIdx.y;
    const int m = o / NB, sidx = o % NB, i = sidx / nbx, l = sidx % nbx;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nac * nac || m >= nm) return;
    const int r = (int)(e % nac), cc = (int)(e / nac);
    const double2* A = M + (long)m * n * n;
    double ar = 0.0, ai = 0.0;
    for (int j = 0; j < nby; ++j)
        for (int q = 0; q < nbx; ++q) {
            const int col = j * na + q * nac + cc;
            double2 v = dag ? A[col + (long)r * n] : A[r + (long)col * n];
            if (dag) v.y = -v.y;
            const double ph = thy * (double)i / (double)nby * (double)j + thx * (double)l / (double)nbx * (double)q;
            double sn, cs;
            sincos(ph, &sn, &cs);
            ar += v.x * cs - v.y * sn;
            ai += v.x * sn + v.y * cs;
        }
    out[(long)o * nac * nac + r + (long)cc * nac] = make_double2(ar, ai);
}

// item expansion: small problem q*NB + s solves block s of lead matrix idx[q]
__global__ void k_circ_items(const int* idx, const double* z, int Q, int NB, int* idx_c, double* z_c)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Q * NB) return;
    const int q = t / NB, sidx = t % NB;
    idx_c[t] = (idx ? idx[q] : 0) * NB + sidx;
    z_c[2 * t] = z[2 * q];
    z_c[2 * t + 1] = z[2 * q + 1];
}

__global__ void k_circ_assemble(int n, int nac, int NB, const double2* gs, double2* D, int Q)
{
    const int q = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * n || q >= Q) return;
    const int r = (int)(e % n), c = (int)(e / n);
    const int sr = r / nac, sc = c / nac;
    double2 v = make_double2(0.0, 0.0);
    if (sr == sc) v = gs[((long)q * NB + sr) * nac * nac + (r % nac) + (long)(c % nac) * nac];
    D[(long)q * n * n + e] = v;
}

__global__ void k_circ_merge(const int* st_c, const int* it_c, int Q, int NB, int* status, int* iters)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Q) return;
    int st = 0, it = 0;
    for (int sidx = 0; sidx < NB; ++sidx) {
        const int v = st_c[q * NB + sidx];
        if (!st && v) st = v;
        it = max(it, it_c[q * NB + sidx]);
    }
    if (!status[q]) status[q] = st;
    iters[q] = it;
}

const double* circ_matrix(nt_ctx* c, int n, int nbx, int nby)
{
    for (auto& u : c->circ_u)
        if (u.n == n && u.nbx == nbx && u.nby == nby) return u.d;
    // P of set_pmatrix: P[(i na + k), (j na + k)] = e^{i theta i j / nb} / sqrt(nb)
    const double th = 6.283185307179586476925286766559;
    const int na = n / nby, nac = na / nbx;
    std::vector<double> P1((size_t)2 * n * n, 0.0), P2((size_t)2 * na * na, 0.0), U((size_t)2 * n * n, 0.0);
    auto pm = [&](std::vector<double>& P, int N, int nb) {
        const int blk = N / nb;
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                const double ph = th * (double)j / (double)nb * (double)i;
                for (int k = 0; k < blk; ++k) {
                    const size_t e = (size_t)(i * blk + k) + (size_t)(j * blk + k) * N;
                    P[2 * e] = cos(ph) / sqrt((double)nb);
                    P[2 * e + 1] = sin(ph) / sqrt((double)nb);
                }
            }
    };
    pm(P1, n, nby);
    pm(P2, na, nbx);
    // U = P1 * blockdiag(P2, ..., P2): U[:, i na + c] = P1[:, i na : (i+1) na] P2[:, c]
    for (int i = 0; i < nby; ++i)
        for (int cc = 0; cc < na; ++cc)
            for (int r = 0; r < n; ++r) {
                double sr = 0.0, si = 0.0;
                for (int k = 0; k < na; ++k) {
                    const size_t a = (size_t)r + (size_t)(i * na + k) * n, b = (size_t)k + (size_t)cc * na;
                    sr += P1[2 * a] * P2[2 * b] - P1[2 * a + 1] * P2[2 * b + 1];
                    si += P1[2 * a] * P2[2 * b + 1] + P1[2 * a + 1] * P2[2 * b];
                }
                const size_t o = (size_t)r + (size_t)(i * na + cc) * n;
                U[2 * o] = sr;
                U[2 * o + 1] = si;
            }
    (void)nac;
    double* d = nullptr;
    hipcheck(c, hipMalloc(&d, sizeof(double) * U.size()), "hipMalloc circ");
    hipcheck(c, hipMemcpy(d, U.data(), sizeof(double) * U.size(), hipMemcpyHostToDevice), "H2D circ");
    c->circ_u.push_back({n, nbx, nby, d});
    return d;
}

size_t decimate_circular_scratch(int n, int Q, int nm, int NB);

// Lead surface GFs of Q items through circular_g.  H00s/Vs hold nm dense lead
// matrices; item q uses matrix idx[q] (idx may be null when nm == 1).
void decimate_circular(nt_ctx* c, const nt_params* prm, int n, int Q, int nm, const double* H00s, const double* Vs,
                       const int* idx, int alpha_dagger, const double* z, DecimationOut<double> out, int* status,
                       int* iters, Arena& ar, size_t scratch_off)
{
    typedef double T;
    const int nbx = prm->n_bloch_x, nby = prm->n_bloch_y, NB = nbx * nby;
    const int nac = n / NB;
    const long nn = (long)n * n, ns = (long)nac * nac;
    Arena sa = ar;
    sa.off = scratch_off;
    T* Hs = sa.ptr<T>(sa.take(sizeof(T) * 2 * ns * nm * NB));
    T* As = sa.ptr<T>(sa.take(sizeof(T) * 2 * ns * nm * NB));
    int* idx_c = sa.ptr<int>(sa.take(sizeof(int) * Q * NB));
    T* z_c = sa.ptr<T>(sa.take(sizeof(T) * 2 * Q * NB));
    int* st_c = sa.ptr<int>(sa.take(sizeof(int) * Q * NB));
    int* it_c = sa.ptr<int>(sa.take(sizeof(int) * Q * NB));
    T* gs = sa.ptr<T>(sa.take(sizeof(T) * 2 * ns * Q * NB));
    T* D = sa.ptr<T>(sa.take(sizeof(T) * 2 * nn * Q));
    T* T1 = sa.ptr<T>(sa.take(sizeof(T) * 2 * nn * Q));
    const size_t dec_off = sa.off;
    hipStream_t s = c->stream;
    const double th = 6.283185307179586476925286766559;   // G.R of a lattice and its reciprocal
    check_range(c, H00s, sizeof(T) * 2 * nn * nm, "circ H00");
    check_range(c, Vs, sizeof(T) * 2 * nn * nm, "circ V");
    check_range(c, idx, sizeof(int) * Q, "circ idx");
    check_range(c, z, sizeof(T) * 2 * Q, "circ z");
    hipLaunchKernelGGL(k_circ_forward, dim3(nblk(ns), nm * NB), dim3(256), 0, s, n, nby, nbx, th, th,
                       reinterpret_cast<const double2*>(H00s), 0, reinterpret_cast<double2*>(Hs), nm);
    // alpha0 = V^+ or V is transformed itself (its first block row), beta = alpha^+
    hipLaunchKernelGGL(k_circ_forward, dim3(nblk(ns), nm * NB), dim3(256), 0, s, n, nby, nbx, th, th,
                       reinterpret_cast<const double2*>(Vs), alpha_dagger, reinterpret_cast<double2*>(As), nm);
    hipLaunchKernelGGL(k_circ_items, dim3(nblk((long)Q * NB)), dim3(256), 0, s, idx, z, Q, NB, idx_c, z_c);
    hipcheck(c, hipGetLastError(), "circ forward");
    zero_dev(c, st_c, sizeof(int) * Q * NB, s);
    DecimationOut<T> os{bw(gs, ns, nac), nonew<T>(), nonew<T>()};
    decimate<T>(c, nac, Q * NB, bm<T>(Hs, ns, nac, idx_c, 0, nm * NB), bm<T>(As, ns, nac, idx_c, 0, nm * NB), 0, z_c,
                prm->convergence, prm->max_iter, os, st_c, it_c, ar, dec_off);
    hipLaunchKernelGGL(k_circ_assemble, dim3(nblk(nn), Q), dim3(256), 0, s, n, nac, NB,
                       reinterpret_cast<const double2*>(gs), reinterpret_cast<double2*>(D), Q);
    hipLaunchKernelGGL(k_circ_merge, dim3(nblk(Q)), dim3(256), 0, s, st_c, it_c, Q, NB, status, iters);
    hipcheck(c, hipGetLastError(), "circ assemble");
    const T* U = circ_matrix(c, n, nbx, nby);
    reg_region(c, U, sizeof(T) * 2 * nn);
    gemm<T>(c, n, n, n, OP_N, bm<T>(U, 0, n), OP_N, bm<T>(D, nn, n), bw(T1, nn, n), 1.0, none<T>(), 0.0, nullptr, Q);
    gemm<T>(c, n, n, n, OP_N, bm<T>(T1, nn, n), OP_C, bm<T>(U, 0, n), out.g, 1.0, none<T>(), 0.0, nullptr, Q);
    if (out.sigma.base) {
        // Sigma = alpha0 g beta0 with the full lead coupling (surface.f90:504, :565)
        const BMat<T> V = bm<T>(Vs, nn, n, idx, 0, nm);
        gemm<T>(c, n, n, n, alpha_dagger ? OP_C : OP_N, V, OP_N, ro(out.g), bw(D, nn, n), 1.0, none<T>(), 0.0, nullptr,
                Q);
        gemm<T>(c, n, n, n, OP_N, bm<T>(D, nn, n), alpha_dagger ? OP_N : OP_C, V, out.sigma, 1.0, none<T>(), 0.0,
                nullptr, Q);
        if (out.gamma.base) gamma_op<T>(c, n, ro(out.sigma), out.gamma, Q);
    }
}

template <typename T>
size_t decimate_scratch(int n, int Q)
{
    Arena a;
    const size_t mb = sizeof(T) * 2 * (size_t)n * n;
    a.take(mb * Q);
    a.take(mb * Q);
    a.take(2 * mb * Q);
    a.take(2 * mb * Q);
    a.take(mb * Q);
    a.take(2 * mb * Q);
    a.take(mb * Q);
    a.take(sizeof(int) * Q);
    a.take(sizeof(unsigned long long) * Q);
    a.take(sizeof(int) * 2);
    return a.off;
}

size_t decimate_circular_scratch(int n, int Q, int nm, int NB)
{
    typedef double T;
    const int nac = n / NB;
    const size_t nn = (size_t)n * n, ns = (size_t)nac * nac;
    Arena a;
    a.take(sizeof(T) * 2 * ns * nm * NB);
    a.take(sizeof(T) * 2 * ns * nm * NB);
    a.take(sizeof(int) * Q * NB);
    a.take(sizeof(T) * 2 * Q * NB);
    a.take(sizeof(int) * Q * NB);
    a.take(sizeof(int) * Q * NB);
    a.take(sizeof(T) * 2 * ns * Q * NB);
    a.take(sizeof(T) * 2 * nn * Q);
    a.take(sizeof(T) * 2 * nn * Q);
    return a.off + decimate_scratch<T>(nac, Q * NB);
}

// Surface GFs (and Sigma, Gamma) of Q lead items; direct or circular decimation.
void lead_gfs(nt_ctx* c, const nt_params* prm, int n, int Q, int nm, const double* H00s, const double* Vs,
              const int* idx, int alpha_dagger, const double* z, DecimationOut<double> out, int* status, int* iters,
              Arena& ar, size_t scratch_off)
{
    if (prm->circular && prm->n_bloch_x * prm->n_bloch_y > 1) {
        decimate_circular(c, prm, n, Q, nm, H00s, Vs, idx, alpha_dagger, z, out, status, iters, ar, scratch_off);
    } else {
        const long nn = (long)n * n;
        decimate<double>(c, n, Q, bm<double>(H00s, nm > 1 || idx ? nn : 0, n, idx, 0, idx ? nm : 0),
                         bm<double>(Vs, nm > 1 || idx ? nn : 0, n, idx, 0, idx ? nm : 0), alpha_dagger, z,
                         prm->convergence, prm->max_iter, out, status, iters, ar, scratch_off);
    }
}

size_t lead_gfs_scratch(const nt_params* prm, int n, int Q, int nm)
{
    const int NB = prm->circular ? prm->n_bloch_x * prm->n_bloch_y : 1;
    return NB > 1 ? decimate_circular_scratch(n, Q, nm, NB) : decimate_scratch<double>(n, Q);
}

// Adjoint of T_LR (FP64 only).  Layers [lo, hi] of the device (0-based) form
// the window whose blocks get gradients.  All buffers hold P points of
// bmax^2 complex entries per slot (slot-major, then point).
struct AdjSpec {
    int lo, hi;
    int K;               // checkpoint interval: G^L of layers lo + n K kept, the rest recomputed
    const double* w;     // per-point weight of T_LR in the objective
    double* SL;          // left-connected G^L_i, i = lo, lo + K, ...   (ceil((L - lo) / K) slots)
    double* S0;          // left-connected corner G^L_{0,i}, same layers
    double* SegL;        // K > 1: one recomputed segment of G^L_i        (K - 1 slots)
    double* Seg0;        //        and of G^L_{0,i}                       (K - 1 slots)
    double* Wk;          // K > 1: inverse work slot of the recomputation
    int* st;             // K > 1: statuses of the recomputed inverses (P)
    double* CW;          // column G_{j,N+1}, j = lo..hi1              (hi1 - lo + 1 slots)
    double* RW;          // row G_{0,j}, j = lo..hi1                   (hi1 - lo + 1 slots)
    double* CA;          // G_{j,N+1} A                                 (hi1 - lo + 1 slots)
    double* Zd;          // X_{l,l}^+, l = lo..hi                       (hi - lo + 1 slots)
    double* Zc;          // X_{l+1,l}^+ + X_{l,l+1}, l = lo..hc         (hc - lo + 1 slots)
    double* A;           // w Gamma_R G_{0,N+1}^+ Gamma_L               (1 slot)
    double* GR;          // G_{N+1,N+1}                                 (1 slot)
    double* Cb[2];       // recursion ping-pong
    double* Rb[2];
};

inline int adj_hi1(int hi, int L) { return std::min(hi + 1, L - 1); }
inline int adj_hc(int hi, int L) { return std::min(hi, L - 2); }
// checkpoint interval for a setting (0: every layer kept, < 0: about sqrt(L - lo))
inline int adj_interval(int setting, int lo, int L)
{
    if (setting == 0) return 1;
    if (setting > 0) return std::min(setting, std::max(1, L - lo));
    return std::max(1, (int)std::ceil(std::sqrt((double)(L - lo))));
}
inline int adj_nchk(int lo, int L, int K) { return (L - lo + K - 1) / K; }
inline int adj_slots(int lo, int hi, int L, int K)
{
    const int nw1 = adj_hi1(hi, L) - lo + 1;
    const int keep = 2 * adj_nchk(lo, L, K) + (K > 1 ? 2 * (K - 1) + 1 : 0);
    return keep + 3 * nw1 + (hi - lo + 1) + std::max(0, adj_hc(hi, L) - lo + 1) + 6;
}

// carve the adjoint buffers (P points, slot size mm complex) from an arena
inline AdjSpec adj_carve(Arena& a, int lo, int hi, int L, int K, long mm, int P, const double* w)
{
    const size_t sb = sizeof(double) * 2 * (size_t)mm * P;
    const int nw1 = adj_hi1(hi, L) - lo + 1, nc = std::max(0, adj_hc(hi, L) - lo + 1);
    AdjSpec s;
    s.lo = lo;
    s.hi = hi;
    s.K = K;
    s.w = w;
    s.SL = a.ptr<double>(a.take(sb * adj_nchk(lo, L, K)));
    s.S0 = a.ptr<double>(a.take(sb * adj_nchk(lo, L, K)));
    s.SegL = s.Seg0 = s.Wk = nullptr;
    s.st = nullptr;
    if (K > 1) {
        s.SegL = a.ptr<double>(a.take(sb * (K - 1)));
        s.Seg0 = a.ptr<double>(a.take(sb * (K - 1)));
        s.Wk = a.ptr<double>(a.take(sb));
        s.st = a.ptr<int>(a.take(sizeof(int) * (size_t)P));
    }
    s.CW = a.ptr<double>(a.take(sb * nw1));
    s.RW = a.ptr<double>(a.take(sb * nw1));
    s.CA = a.ptr<double>(a.take(sb * nw1));
    s.Zd = a.ptr<double>(a.take(sb * (hi - lo + 1)));
    s.Zc = a.ptr<double>(a.take(sb * std::max(nc, 1)));
    s.A = a.ptr<double>(a.take(sb));
    s.GR = a.ptr<double>(a.take(sb));
    for (int i = 0; i < 2; ++i) {
        s.Cb[i] = a.ptr<double>(a.take(sb));
        s.Rb[i] = a.ptr<double>(a.take(sb));
    }
    return s;
}

__global__ void k_merge_status(const int* src, int* dst, int P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P && dst[p] == 0 && src[p] != 0) dst[p] = src[p];
}

__global__ void k_scale_points(double2* A, long n, long stride, const double* w, int P)
{
    const int p = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || e >= n) return;
    const double f = w[p];
    double2 v = A[(long)p * stride + e];
    A[(long)p * stride + e] = make_double2(v.x * f, v.y * f);
}

// out[k][blk][e] += 2 sum_{p : k(p) = k, point ok} Z[blk][p][e]
__global__ void k_adj_reduce(const double2* Z, long mm, int P, long ne, const int* kidx, const int* stdec,
                             const int* strgf, const int* first_mask, double2* out, int nblk)
{
    const int blk = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne || blk >= nblk) return;
    const double2* z = Z + (long)blk * mm * P + e;
    int kc = -1;
    double ar = 0.0, ai = 0.0;
    for (int p = 0; p < P; ++p) {
        const bool ok = !stdec[2 * p] && !stdec[2 * p + 1] && !strgf[p] && !(first_mask && first_mask[p]);
        if (!ok) continue;
        const int k = kidx ? kidx[p] : 0;
        if (k != kc) {
            if (kc >= 0) {
                double2* o = out + ((long)kc * nblk + blk) * ne + e;
                o->x += 2.0 * ar;
                o->y += 2.0 * ai;
            }
            kc = k;
            ar = ai = 0.0;
        }
        const double2 v = z[(long)p * mm];
        ar += v.x;
        ai += v.y;
    }
    if (kc >= 0) {
        double2* o = out + ((long)kc * nblk + blk) * ne + e;
        o->x += 2.0 * ar;
        o->y += 2.0 * ai;
    }
}

template <typename T>
size_t rgf_scratch(int nl, int nr, int bmax, int P)
{
    const int m = std::max(bmax, std::max(nl, nr));
    const size_t mb = sizeof(T) * 2 * (size_t)m * m;
    Arena a;
    for (int i = 0; i < 11; ++i) a.take(mb * P);
    // the right-to-left sweep's own buffers and status (concurrent sweeps)
    for (int i = 0; i < 10; ++i) a.take(mb * P);
    a.take(sizeof(int) * P);
    // one-sweep corner correction: M11, M22, d G_c0, d G_0c; Y, S, S^-1, T;
    // R, Q (one slot each) and the corrected column ((nl+nr) nl <= 2 slots)
    a.take(12 * mb * P);
    return a.off;
}

// Adjoint of T_LR = Re Tr[G_{0,N+1} Gamma_R G_{0,N+1}^+ Gamma_L] with respect to the
// device blocks.  G is the Green's function of the system the left-to-right
// sweep inverts (layer 0 closed by the left surface GF, layers 1..N+1 at the
// device broadening, Sigma_R on N+1).  For a change dH of the device blocks
// dG = G dH G, so dT = 2 Re Tr[dH X] with X = G Gamma_R G^+ Gamma_L G and
//   X_{j,i} = G_{j,N+1} A G_{0,i},   A = Gamma_R G_{0,N+1}^+ Gamma_L.
// The column and the row come from the stored left-connected blocks by the
// Dyson identities G = g + g V G and G = g + G V g of the cut between layers
// j and j+1:
//   G_{j,N+1} = G^L_j V_{j,j+1} G_{j+1,N+1},
//   G_{0,j}   = G^L_{0,j} + G_{0,j+1} V_{j+1,j} G^L_j.
// The gradients (real objective, complex blocks, d/dRe + i d/dIm) are
// 2 X_{l,l}^+ for Hamd_l and 2 (X_{l+1,l}^+ + X_{l,l+1}) for Hamdv_l, whose
// dagger is the lower block.  The weight w multiplies A.
template <typename T>
void adjoint_backward(nt_ctx* c, int nl, int nr, const RgfInput<T>& in, const T* zd, int P, long mm,
                      BMatW<T> gdlr, BMatW<T> gRo, const BMat<T>& Vc, int opVc, int opVcH, const AdjSpec& adj,
                      T* T1, T* T2)
{
    const int L = in.L;
    const long slot = mm * P;
    auto W = [&](T* p, int ld) { return bw(p, mm, ld); };
    auto R = [&](T* p, int ld) { return bm<T>(p, mm, ld); };
    auto at = [&](double* base, int k) { return reinterpret_cast<T*>(base) + 2 * (long)k * slot; };
    // G^L_j and G^L_{0,j} of a layer: a checkpoint, or its segment recomputed
    // from the checkpoint by the left-to-right recursion (the same launches
    // as the forward sweep, so the same values)
    const int K = adj.K;
    int seg_have = -1;
    auto seg_ensure = [&](int j) {
        const int sg = (j - adj.lo) / K;
        if (K == 1 || sg == seg_have) return;
        const int c0 = adj.lo + sg * K, c1 = std::min(c0 + K - 1, L - 1);
        T* Gp = at(adj.SL, sg);
        T* G0p = at(adj.S0, sg);
        T* Aw = reinterpret_cast<T*>(adj.Wk);
        for (int i = c0 + 1; i <= c1; ++i) {
            const int bp = in.bsz[i - 1], bi = in.bsz[i];
            const BMat<T>& V = in.Hdv[i - 1];
            T* Gi = at(adj.SegL, i - c0 - 1);
            T* G0i = at(adj.Seg0, i - c0 - 1);
            const T* zi = in.zl ? in.zl + 2 * (long)i * P : zd;
            gemm<T>(c, bi, bp, bp, OP_C, V, OP_N, R(Gp, bp), W(T1, bi), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            gemm<T>(c, bi, bi, bp, OP_N, R(T1, bi), OP_N, V, W(Aw, bi), -1.0, in.Hd[i], -1.0, zi, P, -1, in.act);
            inverse<T>(c, bi, W(Aw, bi), none<T>(), nullptr, W(Gi, bi), adj.st, P, -1, in.act);
            gemm<T>(c, nl, bi, bp, OP_N, R(G0p, nl), OP_N, V, W(T2, nl), 1.0, none<T>(), 0.0, nullptr, P, -1,
                    in.act);
            gemm<T>(c, nl, bi, bi, OP_N, R(T2, nl), OP_N, R(Gi, bi), W(G0i, nl), 1.0, none<T>(), 0.0, nullptr, P,
                    -1, in.act);
            Gp = Gi;
            G0p = G0i;
        }
        seg_have = sg;
    };
    auto GLj = [&](int j) -> T* {
        seg_ensure(j);
        const int sg = (j - adj.lo) / K, off = (j - adj.lo) % K;
        return off == 0 ? at(adj.SL, sg) : at(adj.SegL, off - 1);
    };
    auto G0j = [&](int j) -> T* {
        seg_ensure(j);
        const int sg = (j - adj.lo) / K, off = (j - adj.lo) % K;
        return off == 0 ? at(adj.S0, sg) : at(adj.Seg0, off - 1);
    };
    if (K > 1) zero_dev(c, adj.st, sizeof(int) * P, c->stream);
    const int lo = adj.lo, hi = adj.hi, hi1 = adj_hi1(hi, L), hc = adj_hc(hi, L);
    T* A = reinterpret_cast<T*>(adj.A);
    hipStream_t s = c->stream;
    // A = w Gamma_R G_{0,N+1}^+ Gamma_L    (nr x nl)
    gemm<T>(c, nr, nl, nr, OP_N, in.gmr, OP_C, ro(gdlr), W(T1, nr), 1.0, none<T>(), 0.0, nullptr, P);
    gemm<T>(c, nr, nl, nl, OP_N, R(T1, nr), OP_N, in.gml, W(A, nr), 1.0, none<T>(), 0.0, nullptr, P);
    check_range(c, adj.w, sizeof(double) * P, "adjoint weights");
    hipLaunchKernelGGL(k_scale_points, dim3(nblk((long)nr * nl), P), dim3(256), 0, s, reinterpret_cast<double2*>(A),
                       (long)nr * nl, mm, adj.w, P);
    hipcheck(c, hipGetLastError(), "scale points");
    // backward recursion from the closing layer N+1 to the window
    BMat<T> Ccur = ro(gRo);      // G_{N+1,N+1}   (nr x nr)
    BMat<T> Rcur = ro(gdlr);     // G_{0,N+1}     (nl x nr)
    int pp = 0;
    for (int j = L - 1; j >= lo; --j) {
        const int bj = in.bsz[j];
        const bool last = j == L - 1;
        const int bn = last ? nr : in.bsz[j + 1];
        const BMat<T> V = last ? Vc : in.Hdv[j];          // V_{j,j+1} = op(V)
        const int opV = last ? opVc : OP_N, opVH = last ? opVcH : OP_C;
        T* Cn = j <= hi1 ? at(adj.CW, j - lo) : reinterpret_cast<T*>(adj.Cb[pp]);
        T* Rn = j <= hi1 ? at(adj.RW, j - lo) : reinterpret_cast<T*>(adj.Rb[pp]);
        pp ^= 1;
        // (the recomputation of a segment uses T1 / T2: it runs first)
        T* glj = GLj(j);
        T* g0j = G0j(j);
        // G_{j,N+1} = G^L_j (V_{j,j+1} G_{j+1,N+1})
        gemm<T>(c, bj, nr, bn, opV, V, OP_N, Ccur, W(T1, bj), 1.0, none<T>(), 0.0, nullptr, P);
        gemm<T>(c, bj, nr, bj, OP_N, R(glj, bj), OP_N, R(T1, bj), W(Cn, bj), 1.0, none<T>(), 0.0, nullptr, P);
        // G_{0,j} = G^L_{0,j} + (G_{0,j+1} V_{j+1,j}) G^L_j
        gemm<T>(c, nl, bj, bn, OP_N, Rcur, opVH, V, W(T2, nl), 1.0, none<T>(), 0.0, nullptr, P);
        gemm<T>(c, nl, bj, bj, OP_N, R(T2, nl), OP_N, R(glj, bj), W(Rn, nl), 1.0, R(g0j, nl), 1.0, nullptr, P);
        Ccur = R(Cn, bj);
        Rcur = R(Rn, nl);
    }
    // X blocks of the window (stored as the gradient operands)
    for (int j = lo; j <= hi1; ++j)
        gemm<T>(c, in.bsz[j], nl, nr, OP_N, R(at(adj.CW, j - lo), in.bsz[j]), OP_N, R(A, nr),
                W(at(adj.CA, j - lo), in.bsz[j]), 1.0, none<T>(), 0.0, nullptr, P);
    for (int l = lo; l <= hi; ++l) {
        const int bl = in.bsz[l];
        // X_{l,l}^+ = G_{0,l}^+ (G_{l,N+1} A)^+
        gemm<T>(c, bl, bl, nl, OP_C, R(at(adj.RW, l - lo), nl), OP_C, R(at(adj.CA, l - lo), bl),
                W(at(adj.Zd, l - lo), bl), 1.0, none<T>(), 0.0, nullptr, P);
    }
    for (int l = lo; l <= hc; ++l) {
        const int bl = in.bsz[l], bn = in.bsz[l + 1];
        T* Z = at(adj.Zc, l - lo);
        // X_{l+1,l}^+ = G_{0,l}^+ (G_{l+1,N+1} A)^+ ; + X_{l,l+1} = (G_{l,N+1} A) G_{0,l+1}
        gemm<T>(c, bl, bn, nl, OP_C, R(at(adj.RW, l - lo), nl), OP_C, R(at(adj.CA, l + 1 - lo), bn), W(Z, bl), 1.0,
                none<T>(), 0.0, nullptr, P);
        gemm<T>(c, bl, bn, nl, OP_N, R(at(adj.CA, l - lo), bl), OP_N, R(at(adj.RW, l + 1 - lo), nl), W(Z, bl), 1.0,
                R(Z, bl), 1.0, nullptr, P);
    }
}

template <typename T>
void rgf_run(nt_ctx* c, int nl, int nr, const RgfInput<T>& in, const T* zd, int P, RgfOutput<T> out, Arena& ar,
             size_t scratch_off, const AdjSpec* adj = nullptr)
{
    int bmax = std::max(nl, nr);
    for (int b : in.bsz) bmax = std::max(bmax, b);
    const long mm = (long)bmax * bmax;
    const size_t mb = sizeof(T) * 2 * mm;
    Arena sa = ar;
    sa.off = scratch_off;
    T* G = sa.ptr<T>(sa.take(mb * P));
    T* Gn = sa.ptr<T>(sa.take(mb * P));
    T* Aw = sa.ptr<T>(sa.take(mb * P));
    T* T1 = sa.ptr<T>(sa.take(mb * P));
    T* G0 = sa.ptr<T>(sa.take(mb * P));
    T* G0n = sa.ptr<T>(sa.take(mb * P));
    T* T2 = sa.ptr<T>(sa.take(mb * P));
    T* GLR = sa.ptr<T>(sa.take(mb * P));
    T* GRL = sa.ptr<T>(sa.take(mb * P));
    T* X = sa.ptr<T>(sa.take(mb * P));
    T* Y = sa.ptr<T>(sa.take(mb * P));
    // buffers of the right-to-left sweep
    T* G_b = sa.ptr<T>(sa.take(mb * P));
    T* Gn_b = sa.ptr<T>(sa.take(mb * P));
    T* Aw_b = sa.ptr<T>(sa.take(mb * P));
    T* T1_b = sa.ptr<T>(sa.take(mb * P));
    T* G0_b = sa.ptr<T>(sa.take(mb * P));
    T* G0n_b = sa.ptr<T>(sa.take(mb * P));
    T* T2_b = sa.ptr<T>(sa.take(mb * P));
    T* GRL_b = sa.ptr<T>(sa.take(mb * P));
    T* X_b = sa.ptr<T>(sa.take(mb * P));
    T* Y_b = sa.ptr<T>(sa.take(mb * P));
    int* st_b = sa.ptr<int>(sa.take(sizeof(int) * P));
    T* CM = sa.ptr<T>(sa.take(4 * mb * P));      // corner blocks M11, M22, d G_c0, d G_0c
    T* CMi = sa.ptr<T>(sa.take(4 * mb * P));     // Y = M11^-1, S, S^-1, T
    T* CC0 = sa.ptr<T>(sa.take(2 * mb * P));     // R, Q
    T* CGp = sa.ptr<T>(sa.take(2 * mb * P));     // corrected column [G_L; G_RL]
    (void)GRL;
    hipStream_t s = c->stream;
    const int L = in.L;
    auto W = [&](T* p, int ld) { return bw(p, mm, ld); };
    auto R = [&](T* p, int ld) { return bm<T>(p, mm, ld); };
    // diagonal shift of device layer i
    auto zdev = [&](int i) -> const T* { return in.zl ? in.zl + 2 * (long)i * P : zd; };
    auto keep = [&](int i) -> T* { return in.keep ? in.keep + 2 * (long)i * mm * P : nullptr; };
    BMatW<T> gdlr = out.G_D_LR.base ? out.G_D_LR : W(GLR, nl);
    BMatW<T> gdrl = out.G_D_RL.base ? out.G_D_RL : W(GRL, nr);
    BMatW<T> gLo = out.G_L.base ? out.G_L : W(Gn, nl);
    BMatW<T> gRo = out.G_R.base ? out.G_R : W(Gn, nr);
    int* st = out.status;

    // adjoint storage slots of device layer i (null when layer i is not stored)
    const long slot = mm * P;
    // adjoint storage: the checkpoint layers lo + n K (every layer when K = 1)
    auto kept = [&](int i) { return adj && i >= adj->lo && (i - adj->lo) % adj->K == 0; };
    auto slotL = [&](int i) -> T* {
        return kept(i) ? reinterpret_cast<T*>(adj->SL) + 2 * (long)((i - adj->lo) / adj->K) * slot : nullptr;
    };
    auto slot0 = [&](int i) -> T* {
        return kept(i) ? reinterpret_cast<T*>(adj->S0) + 2 * (long)((i - adj->lo) / adj->K) * slot : nullptr;
    };
    // one sweep: the right-to-left quantities follow from the left-to-right
    // system's corner blocks (see one_sweep below); its buffers are free
    const bool one = std::is_same<T, double>::value && c->one_sweep && !in.keep && !in.zl &&
                     in.zlead && !in.Hrd_fwd.base && in.Hdv_back.empty();
    // the right-to-left sweep (two sweeps) runs on the auxiliary stream from
    // here; with one sweep only the main stream is used
    if (!one && c->overlap && c->aux) {
        hipcheck(c, hipEventRecord(c->fork, c->stream), "event fork");
        hipcheck(c, hipStreamWaitEvent(c->aux, c->fork, 0), "wait fork");
    }
    T* Gc = G_b;         // left-connected column G^L_{i,0}   (b x nl), ping-pong with Gcn
    T* Gcn = Gn_b;
    T* G00 = Aw_b;       // left-connected corner G^L_{00,i}  (nl x nl)
    T* T3 = T1_b;        // V^+ G^L_{i-1,0}                   (b x nl)
    T* const gbufs[2] = {G, Gn};
    T* const g0bufs[2] = {G0, G0n};
    auto other = [](T* const* bufs, T* cur) { return cur == bufs[0] ? bufs[1] : bufs[0]; };
    if (adj) gRo = W(reinterpret_cast<T*>(adj->GR), nr);

    // ---------------- left to right
    {
        const int b = in.bsz[0];
        if (T* sl = slotL(0)) G = sl;
        if (T* s0 = slot0(0)) G0 = s0;
        if (L > 1 && keep(0)) G = keep(0);
        gemm<T>(c, b, nl, nl, OP_C, in.Hld, OP_N, in.gl, W(T1, b), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        gemm<T>(c, b, b, nl, OP_N, R(T1, b), OP_N, in.Hld, W(Aw, b), -1.0, in.Hd[0], -1.0, zdev(0), P, -1, in.act);
        if (L == 1 && keep(0)) {
            // one layer: its full G, closed by the right lead's surface
            // function, inv(Aw - Hrd^+ gr Hrd) (T1 is free: no corner column
            // is formed when blocks are kept)
            gemm<T>(c, b, nr, nr, OP_C, in.Hrd, OP_N, in.gr, W(T1, b), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            gemm<T>(c, b, b, nr, OP_N, R(T1, b), OP_N, in.Hrd, W(X, b), -1.0, R(Aw, b), 1.0, nullptr, P, -1, in.act);
            inverse<T>(c, b, W(X, b), none<T>(), nullptr, W(keep(0), b), st, P, -1, in.act);
        }
        inverse<T>(c, b, W(Aw, b), none<T>(), nullptr, W(G, b), st, P, -1, in.act);
        gemm<T>(c, nl, b, nl, OP_N, in.gl, OP_N, in.Hld, W(T2, nl), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        gemm<T>(c, nl, b, b, OP_N, R(T2, nl), OP_N, R(G, b), W(G0, nl), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        if (one) {
            // G^L_{1,0} = G^L_1 (Hld^+ gl), G^L_{00,1} = gl + G^L_{0,1} (Hld^+ gl)
            gemm<T>(c, b, nl, b, OP_N, R(G, b), OP_N, R(T1, b), W(Gc, b), 1.0, none<T>(), 0.0, nullptr, P, -1,
                    in.act);
            gemm<T>(c, nl, nl, b, OP_N, R(G0, nl), OP_N, R(T1, b), W(G00, nl), 1.0, in.gl, 1.0, nullptr, P, -1,
                    in.act);
        }
        for (int i = 1; i < L; ++i) {
            const int bp = in.bsz[i - 1], bi = in.bsz[i];
            const BMat<T>& V = in.Hdv[i - 1];
            T* Gi = slotL(i) ? slotL(i) : (keep(i) && i < L - 1 ? keep(i) : other(gbufs, G));
            T* G0i = slot0(i) ? slot0(i) : other(g0bufs, G0);
            gemm<T>(c, bi, bp, bp, OP_C, V, OP_N, R(G, bp), W(T1, bi), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            gemm<T>(c, bi, bi, bp, OP_N, R(T1, bi), OP_N, V, W(Aw, bi), -1.0, in.Hd[i], -1.0, zdev(i), P, -1, in.act);
            if (i == L - 1 && keep(i)) {
                // full G of the last layer, closed by the right lead's surface
                // function: inv(Aw - Hrd^+ gr Hrd)  (negf/lesser.f90:587-600)
                gemm<T>(c, bi, nr, nr, OP_C, in.Hrd, OP_N, in.gr, W(T1, bi), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
                gemm<T>(c, bi, bi, nr, OP_N, R(T1, bi), OP_N, in.Hrd, W(X, bi), -1.0, R(Aw, bi), 1.0, nullptr, P, -1, in.act);
                inverse<T>(c, bi, W(X, bi), none<T>(), nullptr, W(keep(i), bi), st, P, -1, in.act);
            }
            inverse<T>(c, bi, W(Aw, bi), none<T>(), nullptr, W(Gi, bi), st, P, -1, in.act);
            gemm<T>(c, nl, bi, bp, OP_N, R(G0, nl), OP_N, V, W(T2, nl), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            gemm<T>(c, nl, bi, bi, OP_N, R(T2, nl), OP_N, R(Gi, bi), W(G0i, nl), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            if (one) {
                // T3 = V^+ G^L_{i-1,0}; G^L_{i,0} = G^L_i T3; G^L_{00,i} += G^L_{0,i} T3
                gemm<T>(c, bi, nl, bp, OP_C, V, OP_N, R(Gc, bp), W(T3, bi), 1.0, none<T>(), 0.0, nullptr, P, -1,
                        in.act);
                gemm<T>(c, bi, nl, bi, OP_N, R(Gi, bi), OP_N, R(T3, bi), W(Gcn, bi), 1.0, none<T>(), 0.0, nullptr, P,
                        -1, in.act);
                gemm<T>(c, nl, nl, bi, OP_N, R(G0i, nl), OP_N, R(T3, bi), W(G00, nl), 1.0, R(G00, nl), 1.0, nullptr,
                        P, -1, in.act);
                std::swap(Gc, Gcn);
            }
            G = Gi;
            G0 = G0i;
        }
        if (!out.G_R.base && !adj) gRo = W(other(gbufs, G), nr);
        const int bN = in.bsz[L - 1];
        // closing lead layer N+1: Vc = V_{N,N+1} = Hrd_fwd or Hrd^+; (Vc^+ G) Vc
        const bool fwd = in.Hrd_fwd.base != nullptr;
        const BMat<T> Vc = fwd ? in.Hrd_fwd : in.Hrd;
        const int opVc = fwd ? OP_N : OP_C;      // op(.) giving Vc      (bN x nr)
        const int opVcH = fwd ? OP_C : OP_N;     // op(.) giving Vc^+    (nr x bN)
        gemm<T>(c, nr, bN, bN, opVcH, Vc, OP_N, R(G, bN), W(T1, nr), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        // Aw = (z - Hamr - sr) - T1 Vc   (lesser.f90:616-633)
        shift_sub<T>(c, nr, zd, in.Hamr, in.sr, W(X, nr), P, in.act);
        gemm<T>(c, nr, nr, bN, OP_N, R(T1, nr), opVc, Vc, W(Aw, nr), -1.0, R(X, nr), 1.0, nullptr, P, -1, in.act);
        inverse<T>(c, nr, W(Aw, nr), none<T>(), nullptr, gRo, st, P, -1, in.act);
        gemm<T>(c, nl, nr, bN, OP_N, R(G0, nl), opVc, Vc, W(T2, nl), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        gemm<T>(c, nl, nr, nr, OP_N, R(T2, nl), OP_N, ro(gRo), gdlr, 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        // Caroli T_LR = Re Tr[(G_LR gmr)(G_LR^+ gml)]
        gemm<T>(c, nl, nr, nr, OP_N, ro(gdlr), OP_N, in.gmr, W(X, nl), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        gemm<T>(c, nr, nl, nl, OP_C, ro(gdlr), OP_N, in.gml, W(Y, nr), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        trace_prod<T>(c, nl, nr, R(X, nl), R(Y, nr), out.T_LR, P, in.act);
        if (adj) adjoint_backward<T>(c, nl, nr, in, zd, P, mm, gdlr, gRo, Vc, opVc, opVcH, *adj, T1, T2);
        if (one) {
            // One sweep.  The right-to-left sweep inverts the same block-
            // tridiagonal system as this one except at its two ends: there the
            // left lead's surface layer is an explicit block z_dev - Haml - sl
            // (here folded in through gl, whose inverse is z_lead - Haml - sl
            // by the decimation's fixed point) and the right lead's surface
            // layer is folded in through gr (here the explicit closing block
            // z_dev - Hamr - sr).  The two systems differ by
            // Delta = diag(d I, -d I) on these end blocks, d = z_dev - z_lead,
            // so with the corner blocks G_CC of this system's inverse
            // (C = {left surface, closing layer}) the other's are
            //   G'_CC = (I + G_CC Delta)^-1 G_CC,
            // G_RL = G'_c0 and G_L = G'_00 (negf/lesser.f90:694-793 computes
            // them by the second sweep).
            const int n2 = nl + nr;
            T* Tc = T2_b;
            T* Gc0 = X_b;
            // Tc = Vc^+ G^L_{N,0}; G_c0 = G_cc Tc; G_00 = G^L_{00,N} + G_0c Tc
            gemm<T>(c, nr, nl, bN, opVcH, Vc, OP_N, R(Gc, bN), W(Tc, nr), 1.0, none<T>(), 0.0, nullptr, P, -1,
                    in.act);
            gemm<T>(c, nr, nl, nr, OP_N, ro(gRo), OP_N, R(Tc, nr), W(Gc0, nr), 1.0, none<T>(), 0.0, nullptr, P, -1,
                    in.act);
            gemm<T>(c, nl, nl, nr, OP_N, ro(gdlr), OP_N, R(Tc, nr), W(G00, nl), 1.0, R(G00, nl), 1.0, nullptr, P, -1,
                    in.act);
            // G'_CC [I; 0] = M^-1 [G_00; G_c0] with M = I + G_CC Delta, by block
            // elimination over the nl and nr blocks (M11 = I + d G_00 is I up
            // to d G_00, d of the order of the broadening):
            //   Y = M11^-1, T = d G_c0 Y, S = (I - d G_cc) + T d G_0c,
            //   X2 = S^-1 (G_c0 - T G_00), X1 = Y (G_00 + d G_0c X2);
            // G_RL = X2, G_L = X1 (two inverses of nl and nr and six GEMMs
            // instead of an inverse of nl + nr)
            const long m4 = 4 * mm, m2 = 2 * mm;
            auto sl = [&](T* base, long stride, int k, int ld) { return bw(base + 2 * (long)k * mm, stride, ld); };
            const BMatW<T> M11 = sl(CM, m4, 0, nl), M22 = sl(CM, m4, 1, nr), dGc0 = sl(CM, m4, 2, nr),
                           dG0c = sl(CM, m4, 3, nl);
            const BMatW<T> Yw = sl(CMi, m4, 0, nl), Sw = sl(CMi, m4, 1, nr), Siw = sl(CMi, m4, 2, nr),
                           Tw = sl(CMi, m4, 3, nr);
            const BMatW<T> Rw = sl(CC0, m2, 0, nr), Qw = sl(CC0, m2, 1, nl);
            const BMatW<T> X1w = bw(CGp, m2, n2), X2w = bw(CGp + 2 * (long)nl, m2, n2);
            audit<T>(c, R(G00, nl), nl, nl, P, 1, "corner G00");
            audit<T>(c, ro(gdlr), nl, nr, P, 1, "corner G0c");
            audit<T>(c, R(Gc0, nr), nr, nl, P, 1, "corner Gc0");
            audit<T>(c, ro(gRo), nr, nr, P, 1, "corner Gcc");
            audit<T>(c, M11, nl, nl, P, 1, "corner M11");
            audit<T>(c, M22, nr, nr, P, 1, "corner M22");
            audit<T>(c, dGc0, nr, nl, P, 1, "corner dGc0");
            audit<T>(c, dG0c, nl, nr, P, 1, "corner dG0c");
            audit<T>(c, X1w, n2, nl, P, 1, "corner X");
            check_range(c, in.zlead, sizeof(T) * ((size_t)in.zlead_stride * (P - 1) + 2), "corner z_lead");
            check_range(c, zd, sizeof(T) * 2 * P, "corner z_dev");
            check_range(c, in.act, sizeof(int) * P, "corner active");
            const int nx = std::max(nl, nr);
            launch(c, NT_K_ELEM, 0, "corner blocks", [&] {
                hipLaunchKernelGGL(k_corner_blocks, dim3(nblk((long)nx * nx), P, 4), dim3(256), 0, s, nl, nr,
                                   R(G00, nl), ro(gdlr), R(Gc0, nr), ro(gRo), zd, in.zlead, in.zlead_stride, M11, M22,
                                   dGc0, dG0c, in.act);
                return hipGetLastError();
            });
            inverse<T>(c, nl, M11, none<T>(), nullptr, Yw, st, P, -1, in.act);
            gemm<T>(c, nr, nl, nl, OP_N, ro(dGc0), OP_N, ro(Yw), Tw, 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            gemm<T>(c, nr, nr, nl, OP_N, ro(Tw), OP_N, ro(dG0c), Sw, 1.0, ro(M22), 1.0, nullptr, P, -1, in.act);
            gemm<T>(c, nr, nl, nl, OP_N, ro(Tw), OP_N, R(G00, nl), Rw, -1.0, R(Gc0, nr), 1.0, nullptr, P, -1, in.act);
            inverse<T>(c, nr, Sw, none<T>(), nullptr, Siw, st, P, -1, in.act);
            gemm<T>(c, nr, nl, nr, OP_N, ro(Siw), OP_N, ro(Rw), X2w, 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            gemm<T>(c, nl, nl, nr, OP_N, ro(dG0c), OP_N, ro(X2w), Qw, 1.0, R(G00, nl), 1.0, nullptr, P, -1, in.act);
            gemm<T>(c, nl, nl, nl, OP_N, ro(Yw), OP_N, ro(Qw), X1w, 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            // G_RL = rows nl.. of the corrected column, G_L = rows 0..nl-1
            const BMat<T> grl = bm<T>(CGp + 2 * (long)nl, m2, n2);
            const BMat<T> gll = bm<T>(CGp, m2, n2);
            if (out.G_D_RL.base) copy_op<T>(c, nr, nl, grl, out.G_D_RL, 0, (T)1, P, in.act);
            if (out.G_L.base) copy_op<T>(c, nl, nl, gll, out.G_L, 0, (T)1, P, in.act);
            // Caroli T_RL = Re Tr[(G_RL gml)(G_RL^+ gmr)]
            gemm<T>(c, nr, nl, nl, OP_N, grl, OP_N, in.gml, W(X_b, nr), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            gemm<T>(c, nl, nr, nr, OP_C, grl, OP_N, in.gmr, W(Y_b, nl), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            trace_prod<T>(c, nr, nl, R(X_b, nr), R(Y_b, nl), out.T_RL, P, in.act);
            (void)gLo;
            return;
        }
    }
    // ---------------- right to left, on the auxiliary stream when enabled: it
    // needs only the inputs, so it runs beside the left-to-right sweep and
    // fills the CUs the latency-bound inverses of either sweep leave idle.
    // (The left-to-right part above was only enqueued; the fork event is
    // recorded before it on the main stream.)
    G = G_b;
    Gn = Gn_b;
    G0 = G0_b;
    G0n = G0n_b;
    Aw = Aw_b;
    T1 = T1_b;
    T2 = T2_b;
    X = X_b;
    Y = Y_b;
    hipStream_t main_stream = c->stream;
    const bool two = c->overlap && c->aux;
    if (two) {
        c->stream = c->aux;
        s = c->aux;
    }
    zero_dev(c, st_b, sizeof(int) * P, s);
    st = st_b;
    gdrl = out.G_D_RL.base ? out.G_D_RL : W(GRL_b, nr);
    {
        const int bN = in.bsz[L - 1];
        gemm<T>(c, bN, nr, nr, OP_C, in.Hrd, OP_N, in.gr, W(T1, bN), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        gemm<T>(c, bN, bN, nr, OP_N, R(T1, bN), OP_N, in.Hrd, W(Aw, bN), -1.0, in.Hd[L - 1], -1.0, zdev(L - 1), P, -1, in.act);
        inverse<T>(c, bN, W(Aw, bN), none<T>(), nullptr, W(G, bN), st, P, -1, in.act);
        gemm<T>(c, nr, bN, nr, OP_N, in.gr, OP_N, in.Hrd, W(T2, nr), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        gemm<T>(c, nr, bN, bN, OP_N, R(T2, nr), OP_N, R(G, bN), W(G0, nr), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        for (int i = L - 1; i >= 1; --i) {
            const int bi = in.bsz[i], bm_ = in.bsz[i - 1];
            // V = V_{i,i-1}: Hdv_back[i-1] or Hdv[i-1]^+ (bi x bm_)
            const bool back = in.Hdv_back.size() > (size_t)(i - 1) && in.Hdv_back[i - 1].base;
            const BMat<T> Vb = back ? in.Hdv_back[i - 1] : in.Hdv[i - 1];
            const int opV = back ? OP_N : OP_C;    // gives V   (bi x bm_)
            const int opVH = back ? OP_C : OP_N;   // gives V^+ (bm_ x bi)
            gemm<T>(c, bm_, bi, bi, opVH, Vb, OP_N, R(G, bi), W(T1, bm_), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            gemm<T>(c, bm_, bm_, bi, OP_N, R(T1, bm_), opV, Vb, W(Aw, bm_), -1.0, in.Hd[i - 1], -1.0, zdev(i - 1),
                    P, -1, in.act);
            inverse<T>(c, bm_, W(Aw, bm_), none<T>(), nullptr, W(Gn, bm_), st, P, -1, in.act);
            gemm<T>(c, nr, bm_, bi, OP_N, R(G0, nr), opV, Vb, W(T2, nr), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
            gemm<T>(c, nr, bm_, bm_, OP_N, R(T2, nr), OP_N, R(Gn, bm_), W(G0n, nr), 1.0, none<T>(), 0.0, nullptr,
                    P, -1, in.act);
            std::swap(G, Gn);
            std::swap(G0, G0n);
        }
        const int b1 = in.bsz[0];
        // closing lead layer 0: V = Hld (nl x b1); (V G) V^+  (lesser.f90:757-771)
        gemm<T>(c, nl, b1, b1, OP_N, in.Hld, OP_N, R(G, b1), W(T1, nl), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        shift_sub<T>(c, nl, zd, in.Haml, in.sl, W(X, nl), P, in.act);
        gemm<T>(c, nl, nl, b1, OP_N, R(T1, nl), OP_C, in.Hld, W(Aw, nl), -1.0, R(X, nl), 1.0, nullptr, P, -1, in.act);
        BMatW<T> gL = out.G_L.base ? out.G_L : W(Gn, nl);
        inverse<T>(c, nl, W(Aw, nl), none<T>(), nullptr, gL, st, P, -1, in.act);
        gemm<T>(c, nr, nl, b1, OP_N, R(G0, nr), OP_C, in.Hld, W(T2, nr), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        gemm<T>(c, nr, nl, nl, OP_N, R(T2, nr), OP_N, ro(gL), gdrl, 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        // Caroli T_RL = Re Tr[(G_RL gml)(G_RL^+ gmr)]
        gemm<T>(c, nr, nl, nl, OP_N, ro(gdrl), OP_N, in.gml, W(X, nr), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        gemm<T>(c, nl, nr, nr, OP_C, ro(gdrl), OP_N, in.gmr, W(Y, nl), 1.0, none<T>(), 0.0, nullptr, P, -1, in.act);
        trace_prod<T>(c, nr, nl, R(X, nr), R(Y, nl), out.T_RL, P, in.act);
    }
    if (two) {
        hipcheck(c, hipEventRecord(c->join, c->aux), "event join");
        c->stream = main_stream;
        hipcheck(c, hipStreamWaitEvent(main_stream, c->join, 0), "wait join");
    }
    // merge the right-to-left statuses (first error of a point wins)
    hipLaunchKernelGGL(k_merge_status, dim3(nblk(P)), dim3(256), 0, main_stream, st_b, out.status, P);
    hipcheck(c, hipGetLastError(), "merge status");
    (void)gLo;
}

// ------------------------------------------------------- Buttiker probes
// Probe outputs of a batched solve (negf/lesser.f90:1199-1503).

// complex traces kept per point by buttiker_run
inline long probe_acc_n(int L) { return (long)L * L + 5 * (long)L + 6; }
// slots (of the RGF slot size, per point) buttiker_run uses after the sweep
inline long probe_slots(int L) { return 7 * (long)L + 2; }

// Device DOS: kept blocks (L) + P_i, Q_i (L - 1 each) + one product slot.
inline long ldos_slots(int L) { return 3 * (long)L; }

// ldos[p][i][j] = -Im G_ii(j, j) / pi of every layer i (tb/device.f90:845-850)
__global__ void k_ldos_diag(const double2* __restrict__ G, long mm, int P, int L, int b, double* __restrict__ ldos)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)P * L * b) return;
    const int j = (int)(e % b);
    const long pi = e / b;
    const int i = (int)(pi % L), p = (int)(pi / L);
    const double2 g = G[((long)i * P + p) * mm + (long)j * b + j];
    ldos[e] = -g.y / 3.14159265358979323846;
}

// ddos[p] = sum of the point's L b entries in index order (the reference's
// sum(sum(ldos, dim=1)) of tb/device.f90:853 up to the order of the sum)
__global__ void k_ldos_sum(const double* __restrict__ ldos, int P, long n, double* __restrict__ ddos)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double acc = 0.0;
    for (long e = 0; e < n; ++e) acc += ldos[(long)p * n + e];
    ddos[p] = acc;
}

// Layer-resolved device DOS of get_ldos (tb/device.f90:610-855) for a batch.
// On entry keep[i] holds the left-connected G^L_i of layers i < L - 1 and the
// full G of the last layer (rgf_run with in.keep).  The backward recursion
// G_ii = G^L_i + (G^L_i V_i) G_{i+1,i+1} (V_i^+ G^L_i) replaces them by the
// diagonal blocks of the full G; it equals the reference's
// (I - G^L_i V_i G^R_{i+1} V_i^+)^-1 G^L_i with G^R from its right-to-left
// sweep (both are the (i, i) block of the same inverse).
void ldos_run(nt_ctx* c, const RgfInput<double>& in, int P, long mm, double* work, double* ldos, double* ddos)
{
    typedef double T;
    const int L = in.L, b = in.bsz[0];
    const long slot = mm * P, bb = (long)b * b;
    T* keep = in.keep;
    auto at = [&](T* base, long k) { return base + 2 * k * slot; };
    auto R = [&](T* p) { return bm<T>(p, mm, b); };
    auto W = [&](T* p) { return bw(p, mm, b); };
    auto RG = [&](T* p) { return bm<T>(p, mm, b, nullptr, slot); };
    auto WG = [&](T* p) { return bw(p, mm, b, nullptr, slot); };
    hipStream_t s = c->stream;
    if (L > 1) {
        T* Pm = work;
        T* Qm = at(work, L - 1);
        T* Dn = at(work, 2 * (L - 1));
        BMat<T> Vg = in.Hdv[0];
        Vg.gstride = bb;      // grouped over layers
        gemm<T>(c, b, b, b, OP_N, RG(keep), OP_N, Vg, WG(Pm), 1.0, none<T>(), 0.0, nullptr, P, -1, nullptr, L - 1);
        gemm<T>(c, b, b, b, OP_C, Vg, OP_N, RG(keep), WG(Qm), 1.0, none<T>(), 0.0, nullptr, P, -1, nullptr, L - 1);
        for (int i = L - 2; i >= 0; --i) {
            gemm<T>(c, b, b, b, OP_N, R(at(keep, i + 1)), OP_N, R(at(Qm, i)), W(Dn), 1.0, none<T>(), 0.0, nullptr, P);
            gemm<T>(c, b, b, b, OP_N, R(at(Pm, i)), OP_N, R(Dn), W(at(keep, i)), 1.0, R(at(keep, i)), 1.0, nullptr,
                    P);
        }
    }
    // ldos goes to the caller's buffer, or (ddos alone) to the first work slot
    double* out = ldos ? ldos : work;
    const long n = (long)L * b;
    if (ldos) check_range(c, ldos, sizeof(double) * n * P, "ldos");
    if (ddos) check_range(c, ddos, sizeof(double) * P, "ddos");
    audit<T>(c, RG(keep), b, b, P, L, "ldos G");
    hipLaunchKernelGGL(k_ldos_diag, dim3(nblk(n * P)), dim3(256), 0, s, reinterpret_cast<const double2*>(keep), mm, P,
                       L, b, out);
    if (ddos) hipLaunchKernelGGL(k_ldos_sum, dim3(nblk(P)), dim3(256), 0, s, out, P, n, ddos);
    hipcheck(c, hipGetLastError(), "ldos");
}

// From the kept blocks of the left-to-right sweep (left-connected G^L_i of
// layers 0..L-2, full G_{N,N} of the last layer) build every block of the full
// retarded function and reduce it to the traces the probe currents need:
//   G_{i,i}   = G^L_i + P_i G_{i+1,i},  P_i = G^L_i V_i,  G_{i+1,i} = G_{i+1,i+1} Q_i,
//               Q_i = V_i^+ G^L_i         (negf/lesser.f90:724-749)
//   G_{i,j}   = P_i G_{i+1,j}   (i < j),   G_{i,j} = G_{i,j+1} Q_j   (i > j),
// each distance t = |i - j| one grouped GEMM over all pairs and points.
// acc (per point, complex): |G_ij|_F^2 at i L + j; Tr(G_i0 gf G_i0^+),
// Tr(G_iN gl G_iN^+), Tr(gf G_0j G_0j^+), Tr(gl G_Nj G_Nj^+), Tr G_ii (L each);
// Tr(gf G_00), Tr(gl G_NN), Tr(gf G_00 gf G_00^+), Tr(gf G_0N gl G_0N^+),
// Tr(gl G_N0 gf G_N0^+), Tr(gl G_NN gl G_NN^+), with gf, gl the broadenings
// of the two contacts at the end layers.
void buttiker_run(nt_ctx* c, int nl, int nr, const RgfInput<double>& in, int P, long mm, double* work,
                  double* acc)
{
    typedef double T;
    const int L = in.L, b = in.bsz[0];
    const long slot = mm * P, bb = (long)b * b;
    const long as = probe_acc_n(L);
    long next = 0;
    auto take = [&](long n) {
        T* p = work + 2 * next * slot;
        next += n;
        return p;
    };
    T* keep = in.keep;
    T* Pm = take(L - 1);
    T* Qm = take(L - 1);
    T* U[2] = {take(L - 1), take(L - 1)};
    T* D[2] = {take(L - 1), take(L - 1)};
    T* gf = take(1);
    T* gl = take(1);
    T* X = take(1);
    T* Xp = take(1);
    T* Y = take(1);
    T* Z = take(1);
    T* S = take(1);
    (void)take(1);
    auto at = [&](T* base, long k) { return base + 2 * k * slot; };
    auto R = [&](T* p) { return bm<T>(p, mm, b); };
    auto W = [&](T* p) { return bw(p, mm, b); };
    auto RG = [&](T* p) { return bm<T>(p, mm, b, nullptr, slot); };      // grouped over slots
    auto WG = [&](T* p) { return bw(p, mm, b, nullptr, slot); };
    BMat<T> Vg = in.Hdv[0];
    Vg.gstride = bb;                                                        // grouped over layers
    hipStream_t s = c->stream;
    auto frob = [&](BMat<T> A, BMat<T> B2, long off, long ogs, int groups) {
        audit<T>(c, A, b, b, P, groups, "frob X");
        audit<T>(c, B2, b, b, P, groups, "frob Y");
        check_range(c, acc + 2 * off, sizeof(T) * 2 * ((P - 1) * as + (long)(groups - 1) * ogs + 1), "frob out");
        launch(c, NT_K_ELEM, 0, "frob",
               [&] { return launch_frob_dot<T>(b, b, A, B2, acc + 2 * off, as, ogs, P, groups, s); });
    };
    auto trace = [&](BMat<T> A, long off, long ogs, int groups) {
        audit<T>(c, A, b, b, P, groups, "trace X");
        check_range(c, acc + 2 * off, sizeof(T) * 2 * ((P - 1) * as + (long)(groups - 1) * ogs + 1), "trace out");
        launch(c, NT_K_ELEM, 0, "trace", [&] { return launch_trace_diag<T>(b, A, acc + 2 * off, as, ogs, P, groups, s); });
    };
    const long oN = (long)L * L, ocL = oN, ocR = oN + L, om = oN + 2 * L, orr = oN + 3 * L, otr = oN + 4 * L,
               omisc = oN + 5 * L;
    // contact broadenings at the end layers: gf = i(S - S^+), S = Hld^+ gl Hld;
    // gl likewise from Hrd^+ gr Hrd (negf/lesser.f90:1288-1292, 1385-1390)
    gemm<T>(c, b, nl, nl, OP_C, in.Hld, OP_N, in.gl, bw(X, mm, b), 1.0, none<T>(), 0.0, nullptr, P);
    gemm<T>(c, b, b, nl, OP_N, R(X), OP_N, in.Hld, W(S), 1.0, none<T>(), 0.0, nullptr, P);
    gamma_op<T>(c, b, R(S), W(gf), P);
    gemm<T>(c, b, nr, nr, OP_C, in.Hrd, OP_N, in.gr, bw(X, mm, b), 1.0, none<T>(), 0.0, nullptr, P);
    gemm<T>(c, b, b, nr, OP_N, R(X), OP_N, in.Hrd, W(S), 1.0, none<T>(), 0.0, nullptr, P);
    gamma_op<T>(c, b, R(S), W(gl), P);
    // P_i = G^L_i V_i and Q_i = V_i^+ G^L_i, all layers at once
    gemm<T>(c, b, b, b, OP_N, RG(keep), OP_N, Vg, WG(Pm), 1.0, none<T>(), 0.0, nullptr, P, -1, nullptr, L - 1);
    gemm<T>(c, b, b, b, OP_C, Vg, OP_N, RG(keep), WG(Qm), 1.0, none<T>(), 0.0, nullptr, P, -1, nullptr, L - 1);
    // backward: G_{i+1,i} (distance-1 lower blocks) and G_{i,i}
    for (int i = L - 2; i >= 0; --i) {
        T* Dn = at(D[1], i);
        gemm<T>(c, b, b, b, OP_N, R(at(keep, i + 1)), OP_N, R(at(Qm, i)), W(Dn), 1.0, none<T>(), 0.0, nullptr, P);
        gemm<T>(c, b, b, b, OP_N, R(at(Pm, i)), OP_N, R(Dn), W(at(keep, i)), 1.0, R(at(keep, i)), 1.0, nullptr, P);
    }
    // distance 0: |G_ii|^2, Tr G_ii and the end-layer contact products
    frob(RG(keep), RG(keep), 0, L + 1, L);
    trace(RG(keep), otr, 1, L);
    const int N = L - 1;
    auto right_gf = [&](T* G, T* out) {   // out = G gf
        gemm<T>(c, b, b, b, OP_N, R(G), OP_N, R(gf), W(out), 1.0, none<T>(), 0.0, nullptr, P);
    };
    auto right_gl = [&](T* G, T* out) {
        gemm<T>(c, b, b, b, OP_N, R(G), OP_N, R(gl), W(out), 1.0, none<T>(), 0.0, nullptr, P);
    };
    auto left_gf = [&](T* G, T* out) {
        gemm<T>(c, b, b, b, OP_N, R(gf), OP_N, R(G), W(out), 1.0, none<T>(), 0.0, nullptr, P);
    };
    auto left_gl = [&](T* G, T* out) {
        gemm<T>(c, b, b, b, OP_N, R(gl), OP_N, R(G), W(out), 1.0, none<T>(), 0.0, nullptr, P);
    };
    {
        T* G00 = at(keep, 0);
        T* GNN = at(keep, N);
        right_gf(G00, X);
        left_gf(G00, Y);
        right_gl(GNN, Xp);
        left_gl(GNN, Z);
        frob(R(G00), R(X), ocL, 0, 1);
        frob(R(G00), R(Y), om, 0, 1);
        frob(R(GNN), R(Xp), ocR + N, 0, 1);
        frob(R(GNN), R(Z), orr + N, 0, 1);
        trace(R(Y), omisc + 0, 0, 1);
        trace(R(Z), omisc + 1, 0, 1);
        frob(R(Y), R(X), omisc + 2, 0, 1);
        frob(R(Z), R(Xp), omisc + 5, 0, 1);
    }
    int cur = 1;
    for (int t = 1; t <= N; ++t) {
        const int ng = L - t;
        if (t == 1) {
            // upper distance 1: G_{i,i+1} = P_i G_{i+1,i+1}
            gemm<T>(c, b, b, b, OP_N, RG(Pm), OP_N, RG(at(keep, 1)), WG(U[1]), 1.0, none<T>(), 0.0, nullptr, P, -1,
                    nullptr, ng);
        } else {
            const int prev = cur;
            cur ^= 1;
            gemm<T>(c, b, b, b, OP_N, RG(Pm), OP_N, RG(at(U[prev], 1)), WG(U[cur]), 1.0, none<T>(), 0.0, nullptr, P,
                    -1, nullptr, ng);
            gemm<T>(c, b, b, b, OP_N, RG(at(D[prev], 1)), OP_N, RG(Qm), WG(D[cur]), 1.0, none<T>(), 0.0, nullptr, P,
                    -1, nullptr, ng);
        }
        T* Ut = U[cur];
        T* Dt = D[cur];
        frob(RG(Ut), RG(Ut), t, L + 1, ng);
        frob(RG(Dt), RG(Dt), (long)t * L, L + 1, ng);
        T* G0t = at(Ut, 0);           // G_{0,t}
        T* GcN = at(Ut, N - t);       // G_{N-t,N}
        T* Gt0 = at(Dt, 0);           // G_{t,0}
        T* GNc = at(Dt, N - t);       // G_{N,N-t}
        left_gf(G0t, Y);
        right_gl(GcN, Xp);
        right_gf(Gt0, X);
        left_gl(GNc, Z);
        frob(R(G0t), R(Y), om + t, 0, 1);
        frob(R(GcN), R(Xp), ocR + N - t, 0, 1);
        frob(R(Gt0), R(X), ocL + t, 0, 1);
        frob(R(GNc), R(Z), orr + N - t, 0, 1);
        if (t == N) {
            frob(R(Y), R(Xp), omisc + 3, 0, 1);   // Tr(gf G_0N gl G_0N^+)
            frob(R(Z), R(X), omisc + 4, 0, 1);    // Tr(gl G_N0 gf G_N0^+)
        }
    }
}

// ------------------------------------------------------- tdb batched solve
// adjoint storage is capped so that a chunk of points fits this budget
constexpr double ADJ_BUDGET_BYTES = 24.0 * 1024 * 1024 * 1024;
// so is the mixed-precision solve's (FP32 factors and FP64 columns of both directions)
constexpr double MIXED_BUDGET_BYTES = 64.0 * 1024 * 1024 * 1024;

}  // namespace

void solve_tdb_dev(nt_ctx* c, const nt_params* prm, const TdbArgs& a)
{
    typedef double T;
    const int nl = a.nl, nr = a.nr, b = a.b, L = a.L;
    const long nll = (long)nl * nl, nrr = (long)nr * nr, bb = (long)b * b;
    const bool same = nl == nr;
    const bool pm = a.ex.gl_plus || a.ex.gr_minus;
    const bool adjoint = a.adj_w != nullptr;
    const int nmax = std::max(nl, nr);
    const long nmm = (long)nmax * nmax;
    int bplan = b;       // block size the RGF scratch is planned for (largest of the siblings)
    for (const TdbArgs* sb : a.siblings) bplan = std::max(bplan, sb->b);
    const long mmr = (long)std::max(nmax, b) * std::max(nmax, b);   // RGF slot size
    int Pmax = std::min(a.npts, c->batch_limit);
    const int adjK = adjoint ? adj_interval(c->adj_ckpt, a.adj_lo, L) : 1;
    if (prm->precision == NT_PREC_MIXED && !adjoint && !a.probe) {
        // the mixed solve keeps both directions' FP32 factors and FP64 columns
        const double per_pt = (double)mixed_caroli_scratch(nl, nr, b, L, 1);
        Pmax = std::max(1, std::min(Pmax, (int)std::min(1.0e9, MIXED_BUDGET_BYTES / per_pt)));
    }
    if (adjoint) {
        const double per_pt = (double)adj_slots(a.adj_lo, a.adj_hi, L, adjK) * mmr * sizeof(nt_c128);
        Pmax = std::max(1, std::min(Pmax, (int)std::min(1.0e9, ADJ_BUDGET_BYTES / per_pt)));
    }
    const ProbeSpec* probe = a.probe;
    if (probe) {
        const double per_pt = (double)(probe_slots(L) * mmr + probe_acc_n(L) + 2 * L) * sizeof(nt_c128);
        Pmax = std::max(1, std::min(Pmax, (int)std::min(1.0e9, ADJ_BUDGET_BYTES / per_pt)));
    }
    // device DOS: every layer's G^L kept, then the diagonal blocks of G
    const bool dos = a.ex.ldos || a.ex.ddos;
    if (dos) {
        const double per_pt = (double)ldos_slots(L) * mmr * sizeof(nt_c128);
        Pmax = std::max(1, std::min(Pmax, (int)std::min(1.0e9, ADJ_BUDGET_BYTES / per_pt)));
    }

    // the probe solve and the device DOS run in FP64
    const bool mixed = prm->precision == NT_PREC_MIXED && !probe && !dos;
    const size_t c32 = 2 * sizeof(float);
    const long Ldv = std::max(L - 1, 1);
    const bool sweep32 = mixed && !adjoint;   // the adjoint runs the FP64 sweep
    // ---- plan the workspace for one chunk of Pmax points.  Without
    // allocation (the static scratch) the chunk shrinks until the plan fits.
    Arena ar, pa;
    size_t o_idx, o_zdec, o_zpm, o_zdev, o_stdec, o_itdec, o_strgf, o_stpm, o_itpm, o_lh, o_lv, o_g, o_sig, o_gam, o_gpm, o_glr, o_slr, o_gmlr, o_zdev32, o_acc, o_fb, o_cnt, o_zs, o_zpms, o_idx1, o_sts, o_its, o_scr, o_adj, need_dec, need_rgf, need_adj = 0;
    size_t o_h32[6] = {0, 0, 0, 0, 0, 0};
    size_t o_pzl = 0, o_pgam = 0, o_pdgam = 0, o_pacc = 0, o_pkeep = 0, o_pwork = 0;
    for (;;) {
        ar = Arena();
        o_idx = ar.take(sizeof(int) * 2 * Pmax);
        o_zdec = ar.take(sizeof(double) * 4 * Pmax);
        o_zpm = ar.take(sizeof(double) * 4 * Pmax);
        o_zdev = ar.take(sizeof(double) * 2 * Pmax);
        o_stdec = ar.take(sizeof(int) * 2 * Pmax);
        o_itdec = ar.take(sizeof(int) * 2 * Pmax);
        o_strgf = ar.take(sizeof(int) * Pmax);
        o_stpm = ar.take(sizeof(int) * 2 * Pmax);
        o_itpm = ar.take(sizeof(int) * 2 * Pmax);
        o_lh = ar.take(sizeof(nt_c128) * 2 * nmm * a.nk);
        o_lv = ar.take(sizeof(nt_c128) * 2 * nmm * a.nk);
        o_g = ar.take(sizeof(nt_c128) * 2 * nmm * Pmax);      // [gl_adv, gr_ret] per point
        o_sig = ar.take(sizeof(nt_c128) * 2 * nmm * Pmax);
        o_gam = ar.take(sizeof(nt_c128) * 2 * nmm * Pmax);
        o_gpm = ar.take(sizeof(nt_c128) * 2 * nmm * Pmax);    // [gl_plus, gr_minus]
        o_glr = ar.take(sizeof(nt_c128) * nll * Pmax);        // gl = gl_adv^+
        o_slr = ar.take(sizeof(nt_c128) * nll * Pmax);        // sl = sl_adv^+
        o_gmlr = ar.take(sizeof(nt_c128) * nll * Pmax);       // gml = -gml_adv
        // mixed precision: FP32 copies of the H(k) blocks (once per call) and of
        // the lead quantities and device shifts (per chunk)
        if (mixed) {
            o_h32[0] = ar.take(c32 * bb * L * a.nk);
            o_h32[1] = ar.take(c32 * bb * Ldv * a.nk);
            o_h32[2] = ar.take(c32 * nl * b * a.nk);
            o_h32[3] = ar.take(c32 * nr * b * a.nk);
            o_h32[4] = ar.take(c32 * nll * a.nk);
            o_h32[5] = ar.take(c32 * nrr * a.nk);
        }
        o_zdev32 = ar.take(c32 * Pmax);
        // mixed precision: per-point acceptance, FP64-recompute mask, counters
        o_acc = ar.take(sizeof(int) * Pmax);
        o_fb = ar.take(sizeof(int) * Pmax);
        o_cnt = ar.take(sizeof(int) * 8);
        // per-side shifts, k indices, statuses and iterations (nl != nr only)
        o_zs = same ? 0 : ar.take(sizeof(double) * 4 * Pmax);
        o_zpms = same ? 0 : ar.take(sizeof(double) * 4 * Pmax);
        o_idx1 = same ? 0 : ar.take(sizeof(int) * Pmax);
        o_sts = same ? 0 : ar.take(sizeof(int) * 2 * Pmax);
        o_its = same ? 0 : ar.take(sizeof(int) * 2 * Pmax);
        o_scr = ar.off;
        need_dec = lead_gfs_scratch(prm, nmax, 2 * Pmax, 2 * a.nk);
        // the mixed solve, then (for rejected points) the FP64 sweep, share the region
        need_rgf = sweep32 ? std::max(mixed_caroli_scratch(nl, nr, b, L, Pmax), rgf_scratch<T>(nl, nr, b, Pmax))
                                        : rgf_scratch<T>(nl, nr, bplan, Pmax);
        o_adj = o_scr + std::max(need_dec, need_rgf);
        need_adj = 0;
        if (adjoint) {
            Arena probe;
            (void)adj_carve(probe, a.adj_lo, a.adj_hi, L, adjK, mmr, Pmax, nullptr);
            need_adj = probe.off;
        }
        // probe region: per-layer shifts and broadenings, traces, kept blocks and
        // the work slots of buttiker_run
        pa = Arena();
        pa.off = o_adj + need_adj;
        if (probe) {
            o_pzl = pa.take(sizeof(double) * 2 * L * Pmax);
            o_pgam = pa.take(sizeof(double) * L * Pmax);
            o_pdgam = pa.take(sizeof(double) * L * Pmax);
            o_pacc = pa.take(sizeof(double) * 2 * probe_acc_n(L) * Pmax);
            o_pkeep = pa.take(sizeof(nt_c128) * mmr * L * Pmax);
            o_pwork = pa.take(sizeof(nt_c128) * mmr * (probe_slots(L) - L) * Pmax);
        } else if (dos) {
            o_pkeep = pa.take(sizeof(nt_c128) * mmr * L * Pmax);
            o_pwork = pa.take(sizeof(nt_c128) * mmr * (ldos_slots(L) - L) * Pmax);
        }
        // the static scratch, or a workspace cap of the context (worker
        // contexts of a multi-device call share the device's memory)
        const size_t cap = a.noalloc ? NT_SCRATCH_BYTES : c->ws_cap;
        if (cap == 0 || pa.off <= cap) break;
        if (Pmax == 1) {
            if (!a.noalloc) break;     // one point over a workspace cap: solve it anyway
            char buf[200];
            snprintf(buf, sizeof(buf), "one point needs %zu bytes of workspace, more than the %zu-byte static scratch",
                     pa.off, (size_t)NT_SCRATCH_BYTES);
            c->err = buf;
            throw Fail{NT_ERR_ARG};
        }
        Pmax = std::max(1, std::min(Pmax - 1, (int)((double)Pmax * cap / pa.off)));
    }
    if (a.noalloc) {
        ar.base = static_cast<char*>(scratch_ptr(c));
        if (c->ws_poison) poison_dev(c, ar.base, pa.off, c->stream);
    } else {
        ensure_ws(c, pa.off);
        ar.base = static_cast<char*>(c->ws);
    }
    hipStream_t s = c->stream;

    int* idx = ar.ptr<int>(o_idx);
    T* zdec = ar.ptr<T>(o_zdec);
    T* zpm = ar.ptr<T>(o_zpm);
    T* zdev = ar.ptr<T>(o_zdev);
    int* stdec = ar.ptr<int>(o_stdec);
    int* itdec = ar.ptr<int>(o_itdec);
    int* strgf = ar.ptr<int>(o_strgf);
    int* stpm = ar.ptr<int>(o_stpm);
    int* itpm = ar.ptr<int>(o_itpm);
    T* lh = ar.ptr<T>(o_lh);
    T* lv = ar.ptr<T>(o_lv);
    T* gbuf = ar.ptr<T>(o_g);
    T* sigbuf = ar.ptr<T>(o_sig);
    T* gambuf = ar.ptr<T>(o_gam);
    T* gpmbuf = ar.ptr<T>(o_gpm);
    T* glr = ar.ptr<T>(o_glr);
    T* slr = ar.ptr<T>(o_slr);
    T* gmlr = ar.ptr<T>(o_gmlr);

    const T* Haml = reinterpret_cast<const T*>(a.Haml);
    const T* Hamlv = reinterpret_cast<const T*>(a.Hamlv);
    const T* Hamr = reinterpret_cast<const T*>(a.Hamr);
    const T* Hamrv = reinterpret_cast<const T*>(a.Hamrv);
    const T* Hamd = reinterpret_cast<const T*>(a.Hamd);
    const T* Hamdv = reinterpret_cast<const T*>(a.Hamdv);
    const T* Hamld = reinterpret_cast<const T*>(a.Hamld);
    const T* Hamrd = reinterpret_cast<const T*>(a.Hamrd);

    if (same) {
        // lead stack [left k-points..., right k-points...] so one decimation
        // batch serves both leads
        const long n_a = nll * a.nk;
        check_range(c, Haml, sizeof(T) * 2 * n_a, "Haml");
        check_range(c, Hamr, sizeof(T) * 2 * n_a, "Hamr");
        check_range(c, Hamlv, sizeof(T) * 2 * n_a, "Hamlv");
        check_range(c, Hamrv, sizeof(T) * 2 * n_a, "Hamrv");
        hipLaunchKernelGGL(k_copy_lead_stack, dim3(nblk(n_a)), dim3(256), 0, s,
                           reinterpret_cast<const double2*>(Haml), reinterpret_cast<const double2*>(Hamr),
                           reinterpret_cast<double2*>(lh), n_a);
        hipLaunchKernelGGL(k_copy_lead_stack, dim3(nblk(n_a)), dim3(256), 0, s,
                           reinterpret_cast<const double2*>(Hamlv), reinterpret_cast<const double2*>(Hamrv),
                           reinterpret_cast<double2*>(lv), n_a);
        hipcheck(c, hipGetLastError(), "lead stack");
    }

    float* H32[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (mixed) {
        for (int i = 0; i < 6; ++i) H32[i] = ar.ptr<float>(o_h32[i]);
        const float one = 1.0f;
        convert_op<T, float>(c, b, b, bm<T>(Hamd, bb, b), bw<float>(H32[0], bb, b), 0, one, L * a.nk);
        if (L > 1) convert_op<T, float>(c, b, b, bm<T>(Hamdv, bb, b), bw<float>(H32[1], bb, b), 0, one, Ldv * a.nk);
        convert_op<T, float>(c, nl, b, bm<T>(Hamld, (long)nl * b, nl), bw<float>(H32[2], (long)nl * b, nl), 0, one,
                             a.nk);
        convert_op<T, float>(c, nr, b, bm<T>(Hamrd, (long)nr * b, nr), bw<float>(H32[3], (long)nr * b, nr), 0, one,
                             a.nk);
        convert_op<T, float>(c, nl, nl, bm<T>(Haml, nll, nl), bw<float>(H32[4], nll, nl), 0, one, a.nk);
        convert_op<T, float>(c, nr, nr, bm<T>(Hamr, nrr, nr), bw<float>(H32[5], nrr, nr), 0, one, a.nk);
    }
    float* zdev32 = ar.ptr<float>(o_zdev32);

    double iters_sum = 0.0;
    std::vector<int> hit;
    for (int p0 = 0; p0 < a.npts; p0 += Pmax) {
        const int P = std::min(Pmax, a.npts - p0);
        const double* E = a.E + p0;
        const int* kidx = a.kidx ? a.kidx + p0 : nullptr;
        hipLaunchKernelGGL(k_lead_setup, dim3(nblk(P)), dim3(256), 0, s, E, kidx, a.nk, P, 1, prm->eta,
                           prm->eta0, prm->eta_mode, prm->device_eta_factor, idx, zdec, pm ? zpm : nullptr, zdev);
        hipcheck(c, hipGetLastError(), "lead setup");
        if (probe) {
            check_range(c, probe->temp, sizeof(double) * L, "probe temperatures");
            launch(c, NT_K_ELEM, 0, "probe shifts", [&] {
                return launch_probe_shifts(E, zdev, probe->temp, L, P, probe->B, probe->C, probe->lifetime,
                                           ar.ptr<T>(o_pzl), ar.ptr<T>(o_pgam), ar.ptr<T>(o_pdgam), s);
            });
        }
        zero_dev(c, stdec, sizeof(int) * 2 * P, s);
        zero_dev(c, strgf, sizeof(int) * P, s);
        zero_dev(c, stpm, sizeof(int) * 2 * P, s);

        // ---- surface Green's functions (FP64)
        if (same) {
            const long n2 = nll;
            DecimationOut<T> o{bw(gbuf, n2, nl), bw(sigbuf, n2, nl), bw(gambuf, n2, nl)};
            lead_gfs(c, prm, nl, 2 * P, 2 * a.nk, lh, lv, idx, 1, zdec, o, stdec, itdec, ar, o_scr);
            if (pm) {
                DecimationOut<T> o2{bw(gpmbuf, n2, nl), nonew<T>(), nonew<T>()};
                lead_gfs(c, prm, nl, 2 * P, 2 * a.nk, lh, lv, idx, 0, zpm, o2, stpm, itpm, ar, o_scr);
            }
        } else {
            // one batch per lead (nl != nr): per-point slots of nmm keep the
            // [left, right] pairs; the left and right matrices of k-point k
            // are Haml/Hamlv[k] and Hamr/Hamrv[k]
            double* zs = ar.ptr<T>(o_zs);
            double* zpms = ar.ptr<T>(o_zpms);
            int* idx1 = ar.ptr<int>(o_idx1);
            int* sts = ar.ptr<int>(o_sts);
            int* its = ar.ptr<int>(o_its);
            hipLaunchKernelGGL(k_lead_split, dim3(nblk(P)), dim3(256), 0, s, zdec, pm ? zpm : nullptr, idx, P, zs,
                               pm ? zpms : nullptr, idx1);
            hipcheck(c, hipGetLastError(), "lead split");
            auto sides = [&](T* g, T* sg, T* gm, const T* zz, int adag, int* st2, int* it2) {
                zero_dev(c, sts, sizeof(int) * 2 * P, s);
                DecimationOut<T> ol{bw(g, 2 * nmm, nl), sg ? bw(sg, 2 * nmm, nl) : nonew<T>(),
                                    gm ? bw(gm, 2 * nmm, nl) : nonew<T>()};
                lead_gfs(c, prm, nl, P, a.nk, Haml, Hamlv, idx1, adag, zz, ol, sts, its, ar, o_scr);
                DecimationOut<T> orr{bw(g + 2 * nmm, 2 * nmm, nr), sg ? bw(sg + 2 * nmm, 2 * nmm, nr) : nonew<T>(),
                                     gm ? bw(gm + 2 * nmm, 2 * nmm, nr) : nonew<T>()};
                lead_gfs(c, prm, nr, P, a.nk, Hamr, Hamrv, idx1, adag, zz + 2 * P, orr, sts + P, its + P, ar, o_scr);
                hipLaunchKernelGGL(k_lead_merge, dim3(nblk(P)), dim3(256), 0, s, sts, its, P, st2, it2);
                hipcheck(c, hipGetLastError(), "lead merge");
            };
            sides(gbuf, sigbuf, gambuf, zs, 1, stdec, itdec);
            if (pm) sides(gpmbuf, nullptr, nullptr, zpms, 0, stpm, itpm);
        }
        // ---- adv -> ret (lesser.f90:141-144): gl = gl_adv^+, sl = sl_adv^+, gml = -gml_adv
        copy_op<T>(c, nl, nl, bm<T>(gbuf, 2 * nmm, nl), bw(glr, nll, nl), 1, (T)1, P);
        copy_op<T>(c, nl, nl, bm<T>(sigbuf, 2 * nmm, nl), bw(slr, nll, nl), 1, (T)1, P);
        copy_op<T>(c, nl, nl, bm<T>(gambuf, 2 * nmm, nl), bw(gmlr, nll, nl), 0, (T)-1, P);
        // ---- RGF + Caroli
        // Per-k operands: item p reads k-point kidx[p]; without a k map there is
        // one k-point and every item shares it (stride 0).
        auto kst = [&](long per_k) { return kidx ? per_k : 0L; };
        const int kn = a.nk;
        RgfInput<T> in;
        in.L = L;
        in.bsz.assign(L, b);
        for (int i = 0; i < L; ++i) in.Hd.push_back(bm<T>(Hamd + 2 * i * bb, kst((long)L * bb), b, kidx, 0, kn));
        for (int i = 0; i + 1 < L; ++i)
            in.Hdv.push_back(bm<T>(Hamdv + 2 * i * bb, kst((long)std::max(L - 1, 1) * bb), b, kidx, 0, kn));
        in.Hld = bm<T>(Hamld, kst((long)nl * b), nl, kidx, 0, kn);
        in.Hrd = bm<T>(Hamrd, kst((long)nr * b), nr, kidx, 0, kn);
        in.Hrd_fwd = none<T>();
        in.Haml = bm<T>(Haml, kst(nll), nl, kidx, 0, kn);
        in.Hamr = bm<T>(Hamr, kst(nrr), nr, kidx, 0, kn);
        in.gl = bm<T>(glr, nll, nl);
        in.sl = bm<T>(slr, nll, nl);
        in.gml = bm<T>(gmlr, nll, nl);
        in.gr = bm<T>(gbuf + 2 * nmm, 2 * nmm, nr);
        in.sr = bm<T>(sigbuf + 2 * nmm, 2 * nmm, nr);
        in.gmr = bm<T>(gambuf + 2 * nmm, 2 * nmm, nr);
        in.zlead = zdec + 2;          // right retarded E + i eta(E) of each point (stride 4)
        in.zlead_stride = 4;
        if (probe) {
            in.zl = ar.ptr<T>(o_pzl);
            in.keep = ar.ptr<T>(o_pkeep);
        } else if (dos) {
            in.keep = ar.ptr<T>(o_pkeep);
        }
        RgfOutput<T> out;
        out.T_LR = a.T_LR + p0;
        out.T_RL = a.T_RL + p0;
        out.status = strgf;
        auto exw = [&](nt_c128* base, int n1, int n2) {
            return base ? bw(reinterpret_cast<T*>(base) + 2 * (long)p0 * n1 * n2, (long)n1 * n2, n1) : nonew<T>();
        };
        out.G_D_LR = exw(a.ex.G_D_LR, nl, nr);
        out.G_D_RL = exw(a.ex.G_D_RL, nr, nl);
        out.G_L = exw(a.ex.G_L, nl, nl);
        out.G_R = exw(a.ex.G_R, nr, nr);
        if (!sweep32) {
            AdjSpec adj{};
            if (adjoint) {
                Arena aa = ar;
                aa.off = o_adj;
                adj = adj_carve(aa, a.adj_lo, a.adj_hi, L, adjK, mmr, P, a.adj_w + p0);
            }
            rgf_run<T>(c, nl, nr, in, zdev, P, out, ar, o_scr, adjoint ? &adj : nullptr);
            if (dos)
                ldos_run(c, in, P, mmr, ar.ptr<T>(o_pwork), a.ex.ldos ? a.ex.ldos + (long)p0 * L * b : nullptr,
                         a.ex.ddos ? a.ex.ddos + p0 : nullptr);
            if (probe) {
                buttiker_run(c, nl, nr, in, P, mmr, ar.ptr<T>(o_pwork), ar.ptr<T>(o_pacc));
                ProbeCombineArgs pc;
                pc.L = L;
                pc.P = P;
                pc.p0 = p0;
                pc.E = E;
                pc.temp = probe->temp;
                pc.t_left = probe->t_left;
                pc.t_right = probe->t_right;
                pc.gam = ar.ptr<T>(o_pgam);
                pc.dgam = ar.ptr<T>(o_pdgam);
                pc.acc = ar.ptr<T>(o_pacc);
                pc.acc_stride = probe_acc_n(L);
                pc.first_mask = a.first_mask ? a.first_mask + p0 : nullptr;
                pc.flux = probe->flux;
                pc.jac = probe->jac;
                pc.flux_left = probe->flux_left;
                pc.flux_right = probe->flux_right;
                pc.a_sum = probe->a_sum;
                pc.gn_sum = probe->gn_sum;
                check_range(c, probe->flux + (long)p0 * L, sizeof(double) * L * P, "probe flux");
                check_range(c, probe->jac + (long)p0 * L * L, sizeof(double) * L * L * P, "probe jacobian");
                check_range(c, probe->a_sum + (long)p0 * L, sizeof(double) * L * P, "probe A");
                check_range(c, probe->gn_sum + (long)p0 * L, sizeof(double) * L * P, "probe Gn");
                check_range(c, probe->flux_left + p0, sizeof(double) * P, "flux_left");
                check_range(c, probe->flux_right + p0, sizeof(double) * P, "flux_right");
                launch(c, NT_K_ELEM, 0, "probe combine", [&] { return launch_probe_combine(pc, s); });
            }
            if (adjoint) {
                const long ne = bb;
                const int nw = a.adj_hi - a.adj_lo + 1, nc = adj_hc(a.adj_hi, L) - a.adj_lo + 1;
                const int* fm = a.first_mask ? a.first_mask + p0 : nullptr;
                check_range(c, a.dHd, sizeof(nt_c128) * ne * nw * a.nk, "dHd");
                hipLaunchKernelGGL(k_adj_reduce, dim3(nblk(ne), nw), dim3(256), 0, s,
                                   reinterpret_cast<const double2*>(adj.Zd), mmr, P, ne, kidx, stdec, strgf, fm,
                                   reinterpret_cast<double2*>(a.dHd), nw);
                if (nc > 0 && a.dHdv) {
                    check_range(c, a.dHdv, sizeof(nt_c128) * ne * nc * a.nk, "dHdv");
                    hipLaunchKernelGGL(k_adj_reduce, dim3(nblk(ne), nc), dim3(256), 0, s,
                                       reinterpret_cast<const double2*>(adj.Zc), mmr, P, ne, kidx, stdec, strgf, fm,
                                       reinterpret_cast<double2*>(a.dHdv), nc);
                }
                hipcheck(c, hipGetLastError(), "adjoint reduce");
            }
        } else {
            // Mixed precision: the surface GFs above stay FP64 (decimation);
            // each sweep direction is factorised in complex FP32 on MFMA and
            // its block column refined against FP64 residuals (nt_mixed.hip).
            // Points whose refinement does not pass the error estimate are
            // recomputed by the FP64 sweep, masked to those points.
            const float one = 1.0f;
            convert_op<T, float>(c, 1, 1, bm<T>(zdev, 1, 1), bw<float>(zdev32, 1, 1), 0, one, P);
            RgfInput<float> i32;
            i32.L = L;
            i32.bsz.assign(L, b);
            for (int i = 0; i < L; ++i)
                i32.Hd.push_back(bm<float>(H32[0] + 2 * i * bb, kst((long)L * bb), b, kidx, 0, kn));
            for (int i = 0; i + 1 < L; ++i)
                i32.Hdv.push_back(bm<float>(H32[1] + 2 * i * bb, kst(Ldv * bb), b, kidx, 0, kn));
            i32.Hld = bm<float>(H32[2], kst((long)nl * b), nl, kidx, 0, kn);
            i32.Hrd = bm<float>(H32[3], kst((long)nr * b), nr, kidx, 0, kn);
            i32.Hrd_fwd = none<float>();
            int* acc = ar.ptr<int>(o_acc);
            int* fb = ar.ptr<int>(o_fb);
            int* cnt = ar.ptr<int>(o_cnt);
            mixed_caroli(c, nl, nr, in, i32, zdev, zdev32, P, out, ar, o_scr, acc, cnt);
            zero_dev(c, cnt + 6, sizeof(int) * 2, s);
            hipLaunchKernelGGL(k_reject_mask, dim3(nblk(P)), dim3(256), 0, s, acc, fb, cnt + 6, P);
            hipcheck(c, hipGetLastError(), "reject mask");
            int hc[8];
            hipcheck(c, hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s), "memcpy");
            hipcheck(c, hipStreamSynchronize(s), "sync");
            c->mixed_counts[0] += P;
            c->mixed_counts[1] += hc[1] + hc[4];
            c->mixed_counts[2] += hc[6];
            if (hc[6] > 0) {
                RgfInput<T> inf = in;
                inf.act = fb;
                rgf_run<T>(c, nl, nr, inf, zdev, P, out, ar, o_scr);
            }
        }

        // ---- optional lead outputs
        auto put = [&](nt_c128* dst, const T* srcbase, long stride, int n) {
            if (!dst) return;
            copy_op<T>(c, n, n, bm<T>(srcbase, stride, n),
                       bw(reinterpret_cast<T*>(dst) + 2 * (long)p0 * n * n, (long)n * n, n), 0, (T)1, P);
        };
        put(a.ex.gl_adv, gbuf, 2 * nmm, nl);
        put(a.ex.gr_ret, gbuf + 2 * nmm, 2 * nmm, nr);
        put(a.ex.sl_adv, sigbuf, 2 * nmm, nl);
        put(a.ex.sr_ret, sigbuf + 2 * nmm, 2 * nmm, nr);
        put(a.ex.gml_adv, gambuf, 2 * nmm, nl);
        put(a.ex.gmr_ret, gambuf + 2 * nmm, 2 * nmm, nr);
        if (pm) {
            put(a.ex.gl_plus, gpmbuf, 2 * nmm, nl);
            put(a.ex.gr_minus, gpmbuf + 2 * nmm, 2 * nmm, nr);
        }
        if (a.ex.sdos_l) check_range(c, a.ex.sdos_l + p0, sizeof(double) * P, "sdos_l");
        if (a.ex.sdos_r) check_range(c, a.ex.sdos_r + p0, sizeof(double) * P, "sdos_r");
        // surface DOS: negf +-2 sqrt(E) Im Tr g / pi (negf/surface.f90:608-630),
        // tb +-Im Tr g / pi (tb/surface.f90:345-366); by default the driver
        // family the eta convention belongs to
        const bool tb_dos = prm->dos_mode == NT_DOS_TB ||
                            (prm->dos_mode == NT_DOS_FROM_ETA && prm->eta_mode == NT_ETA_ABSOLUTE);
        const double* Edos = tb_dos ? nullptr : E;
        const double cdos = tb_dos ? 1.0 : 2.0;
        if (a.ex.sdos_l)
            launch(c, NT_K_ELEM, 0, "sdos", [&] {
                return launch_trace_im<T>(nl, bm<T>(gbuf, 2 * nmm, nl), Edos, cdos, a.ex.sdos_l + p0, P, s);
            });
        if (a.ex.sdos_r)
            launch(c, NT_K_ELEM, 0, "sdos", [&] {
                return launch_trace_im<T>(nr, bm<T>(gbuf + 2 * nmm, 2 * nmm, nr), Edos, -cdos, a.ex.sdos_r + p0, P,
                                          s);
            });
        // ---- status, iterations, clamps
        hipLaunchKernelGGL(k_point_status, dim3(nblk(P)), dim3(256), 0, s, stdec, itdec, strgf, P,
                           a.status ? a.status + p0 : nullptr, a.iters ? a.iters + p0 : nullptr,
                           a.first_mask ? a.first_mask + p0 : nullptr, a.T_LR + p0, a.T_RL + p0);
        hipcheck(c, hipGetLastError(), "point status");
        // ---- devices sharing these leads: their own sweeps on the same lead GFs
        for (const TdbArgs* sb : a.siblings) {
            const int Ls = sb->L, bs = sb->b;
            const long bbs = (long)bs * bs;
            const T* Hd_s = reinterpret_cast<const T*>(sb->Hamd);
            const T* Hdv_s = reinterpret_cast<const T*>(sb->Hamdv);
            RgfInput<T> ins = in;
            ins.L = Ls;
            ins.bsz.assign(Ls, bs);
            ins.Hd.clear();
            ins.Hdv.clear();
            for (int i = 0; i < Ls; ++i) ins.Hd.push_back(bm<T>(Hd_s + 2 * i * bbs, kst((long)Ls * bbs), bs, kidx, 0, kn));
            for (int i = 0; i + 1 < Ls; ++i)
                ins.Hdv.push_back(bm<T>(Hdv_s + 2 * i * bbs, kst((long)std::max(Ls - 1, 1) * bbs), bs, kidx, 0, kn));
            ins.Hld = bm<T>(reinterpret_cast<const T*>(sb->Hamld), kst((long)nl * bs), nl, kidx, 0, kn);
            ins.Hrd = bm<T>(reinterpret_cast<const T*>(sb->Hamrd), kst((long)nr * bs), nr, kidx, 0, kn);
            RgfOutput<T> outs;
            outs.T_LR = sb->T_LR + p0;
            outs.T_RL = sb->T_RL + p0;
            outs.status = strgf;
            outs.G_D_LR = outs.G_D_RL = outs.G_L = outs.G_R = nonew<T>();
            zero_dev(c, strgf, sizeof(int) * P, s);
            rgf_run<T>(c, nl, nr, ins, zdev, P, outs, ar, o_scr);
            check_range(c, sb->T_LR + p0, sizeof(double) * P, "sibling T_LR");
            check_range(c, sb->T_RL + p0, sizeof(double) * P, "sibling T_RL");
            if (sb->status) check_range(c, sb->status + p0, sizeof(int) * P, "sibling status");
            if (sb->first_mask) check_range(c, sb->first_mask + p0, sizeof(int) * P, "sibling mask");
            hipLaunchKernelGGL(k_point_status, dim3(nblk(P)), dim3(256), 0, s, stdec, itdec, strgf, P,
                               sb->status ? sb->status + p0 : nullptr, nullptr,
                               sb->first_mask ? sb->first_mask + p0 : nullptr, sb->T_LR + p0, sb->T_RL + p0);
            hipcheck(c, hipGetLastError(), "sibling point status");
        }
        // mean iterations (host side, for the flop model); not read when the
        // solve must run without host synchronisation
        if (c->dec_check > 0) {
            hit.resize(2 * P);
            hipcheck(c, hipMemcpyAsync(hit.data(), itdec, sizeof(int) * 2 * P, hipMemcpyDeviceToHost, s), "memcpy");
            hipcheck(c, hipStreamSynchronize(s), "sync");
            for (int q = 0; q < 2 * P; ++q) iters_sum += hit[q];
        }
    }
    c->last_mean_iters = c->dec_check > 0 ? iters_sum / std::max(1, 2 * a.npts) : -1.0;
}

// =================================================================== C ABI
// ------------------------------------------------------------- autotuner
// Deterministic diagonally dominant test matrices: entry hash in [-0.5, 0.5)
// plus n on the diagonal.
template <typename T>
__global__ void k_fill_tune(T* A, int n, int batch)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n * n * batch;
    if (e >= total) return;
    unsigned long long h = (unsigned long long)e * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    h ^= h >> 31;
    h *= 0xD6E8FEB86659FD93ull;
    h ^= h >> 29;
    const double u = (double)(h & 0xFFFFF) / (double)0x100000 - 0.5;
    const double v = (double)((h >> 20) & 0xFFFFF) / (double)0x100000 - 0.5;
    const long m = e % ((long)n * n);
    const bool diag = (m % n) == (m / n);
    A[2 * e] = (T)(u + (diag ? n : 0));
    A[2 * e + 1] = (T)v;
}

// Times every panel width the inverse supports for n x n matrices in batches
// of `batch` and records the fastest for this context.
template <typename T>
void tune_inverse(nt_ctx* c, int n, int batch, int* best_nb, double* best_ms)
{
    const long nn = (long)n * n;
    DevBuf A(c, sizeof(T) * 2 * nn * batch), W(c, sizeof(T) * 2 * nn * batch), X(c, sizeof(T) * 2 * nn * batch),
        info(c, sizeof(int) * batch), z(c, sizeof(T) * 2 * batch);
    hipLaunchKernelGGL(k_fill_tune<T>, dim3(nblk(nn * batch)), dim3(256), 0, c->stream, A.as<T>(), n, batch);
    zero_dev(c, z.p, sizeof(T) * 2 * batch, c->stream);
    hipcheck(c, hipGetLastError(), "fill tune");
    reg_region(c, A.p, sizeof(T) * 2 * nn * batch);
    reg_region(c, W.p, sizeof(T) * 2 * nn * batch);
    reg_region(c, X.p, sizeof(T) * 2 * nn * batch);
    reg_region(c, info.p, sizeof(int) * batch);
    reg_region(c, z.p, sizeof(T) * 2 * batch);
    int widths[INV_MAX_WIDTHS];
    const int k = inv_panel_widths<T>(n, widths);
    hipEvent_t e0, e1;
    hipcheck(c, hipEventCreate(&e0), "event");
    hipcheck(c, hipEventCreate(&e1), "event");
    double best = 1e300;
    int bnb = widths[0];
    const bool prof = c->prof;
    c->prof = false;
    for (int i = 0; i < k; ++i) {
        c->inv_nb[{(int)sizeof(T), n}] = widths[i];
        auto run = [&] {
            inverse<T>(c, n, bw(W.as<T>(), nn, n), bm<T>(A.as<T>(), nn, n), z.as<T>(), bw(X.as<T>(), nn, n),
                       info.as<int>(), batch);
        };
        run();   // warm-up (first launch configures the kernel)
        hipcheck(c, hipEventRecord(e0, c->stream), "event");
        const int reps = 3;
        for (int r = 0; r < reps; ++r) run();
        hipcheck(c, hipEventRecord(e1, c->stream), "event");
        hipcheck(c, hipEventSynchronize(e1), "event sync");
        float ms = 0.f;
        hipcheck(c, hipEventElapsedTime(&ms, e0, e1), "elapsed");
        if (ms / reps < best) {
            best = ms / reps;
            bnb = widths[i];
        }
    }
    c->prof = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->inv_nb[{(int)sizeof(T), n}] = bnb;
    if (best_nb) *best_nb = bnb;
    if (best_ms) *best_ms = best;
}

// Time the n x n x n GEMM of `batch` items with each output tile and keep
// the fastest for GEMMs whose larger output dimension is n.
template <typename T>
void tune_gemm(nt_ctx* c, int n, int batch, int* best_tile, double* best_ms)
{
    const long nn = (long)n * n;
    DevBuf A(c, sizeof(T) * 2 * nn * batch), B(c, sizeof(T) * 2 * nn * batch), X(c, sizeof(T) * 2 * nn * batch);
    hipLaunchKernelGGL(k_fill_tune<T>, dim3(nblk(nn * batch)), dim3(256), 0, c->stream, A.as<T>(), n, batch);
    hipLaunchKernelGGL(k_fill_tune<T>, dim3(nblk(nn * batch)), dim3(256), 0, c->stream, B.as<T>(), n, batch);
    hipcheck(c, hipGetLastError(), "fill tune");
    reg_region(c, A.p, sizeof(T) * 2 * nn * batch);
    reg_region(c, B.p, sizeof(T) * 2 * nn * batch);
    reg_region(c, X.p, sizeof(T) * 2 * nn * batch);
    const int tiles[2] = {64, 32};
    hipEvent_t e0, e1;
    hipcheck(c, hipEventCreate(&e0), "event");
    hipcheck(c, hipEventCreate(&e1), "event");
    double best = 1e300;
    int bt = 64;
    const bool prof = c->prof;
    c->prof = false;
    for (int tb : tiles) {
        c->gemm_tile[{(int)sizeof(T), n}] = tb;
        auto run = [&] {
            gemm<T>(c, n, n, n, OP_N, bm<T>(A.as<T>(), nn, n), OP_N, bm<T>(B.as<T>(), nn, n), bw(X.as<T>(), nn, n),
                    1.0, none<T>(), 0.0, nullptr, batch);
        };
        run();
        hipcheck(c, hipEventRecord(e0, c->stream), "event");
        const int reps = 5;
        for (int r = 0; r < reps; ++r) run();
        hipcheck(c, hipEventRecord(e1, c->stream), "event");
        hipcheck(c, hipEventSynchronize(e1), "event sync");
        float ms = 0.f;
        hipcheck(c, hipEventElapsedTime(&ms, e0, e1), "elapsed");
        if (ms / reps < best) {
            best = ms / reps;
            bt = tb;
        }
    }
    c->prof = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->gemm_tile[{(int)sizeof(T), n}] = bt;
    if (best_tile) *best_tile = bt;
    if (best_ms) *best_ms = best;
}

extern "C" {

int nt_params_validate(const nt_params* p, int nl, int nr)
{
    if (!p || check_params(p) || nl <= 0 || nr <= 0 || check_circular(p, nl, nr)) return NT_ERR_ARG;
    return NT_OK;
}

void nt_params_default(nt_params* p)
{
    if (!p) return;
    p->eta = 1.0e-6;
    p->eta0 = 1.0e-6;
    p->convergence = 1.0e-12;
    p->device_eta_factor = 10.0;
    p->lambda_eta = 5.0e-3;
    p->eta_mode = NT_ETA_RELATIVE;
    p->max_iter = 300;
    p->precision = NT_PREC_FP64;
    p->circular = 0;
    p->n_bloch_x = 1;
    p->n_bloch_y = 1;
    p->path_mode = 2;
    p->dos_mode = NT_DOS_FROM_ETA;
}

int nt_create(nt_ctx** out, int device_ordinal)
{
    if (!out) return NT_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return NT_ERR_HIP;
    if (device_ordinal < 0 || device_ordinal >= ndev) return NT_ERR_ARG;
    nt_ctx* c = new nt_ctx();
    c->device = device_ordinal;
    if (hipSetDevice(device_ordinal) != hipSuccess || hipStreamCreate(&c->own) != hipSuccess ||
        hipStreamCreate(&c->aux) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->chk[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->chk[1], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&c->host_flag), 2 * sizeof(int)) != hipSuccess) {
        nt_destroy(c);        // releases whatever was created
        return NT_ERR_HIP;
    }
    c->stream = c->own;
    {
        // this device's copy of the static scratch, looked up once per device
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (device_ordinal < NT_MAX_DEVICES && !g_scratch_dev[device_ordinal]) {
            void* ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
                nt_destroy(c);
                return NT_ERR_HIP;
            }
            g_scratch_dev[device_ordinal] = ptr;
        }
    }
    *out = c;
    return NT_OK;
}

int nt_destroy(nt_ctx* c)
{
    if (!c) return NT_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    if (c->ws) (void)hipFree(c->ws);
    for (auto& u : c->circ_u) (void)hipFree(u.d);
    if (c->deg_buf) (void)hipFree(c->deg_buf);
    for (nt_ctx* s : c->subs) nt_destroy(s);
    if (c->host_flag) (void)hipHostFree(c->host_flag);
    if (c->own) (void)hipStreamDestroy(c->own);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    for (hipEvent_t e : c->chk)
        if (e) (void)hipEventDestroy(e);
    delete c;
    return NT_OK;
}

int nt_sync(nt_ctx* c)
{
    return guarded(c, [&] {
        hipcheck(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        return 0;
    });
}

void* nt_stream(nt_ctx* c) { return c ? (void*)c->stream : nullptr; }

int nt_set_stream(nt_ctx* c, void* stream)
{
    if (!c) return NT_ERR_ARG;
    c->stream = stream ? (hipStream_t)stream : c->own;
    return NT_OK;
}

int nt_set_batch_limit(nt_ctx* c, int max_points)
{
    if (!c || max_points < 1 || max_points > 65535) return NT_ERR_ARG;
    c->batch_limit = max_points;
    return NT_OK;
}

const char* nt_last_error(nt_ctx* c) { return c ? c->err.c_str() : "null context"; }

const char* nt_status_string(int s)
{
    switch (s) {
    case NT_OK: return "ok";
    case NT_ERR_NAN: return "NaN in a matrix to invert";
    case NT_ERR_NOT_CONVERGED: return "surface decimation did not converge within max_iter";
    case NT_ERR_ARG: return "invalid argument";
    case NT_ERR_HIP: return "HIP runtime error";
    case NT_ERR_NOMEM: return "out of device memory";
    default: return s > 0 ? "Matrix is numerically singular!" : "unknown status";
    }
}

int nt_profile_enable(nt_ctx* c, int on)
{
    if (!c) return NT_ERR_ARG;
    c->prof = on != 0;
    return NT_OK;
}

int nt_profile_reset(nt_ctx* c)
{
    if (!c) return NT_ERR_ARG;
    for (int k = 0; k < NT_K_COUNT; ++k) {
        c->launches[k] = 0;
        c->ms[k] = 0;
        c->flops[k] = 0;
        c->flops_pending[k] = 0;
    }
    for (int k = 0; k < 3; ++k) c->mixed_counts[k] = 0;
    return NT_OK;
}

int nt_mixed_counts(nt_ctx* c, long long* counts)
{
    if (!c || !counts) return NT_ERR_ARG;
    for (int k = 0; k < 3; ++k) counts[k] = c->mixed_counts[k];
    return NT_OK;
}

int nt_profile_get(nt_ctx* c, int kind, long long* launches, double* total_ms, double* flops)
{
    if (!c || kind < 0 || kind >= NT_K_COUNT) return NT_ERR_ARG;
    return guarded(c, [&] {
        harvest_profile(c);
        if (launches) *launches = c->launches[kind];
        if (total_ms) *total_ms = c->ms[kind];
        if (flops) *flops = c->flops[kind];
        return 0;
    });
}

double nt_last_mean_iters(nt_ctx* c) { return c ? c->last_mean_iters : 0.0; }

static int tdb_entry(nt_ctx* c, const nt_params* p, int nl, int nr, int b, int L, int npts, const double* E,
                     const int* kidx, int nk, const nt_c128* Haml, const nt_c128* Hamlv, const nt_c128* Hamr,
                     const nt_c128* Hamrv, const nt_c128* Hamd, const nt_c128* Hamdv, const nt_c128* Hamld,
                     const nt_c128* Hamrd, const int* first_point_mask, double* T_LR, double* T_RL, int* status,
                     int* iters, const nt_solve_extras* extras, const nt_probe_desc* pd, const nt_probe_out* po,
                     bool noalloc = false)
{
    if (!c) return NT_ERR_ARG;
    if (noalloc && (p && p->circular)) return NT_ERR_ARG;   // circular decimation allocates its transforms
    if (check_params(p) || check_circular(p, nl, nr)) return NT_ERR_ARG;
    if (nl <= 0 || nr <= 0 || b <= 0 || L <= 0 || npts < 0 || nk <= 0 || (nk > 1 && !kidx)) return NT_ERR_ARG;
    if (std::max(std::max(nl, nr), b) > inv_max_n()) return NT_ERR_ARG;
    if (!E || !Haml || !Hamlv || !Hamr || !Hamrv || !Hamd || (L > 1 && !Hamdv) || !Hamld || !Hamrd || !T_LR || !T_RL)
        return NT_ERR_ARG;
    if (npts == 0) return NT_OK;
    return guarded(c, [&] {
        // caller buffers, sized from the arguments, for the bounds audit
        const size_t z = sizeof(nt_c128);
        const size_t nll = (size_t)nl * nl, nrr = (size_t)nr * nr, bb = (size_t)b * b;
        reg_region(c, E, sizeof(double) * npts);
        reg_region(c, kidx, sizeof(int) * npts);
        reg_region(c, Haml, z * nll * nk);
        reg_region(c, Hamlv, z * nll * nk);
        reg_region(c, Hamr, z * nrr * nk);
        reg_region(c, Hamrv, z * nrr * nk);
        reg_region(c, Hamd, z * bb * L * nk);
        reg_region(c, Hamdv, z * bb * std::max(L - 1, 1) * nk);
        reg_region(c, Hamld, z * nl * b * nk);
        reg_region(c, Hamrd, z * nr * b * nk);
        reg_region(c, first_point_mask, sizeof(int) * npts);
        reg_region(c, T_LR, sizeof(double) * npts);
        reg_region(c, T_RL, sizeof(double) * npts);
        reg_region(c, status, sizeof(int) * npts);
        reg_region(c, iters, sizeof(int) * npts);
        nt_solve_extras ex = extras ? *extras : nt_solve_extras{};
        reg_region(c, ex.gl_adv, z * nll * npts);
        reg_region(c, ex.gr_ret, z * nrr * npts);
        reg_region(c, ex.gl_plus, z * nll * npts);
        reg_region(c, ex.gr_minus, z * nrr * npts);
        reg_region(c, ex.gml_adv, z * nll * npts);
        reg_region(c, ex.gmr_ret, z * nrr * npts);
        reg_region(c, ex.sl_adv, z * nll * npts);
        reg_region(c, ex.sr_ret, z * nrr * npts);
        reg_region(c, ex.G_D_LR, z * nl * nr * npts);
        reg_region(c, ex.G_D_RL, z * nl * nr * npts);
        reg_region(c, ex.G_L, z * nll * npts);
        reg_region(c, ex.G_R, z * nrr * npts);
        reg_region(c, ex.sdos_l, sizeof(double) * npts);
        reg_region(c, ex.sdos_r, sizeof(double) * npts);
        reg_region(c, ex.ldos, sizeof(double) * npts * L * b);
        reg_region(c, ex.ddos, sizeof(double) * npts);
        if ((ex.ldos || ex.ddos) && pd) {
            c->err = "the device DOS (ldos, ddos) is not produced by the probe solve";
            return (int)NT_ERR_ARG;
        }
        if (kidx) {
            // every k index must name one of the nk block sets
            std::vector<int> hk(npts);
            hipcheck(c, hipMemcpyAsync(hk.data(), kidx, sizeof(int) * npts, hipMemcpyDeviceToHost, c->stream), "D2H");
            hipcheck(c, hipStreamSynchronize(c->stream), "sync");
            for (int v : hk)
                if (v < 0 || v >= nk) {
                    c->err = "kidx out of range [0, nk)";
                    return (int)NT_ERR_ARG;
                }
        }
        TdbArgs a{nl, nr, b, L, npts, nk, E, kidx, Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd,
                  first_point_mask, T_LR, T_RL, status, iters, ex};
        a.noalloc = noalloc;
        ProbeSpec ps{};
        if (pd) {
            reg_region(c, pd->temperature, sizeof(double) * L);
            reg_region(c, po->flux, sizeof(double) * npts * L);
            reg_region(c, po->jacobian, sizeof(double) * npts * L * L);
            reg_region(c, po->flux_left, sizeof(double) * npts);
            reg_region(c, po->flux_right, sizeof(double) * npts);
            reg_region(c, po->a_sum, sizeof(double) * npts * L);
            reg_region(c, po->gn_sum, sizeof(double) * npts * L);
            ps = ProbeSpec{pd->temperature, pd->t_left,  pd->t_right,    pd->B_probe,     pd->T_probe, pd->lifetime,
                           po->flux,        po->jacobian, po->flux_left, po->flux_right, po->a_sum,   po->gn_sum};
            a.probe = &ps;
        }
        // the FP64 solve of many points in parts solved side by side by worker
        // contexts: the latency-bound launches of one part (inverses,
        // decimation steps of few unconverged items) leave compute units to
        // the other's GEMMs.  Not the probe solve, whose grouped GEMMs fill
        // the GPU by themselves (two parts: 287 against 297 probe solves/s at
        // b = 192, L = 40, 128 points)
        const int parts = c->point_split;
        if (parts > 1 && !noalloc && !pd && npts >= parts * NT_SPLIT_MIN_POINTS) {
            solve_split(c, p, a, parts);
            return 0;
        }
        solve_tdb_dev(c, p, a);
        return 0;
    });
}

int nt_solve_tdb_batched_dev(nt_ctx* c, const nt_params* p, int nl, int nr, int b, int L, int npts, const double* E,
                             const int* kidx, int nk, const nt_c128* Haml, const nt_c128* Hamlv, const nt_c128* Hamr,
                             const nt_c128* Hamrv, const nt_c128* Hamd, const nt_c128* Hamdv, const nt_c128* Hamld,
                             const nt_c128* Hamrd, const int* first_point_mask, double* T_LR, double* T_RL,
                             int* status, int* iters, const nt_solve_extras* extras)
{
    return tdb_entry(c, p, nl, nr, b, L, npts, E, kidx, nk, Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd,
                     first_point_mask, T_LR, T_RL, status, iters, extras, nullptr, nullptr);
}

int nt_solve_tdb_batched_noalloc_dev(nt_ctx* c, const nt_params* p, int nl, int nr, int b, int L, int npts,
                                     const double* E, const int* kidx, int nk, const nt_c128* Haml,
                                     const nt_c128* Hamlv, const nt_c128* Hamr, const nt_c128* Hamrv,
                                     const nt_c128* Hamd, const nt_c128* Hamdv, const nt_c128* Hamld,
                                     const nt_c128* Hamrd, const int* first_point_mask, double* T_LR, double* T_RL,
                                     int* status, int* iters, const nt_solve_extras* extras)
{
    return tdb_entry(c, p, nl, nr, b, L, npts, E, kidx, nk, Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd,
                     first_point_mask, T_LR, T_RL, status, iters, extras, nullptr, nullptr, true);
}

int nt_buttiker_tdb_batched_dev(nt_ctx* c, const nt_params* p, const nt_probe_desc* probe, int nl, int nr, int b,
                                int L, int npts, const double* E, const int* kidx, int nk, const nt_c128* Haml,
                                const nt_c128* Hamlv, const nt_c128* Hamr, const nt_c128* Hamrv, const nt_c128* Hamd,
                                const nt_c128* Hamdv, const nt_c128* Hamld, const nt_c128* Hamrd,
                                const int* first_point_mask, double* T_LR, double* T_RL, int* status, int* iters,
                                const nt_solve_extras* extras, const nt_probe_out* out)
{
    if (!c) return NT_ERR_ARG;
    if (!probe || !out || !probe->temperature || L < 2 || nl != nr) return NT_ERR_ARG;
    if (!out->flux || !out->jacobian || !out->flux_left || !out->flux_right || !out->a_sum || !out->gn_sum)
        return NT_ERR_ARG;
    if (!(probe->B_probe > 0.0) || !(probe->t_left > 0.0) || !(probe->t_right > 0.0)) return NT_ERR_ARG;
    return tdb_entry(c, p, nl, nr, b, L, npts, E, kidx, nk, Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd,
                     first_point_mask, T_LR, T_RL, status, iters, extras, probe, out);
}

int nt_solve_tdb_adjoint_batched_dev(nt_ctx* c, const nt_params* p, int nl, int nr, int b, int L, int npts,
                                     const double* E, const int* kidx, int nk, const nt_c128* Haml,
                                     const nt_c128* Hamlv, const nt_c128* Hamr, const nt_c128* Hamrv,
                                     const nt_c128* Hamd, const nt_c128* Hamdv, const nt_c128* Hamld,
                                     const nt_c128* Hamrd, const int* first_point_mask, const double* weight, int lo,
                                     int hi, double* T_LR, double* T_RL, int* status, int* iters, nt_c128* dW_dHamd,
                                     nt_c128* dW_dHamdv)
{
    if (!c) return NT_ERR_ARG;
    if (check_params(p) || check_circular(p, nl, nr)) return NT_ERR_ARG;
    if (nl <= 0 || nr <= 0 || b <= 0 || L <= 0 || npts < 0 || nk <= 0 || (nk > 1 && !kidx)) return NT_ERR_ARG;
    if (std::max(std::max(nl, nr), b) > inv_max_n()) return NT_ERR_ARG;
    if (!E || !Haml || !Hamlv || !Hamr || !Hamrv || !Hamd || (L > 1 && !Hamdv) || !Hamld || !Hamrd || !T_LR || !T_RL)
        return NT_ERR_ARG;
    if (!weight || !dW_dHamd || lo < 0 || hi < lo || hi >= L) return NT_ERR_ARG;
    const int nc = std::min(hi, L - 2) - lo + 1;
    if (nc > 0 && !dW_dHamdv) return NT_ERR_ARG;
    if (npts == 0) return NT_OK;
    return guarded(c, [&] {
        const size_t z = sizeof(nt_c128);
        const size_t nll = (size_t)nl * nl, nrr = (size_t)nr * nr, bb = (size_t)b * b;
        reg_region(c, E, sizeof(double) * npts);
        reg_region(c, kidx, sizeof(int) * npts);
        reg_region(c, Haml, z * nll * nk);
        reg_region(c, Hamlv, z * nll * nk);
        reg_region(c, Hamr, z * nrr * nk);
        reg_region(c, Hamrv, z * nrr * nk);
        reg_region(c, Hamd, z * bb * L * nk);
        reg_region(c, Hamdv, z * bb * std::max(L - 1, 1) * nk);
        reg_region(c, Hamld, z * nl * b * nk);
        reg_region(c, Hamrd, z * nr * b * nk);
        reg_region(c, first_point_mask, sizeof(int) * npts);
        reg_region(c, weight, sizeof(double) * npts);
        reg_region(c, T_LR, sizeof(double) * npts);
        reg_region(c, T_RL, sizeof(double) * npts);
        reg_region(c, status, sizeof(int) * npts);
        reg_region(c, iters, sizeof(int) * npts);
        reg_region(c, dW_dHamd, z * bb * (hi - lo + 1) * nk);
        if (nc > 0) reg_region(c, dW_dHamdv, z * bb * nc * nk);
        if (kidx) {
            std::vector<int> hk(npts);
            hipcheck(c, hipMemcpyAsync(hk.data(), kidx, sizeof(int) * npts, hipMemcpyDeviceToHost, c->stream), "D2H");
            hipcheck(c, hipStreamSynchronize(c->stream), "sync");
            for (int v : hk)
                if (v < 0 || v >= nk) {
                    c->err = "kidx out of range [0, nk)";
                    return (int)NT_ERR_ARG;
                }
        }
        TdbArgs a{nl, nr, b, L, npts, nk, E, kidx, Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd,
                  first_point_mask, T_LR, T_RL, status, iters, nt_solve_extras{}};
        a.adj_w = weight;
        a.adj_lo = lo;
        a.adj_hi = hi;
        a.dHd = dW_dHamd;
        a.dHdv = nc > 0 ? dW_dHamdv : nullptr;
        // many points: parts solved side by side by worker contexts, as the
        // batched solve (the adjoint sweep runs in FP64 whatever the precision)
        const int parts = c->point_split;
        if (parts > 1 && npts >= parts * NT_SPLIT_MIN_POINTS) {
            solve_split(c, p, a, parts);
            return 0;
        }
        solve_tdb_dev(c, p, a);
        return 0;
    });
}

int nt_solve_tdb_batched(nt_ctx* c, const nt_params* p, int nl, int nr, int b, int L, int npts, const double* E,
                         const int* kidx, int nk, const nt_c128* Haml, const nt_c128* Hamlv, const nt_c128* Hamr,
                         const nt_c128* Hamrv, const nt_c128* Hamd, const nt_c128* Hamdv, const nt_c128* Hamld,
                         const nt_c128* Hamrd, const int* first_point_mask, double* T_LR, double* T_RL, int* status,
                         int* iters, const nt_solve_extras* extras)
{
    if (!c) return NT_ERR_ARG;
    if (check_params(p) || check_circular(p, nl, nr)) return NT_ERR_ARG;
    if (nl <= 0 || nr <= 0 || b <= 0 || L <= 0 || npts < 0 || nk <= 0 || (nk > 1 && !kidx)) return NT_ERR_ARG;
    if (!E || !Haml || !Hamlv || !Hamr || !Hamrv || !Hamd || (L > 1 && !Hamdv) || !Hamld || !Hamrd || !T_LR || !T_RL)
        return NT_ERR_ARG;
    if (npts == 0) return NT_OK;
    if (kidx)
        for (int i = 0; i < npts; ++i)
            if (kidx[i] < 0 || kidx[i] >= nk) return NT_ERR_ARG;
    return guarded(c, [&] {
        const size_t z = sizeof(nt_c128);
        const size_t nll = (size_t)nl * nl, nrr = (size_t)nr * nr, bb = (size_t)b * b;
        DevBuf dE(c, sizeof(double) * npts), dk(c, kidx ? sizeof(int) * npts : 0);
        DevBuf dHl(c, z * nll * nk), dHlv(c, z * nll * nk), dHr(c, z * nrr * nk), dHrv(c, z * nrr * nk);
        DevBuf dHd(c, z * bb * L * nk), dHdv(c, z * bb * std::max(L - 1, 1) * nk);
        DevBuf dHld(c, z * nl * b * nk), dHrd(c, z * nr * b * nk);
        DevBuf dfm(c, first_point_mask ? sizeof(int) * npts : 0);
        DevBuf dTl(c, sizeof(double) * npts), dTr(c, sizeof(double) * npts);
        DevBuf dst(c, sizeof(int) * npts), dit(c, sizeof(int) * npts);
        up(c, dE, E, sizeof(double) * npts);
        if (kidx) up(c, dk, kidx, sizeof(int) * npts);
        up(c, dHl, Haml, z * nll * nk);
        up(c, dHlv, Hamlv, z * nll * nk);
        up(c, dHr, Hamr, z * nrr * nk);
        up(c, dHrv, Hamrv, z * nrr * nk);
        up(c, dHd, Hamd, z * bb * L * nk);
        if (L > 1) up(c, dHdv, Hamdv, z * bb * (L - 1) * nk);
        up(c, dHld, Hamld, z * nl * b * nk);
        up(c, dHrd, Hamrd, z * nr * b * nk);
        if (first_point_mask) up(c, dfm, first_point_mask, sizeof(int) * npts);
        // optional extras
        nt_solve_extras ex{};
        std::vector<DevBuf*> owned;
        struct Map {
            nt_c128* host;
            nt_c128** dev;
            size_t n;
        };
        std::vector<Map> maps;
        nt_solve_extras hx = extras ? *extras : nt_solve_extras{};
        auto want = [&](nt_c128* h, nt_c128** d, size_t elems) {
            if (!h) return;
            DevBuf* db = new DevBuf(c, z * elems * npts);
            owned.push_back(db);
            *d = db->as<nt_c128>();
            maps.push_back({h, d, elems});
        };
        want(hx.gl_adv, &ex.gl_adv, nll);
        want(hx.gr_ret, &ex.gr_ret, nrr);
        want(hx.gl_plus, &ex.gl_plus, nll);
        want(hx.gr_minus, &ex.gr_minus, nrr);
        want(hx.gml_adv, &ex.gml_adv, nll);
        want(hx.gmr_ret, &ex.gmr_ret, nrr);
        want(hx.sl_adv, &ex.sl_adv, nll);
        want(hx.sr_ret, &ex.sr_ret, nrr);
        want(hx.G_D_LR, &ex.G_D_LR, (size_t)nl * nr);
        want(hx.G_D_RL, &ex.G_D_RL, (size_t)nl * nr);
        want(hx.G_L, &ex.G_L, nll);
        want(hx.G_R, &ex.G_R, nrr);
        DevBuf dsl(c, hx.sdos_l ? sizeof(double) * npts : 0), dsr(c, hx.sdos_r ? sizeof(double) * npts : 0);
        ex.sdos_l = hx.sdos_l ? dsl.as<double>() : nullptr;
        ex.sdos_r = hx.sdos_r ? dsr.as<double>() : nullptr;
        const size_t nld = (size_t)L * b * npts;
        DevBuf dld(c, hx.ldos ? sizeof(double) * nld : 0), ddd(c, hx.ddos ? sizeof(double) * npts : 0);
        ex.ldos = hx.ldos ? dld.as<double>() : nullptr;
        ex.ddos = hx.ddos ? ddd.as<double>() : nullptr;
        int rc = 0;
        try {
            TdbArgs a{nl, nr, b, L, npts, nk, dE.as<double>(), kidx ? dk.as<int>() : nullptr,
                      dHl.as<nt_c128>(), dHlv.as<nt_c128>(), dHr.as<nt_c128>(), dHrv.as<nt_c128>(),
                      dHd.as<nt_c128>(), dHdv.as<nt_c128>(), dHld.as<nt_c128>(), dHrd.as<nt_c128>(),
                      first_point_mask ? dfm.as<int>() : nullptr, dTl.as<double>(), dTr.as<double>(),
                      dst.as<int>(), dit.as<int>(), ex};
            solve_tdb_dev(c, p, a);
            down(c, T_LR, dTl, sizeof(double) * npts);
            down(c, T_RL, dTr, sizeof(double) * npts);
            std::vector<int> hst(npts);
            down(c, hst.data(), dst, sizeof(int) * npts);
            if (iters) down(c, iters, dit, sizeof(int) * npts);
            for (auto& m : maps) {
                DevBuf tmp(c, 0);
                hipcheck(c, hipMemcpyAsync(m.host, *m.dev, z * m.n * npts, hipMemcpyDeviceToHost, c->stream), "D2H");
            }
            if (hx.sdos_l) down(c, hx.sdos_l, dsl, sizeof(double) * npts);
            if (hx.sdos_r) down(c, hx.sdos_r, dsr, sizeof(double) * npts);
            if (hx.ldos) down(c, hx.ldos, dld, sizeof(double) * nld);
            if (hx.ddos) down(c, hx.ddos, ddd, sizeof(double) * npts);
            hipcheck(c, hipStreamSynchronize(c->stream), "sync");
            if (status) std::memcpy(status, hst.data(), sizeof(int) * npts);
            rc = first_status_host(hst.data(), npts);
        } catch (...) {
            for (auto* d : owned) delete d;
            throw;
        }
        for (auto* d : owned) delete d;
        return rc;
    });
}

int nt_surface_green_tdb(nt_ctx* c, const nt_params* p, int nl, int nr, double E, const nt_c128* Haml,
                         const nt_c128* Hamr, const nt_c128* Hamlv, const nt_c128* Hamrv, nt_c128* gl_adv,
                         nt_c128* gr_ret, nt_c128* gl_plus, nt_c128* gr_minus, nt_c128* gml_adv, nt_c128* gmr_ret,
                         nt_c128* sl_adv, nt_c128* sr_ret, int* iters)
{
    if (!c) return NT_ERR_ARG;
    if (check_params(p) || nl <= 0 || nr <= 0 || !Haml || !Hamr || !Hamlv || !Hamrv) return NT_ERR_ARG;
    if (check_circular(p, nl, nr)) return NT_ERR_ARG;
    if (std::max(nl, nr) > inv_max_n()) return NT_ERR_ARG;
    return guarded(c, [&] {
        typedef double T;
        const long nll = (long)nl * nl, nrr = (long)nr * nr;
        const size_t z = sizeof(nt_c128);
        DevBuf dHl(c, z * nll), dHlv(c, z * nll), dHr(c, z * nrr), dHrv(c, z * nrr);
        up(c, dHl, Haml, z * nll);
        up(c, dHlv, Hamlv, z * nll);
        up(c, dHr, Hamr, z * nrr);
        up(c, dHrv, Hamrv, z * nrr);
        DevBuf dE(c, sizeof(double)), dz(c, sizeof(double) * 6), dst(c, sizeof(int) * 4), dit(c, sizeof(int) * 4);
        up(c, dE, &E, sizeof(double));
        const double et = p->eta_mode == NT_ETA_RELATIVE ? E * p->eta + p->eta0 : p->eta;
        // z: [E - i eta, E + i eta, E + i eta, E - i eta]
        const double zh[8] = {E, -et, E, et, E, et, E, -et};
        DevBuf dzz(c, sizeof(zh));
        up(c, dzz, zh, sizeof(zh));
        zero_dev(c, dst.p, sizeof(int) * 4, c->stream);
        const int nmax = std::max(nl, nr);
        const long nmm = (long)nmax * nmax;
        DevBuf dg(c, z * nmm * 4), dsig(c, z * nmm * 2), dgam(c, z * nmm * 2);
        Arena ar;
        ensure_ws(c, lead_gfs_scratch(p, nmax, 1, 1));
        ar.base = static_cast<char*>(c->ws);
        const T* zz = dzz.as<T>();
        auto one = [&](int n, const T* H, const T* V, int adag, const T* zp, T* g, T* sig, T* gam, int* st, int* it) {
            DecimationOut<T> o{bw(g, (long)n * n, n), sig ? bw(sig, (long)n * n, n) : nonew<T>(),
                               gam ? bw(gam, (long)n * n, n) : nonew<T>()};
            lead_gfs(c, p, n, 1, 1, H, V, nullptr, adag, zp, o, st, it, ar, 0);
        };
        T* g = dg.as<T>();
        one(nl, dHl.as<T>(), dHlv.as<T>(), 1, zz + 0, g, dsig.as<T>(), dgam.as<T>(), dst.as<int>(), dit.as<int>());
        if (gl_plus)
            one(nl, dHl.as<T>(), dHlv.as<T>(), 0, zz + 4, g + 2 * nmm, nullptr, nullptr, dst.as<int>() + 1,
                dit.as<int>() + 1);
        one(nr, dHr.as<T>(), dHrv.as<T>(), 1, zz + 2, g + 4 * nmm, dsig.as<T>() + 2 * nmm, dgam.as<T>() + 2 * nmm,
            dst.as<int>() + 2, dit.as<int>() + 2);
        if (gr_minus)
            one(nr, dHr.as<T>(), dHrv.as<T>(), 0, zz + 6, g + 6 * nmm, nullptr, nullptr, dst.as<int>() + 3,
                dit.as<int>() + 3);
        int hst[4] = {0, 0, 0, 0}, hit[4] = {0, 0, 0, 0};
        hipcheck(c, hipMemcpyAsync(hst, dst.p, sizeof(hst), hipMemcpyDeviceToHost, c->stream), "D2H");
        hipcheck(c, hipMemcpyAsync(hit, dit.p, sizeof(hit), hipMemcpyDeviceToHost, c->stream), "D2H");
        auto cp = [&](nt_c128* h, const T* d, long elems) {
            if (h) hipcheck(c, hipMemcpyAsync(h, d, z * elems, hipMemcpyDeviceToHost, c->stream), "D2H");
        };
        cp(gl_adv, g, nll);
        cp(gl_plus, g + 2 * nmm, nll);
        cp(gr_ret, g + 4 * nmm, nrr);
        cp(gr_minus, g + 6 * nmm, nrr);
        cp(sl_adv, dsig.as<T>(), nll);
        cp(sr_ret, dsig.as<T>() + 2 * nmm, nrr);
        cp(gml_adv, dgam.as<T>(), nll);
        cp(gmr_ret, dgam.as<T>() + 2 * nmm, nrr);
        hipcheck(c, hipStreamSynchronize(c->stream), "sync");
        if (iters) {
            iters[0] = hit[0];
            iters[1] = gl_plus ? hit[1] : 0;
            iters[2] = hit[2];
            iters[3] = gr_minus ? hit[3] : 0;
        }
        for (int i = 0; i < 4; ++i)
            if (hst[i]) return hst[i];
        return 0;
    });
}

// layer LDOS of get_ldos (tb/device.f90:610-855):
//   G^L_i left-connected from gl, G^R_i right-connected from gr,
//   GLR_i = (I - G^L_i V_{i,i+1} G^R_{i+1} V_{i,i+1}^+)^{-1} G^L_i   (:800-852)
// with the right lead (Vc, gr) closing the last layer; ldos(i,j) = -Im GLR_i(j,j)/pi.
__global__ void k_ldos_diag(const double2* X, int n, int layer, int L, int maxlen, double* ldos)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= maxlen) return;
    const double v = j < n ? -X[j + (long)j * n].y / 3.14159265358979323846 : 0.0;
    ldos[layer + (long)j * L] = v;
}

static int ldos_host(nt_ctx* c, const nt_params* p, int nl, int nr, int L, const int* bsz, double E,
                     const nt_c128* gl, const nt_c128* gr, const nt_c128* const* Hd, const nt_c128* const* Hdv,
                     const nt_c128* const* Hdv_back, const nt_c128* Hld, const nt_c128* Hrd, const nt_c128* Hrd_fwd,
                     int maxlen, double* ldos, double* ddos)
{
    typedef double T;
    const size_t z = sizeof(nt_c128);
    std::vector<DevBuf*> owned;
    auto upm = [&](const nt_c128* h, long elems) -> const T* {
        if (!h) return nullptr;
        DevBuf* d = new DevBuf(c, z * elems);
        owned.push_back(d);
        up(c, *d, h, z * elems);
        return d->as<T>();
    };
    auto buf = [&](long elems) -> T* {
        DevBuf* d = new DevBuf(c, z * elems);
        owned.push_back(d);
        return d->as<T>();
    };
    int rc = 0;
    try {
        int bmax = std::max(nl, nr);
        for (int i = 0; i < L; ++i) bmax = std::max(bmax, bsz[i]);
        const long mm = (long)bmax * bmax;
        std::vector<BMat<T>> hd, hv, hb;
        for (int i = 0; i < L; ++i) hd.push_back(bm<T>(upm(Hd[i], (long)bsz[i] * bsz[i]), 0, bsz[i]));
        for (int i = 0; i + 1 < L; ++i) {
            hv.push_back(bm<T>(upm(Hdv[i], (long)bsz[i] * bsz[i + 1]), 0, bsz[i]));
            const nt_c128* b = Hdv_back ? Hdv_back[i] : nullptr;
            hb.push_back(b ? bm<T>(upm(b, (long)bsz[i] * bsz[i + 1]), 0, bsz[i + 1]) : none<T>());
        }
        const BMat<T> hld = bm<T>(upm(Hld, (long)nl * bsz[0]), 0, nl);
        const BMat<T> hrd = bm<T>(upm(Hrd, (long)nr * bsz[L - 1]), 0, nr);
        const BMat<T> hfw = Hrd_fwd ? bm<T>(upm(Hrd_fwd, (long)nr * bsz[L - 1]), 0, bsz[L - 1]) : none<T>();
        const BMat<T> dgl = bm<T>(upm(gl, (long)nl * nl), 0, nl);
        const BMat<T> dgr = bm<T>(upm(gr, (long)nr * nr), 0, nr);
        T* GL = buf(mm * L);
        T* GR = buf(mm);
        T* GRn = buf(mm);
        T* T1 = buf(mm);
        T* M = buf(mm);
        T* A = buf(mm);
        T* X = buf(mm);
        DevBuf dz(c, sizeof(double) * 4), dst(c, sizeof(int)), dl(c, sizeof(double) * (size_t)L * maxlen);
        const double et = p->eta_mode == NT_ETA_RELATIVE ? E * p->eta + p->eta0 : p->eta;
        const double zh[4] = {E, p->device_eta_factor * et, 1.0, 0.0};
        up(c, dz, zh, sizeof(zh));
        const T* zd = dz.as<T>();
        const T* one = dz.as<T>() + 2;
        int* st = dst.as<int>();
        zero_dev(c, st, sizeof(int), c->stream);
        auto W = [&](T* q, int ld) { return bw(q, 0, ld); };
        auto R = [&](T* q, int ld) { return bm<T>(q, 0, ld); };
        auto GLi = [&](int i) { return GL + 2 * (long)i * mm; };
        auto put = [&](int layer, int n) {
            hipLaunchKernelGGL(k_ldos_diag, dim3(nblk(maxlen)), dim3(256), 0, c->stream,
                               reinterpret_cast<const double2*>(X), n, layer, L, maxlen, dl.as<double>());
            hipcheck(c, hipGetLastError(), "ldos diag");
        };
        // left to right
        {
            const int b = bsz[0];
            gemm<T>(c, b, nl, nl, OP_C, hld, OP_N, dgl, W(T1, b), 1.0, none<T>(), 0.0, nullptr, 1);
            gemm<T>(c, b, b, nl, OP_N, R(T1, b), OP_N, hld, W(A, b), -1.0, hd[0], -1.0, zd, 1);
            inverse<T>(c, b, W(A, b), none<T>(), nullptr, W(GLi(0), b), st, 1);
            for (int i = 1; i < L; ++i) {
                const int bp = bsz[i - 1], bi = bsz[i];
                gemm<T>(c, bi, bp, bp, OP_C, hv[i - 1], OP_N, R(GLi(i - 1), bp), W(T1, bi), 1.0, none<T>(), 0.0,
                        nullptr, 1);
                gemm<T>(c, bi, bi, bp, OP_N, R(T1, bi), OP_N, hv[i - 1], W(A, bi), -1.0, hd[i], -1.0, zd, 1);
                inverse<T>(c, bi, W(A, bi), none<T>(), nullptr, W(GLi(i), bi), st, 1);
            }
        }
        // last layer: GLR_N with the right lead, Vc = Hrd_fwd or Hrd^+ ; R->L start G^R_N
        {
            const int bN = bsz[L - 1];
            const bool fwd = hfw.base != nullptr;
            const BMat<T> Vc = fwd ? hfw : hrd;
            const int opVc = fwd ? OP_N : OP_C, opVcH = fwd ? OP_C : OP_N;
            gemm<T>(c, bN, nr, bN, OP_N, R(GLi(L - 1), bN), opVc, Vc, W(T1, bN), 1.0, none<T>(), 0.0, nullptr, 1);
            gemm<T>(c, bN, nr, nr, OP_N, R(T1, bN), OP_N, dgr, W(M, bN), 1.0, none<T>(), 0.0, nullptr, 1);
            gemm<T>(c, bN, bN, nr, OP_N, R(M, bN), opVcH, Vc, W(A, bN), -1.0, none<T>(), 0.0, one, 1);
            inverse<T>(c, bN, W(A, bN), none<T>(), nullptr, W(M, bN), st, 1);
            gemm<T>(c, bN, bN, bN, OP_N, R(M, bN), OP_N, R(GLi(L - 1), bN), W(X, bN), 1.0, none<T>(), 0.0, nullptr,
                    1);
            put(L - 1, bN);
            gemm<T>(c, bN, nr, nr, OP_C, hrd, OP_N, dgr, W(T1, bN), 1.0, none<T>(), 0.0, nullptr, 1);
            gemm<T>(c, bN, bN, nr, OP_N, R(T1, bN), OP_N, hrd, W(A, bN), -1.0, hd[L - 1], -1.0, zd, 1);
            inverse<T>(c, bN, W(A, bN), none<T>(), nullptr, W(GR, bN), st, 1);
        }
        for (int i = L - 1; i >= 1; --i) {
            const int bi = bsz[i], bm_ = bsz[i - 1];
            // GLR_{i-1} = (I - G^L_{i-1} V G^R_i V^+)^{-1} G^L_{i-1}, V = V_{i-1,i}
            gemm<T>(c, bm_, bi, bm_, OP_N, R(GLi(i - 1), bm_), OP_N, hv[i - 1], W(T1, bm_), 1.0, none<T>(), 0.0,
                    nullptr, 1);
            gemm<T>(c, bm_, bi, bi, OP_N, R(T1, bm_), OP_N, R(GR, bi), W(M, bm_), 1.0, none<T>(), 0.0, nullptr, 1);
            gemm<T>(c, bm_, bm_, bi, OP_N, R(M, bm_), OP_C, hv[i - 1], W(A, bm_), -1.0, none<T>(), 0.0, one, 1);
            inverse<T>(c, bm_, W(A, bm_), none<T>(), nullptr, W(M, bm_), st, 1);
            gemm<T>(c, bm_, bm_, bm_, OP_N, R(M, bm_), OP_N, R(GLi(i - 1), bm_), W(X, bm_), 1.0, none<T>(), 0.0,
                    nullptr, 1);
            put(i - 1, bm_);
            // G^R_{i-1} = (z - H_{i-1} - V^+ G^R_i V)^{-1}, V = V_{i,i-1} (lower block or Hdv^+)
            const bool back = hb[i - 1].base != nullptr;
            const BMat<T> Vb = back ? hb[i - 1] : hv[i - 1];
            const int opV = back ? OP_N : OP_C, opVH = back ? OP_C : OP_N;
            gemm<T>(c, bm_, bi, bi, opVH, Vb, OP_N, R(GR, bi), W(T1, bm_), 1.0, none<T>(), 0.0, nullptr, 1);
            gemm<T>(c, bm_, bm_, bi, OP_N, R(T1, bm_), opV, Vb, W(A, bm_), -1.0, hd[i - 1], -1.0, zd, 1);
            inverse<T>(c, bm_, W(A, bm_), none<T>(), nullptr, W(GRn, bm_), st, 1);
            std::swap(GR, GRn);
        }
        int hs = 0;
        down(c, ldos, dl, sizeof(double) * (size_t)L * maxlen);
        hipcheck(c, hipMemcpyAsync(&hs, dst.p, sizeof(int), hipMemcpyDeviceToHost, c->stream), "D2H");
        hipcheck(c, hipStreamSynchronize(c->stream), "sync");
        if (ddos) {
            double acc = 0.0;
            for (size_t k = 0; k < (size_t)L * maxlen; ++k) acc += ldos[k];
            *ddos = acc;
        }
        rc = hs;
    } catch (...) {
        for (auto* d : owned) delete d;
        throw;
    }
    for (auto* d : owned) delete d;
    return rc;
}

static int rgf_host_common(nt_ctx* c, const nt_params* p, int first_point, int nl, int nr, int L, const int* bsz,
                           double E, const nt_c128* gl, const nt_c128* gr, const nt_c128* gml, const nt_c128* gmr,
                           const nt_c128* sl, const nt_c128* sr, const nt_c128* const* Hd, const nt_c128* const* Hdv,
                           const nt_c128* const* Hdv_back, const nt_c128* Hld, const nt_c128* Hrd,
                           const nt_c128* Hrd_fwd, const nt_c128* Haml, const nt_c128* Hamr, double* transl,
                           double* transr, nt_c128* G_D_RL, nt_c128* G_D_LR, nt_c128* G_L, nt_c128* G_R)
{
    typedef double T;
    const size_t z = sizeof(nt_c128);
    const long nll = (long)nl * nl, nrr = (long)nr * nr;
    std::vector<DevBuf*> owned;
    auto upm = [&](const nt_c128* h, long elems) -> const T* {
        if (!h) return nullptr;
        DevBuf* d = new DevBuf(c, z * elems);
        owned.push_back(d);
        up(c, *d, h, z * elems);
        return d->as<T>();
    };
    int rc = 0;
    try {
        RgfInput<T> in;
        in.L = L;
        in.bsz.assign(bsz, bsz + L);
        for (int i = 0; i < L; ++i) in.Hd.push_back(bm<T>(upm(Hd[i], (long)bsz[i] * bsz[i]), 0, bsz[i]));
        for (int i = 0; i + 1 < L; ++i) {
            in.Hdv.push_back(bm<T>(upm(Hdv[i], (long)bsz[i] * bsz[i + 1]), 0, bsz[i]));
            const nt_c128* hb = Hdv_back ? Hdv_back[i] : nullptr;
            in.Hdv_back.push_back(hb ? bm<T>(upm(hb, (long)bsz[i] * bsz[i + 1]), 0, bsz[i + 1]) : none<T>());
        }
        in.Hld = bm<T>(upm(Hld, (long)nl * bsz[0]), 0, nl);
        in.Hrd = bm<T>(upm(Hrd, (long)nr * bsz[L - 1]), 0, nr);
        in.Hrd_fwd = Hrd_fwd ? bm<T>(upm(Hrd_fwd, (long)nr * bsz[L - 1]), 0, bsz[L - 1]) : none<T>();
        in.Haml = bm<T>(upm(Haml, nll), 0, nl);
        in.Hamr = bm<T>(upm(Hamr, nrr), 0, nr);
        in.gl = bm<T>(upm(gl, nll), 0, nl);
        in.gr = bm<T>(upm(gr, nrr), 0, nr);
        in.gml = bm<T>(upm(gml, nll), 0, nl);
        in.gmr = bm<T>(upm(gmr, nrr), 0, nr);
        in.sl = bm<T>(upm(sl, nll), 0, nl);
        in.sr = bm<T>(upm(sr, nrr), 0, nr);
        DevBuf dE(c, sizeof(double)), dz(c, sizeof(double) * 2), dT(c, sizeof(double) * 2), dst(c, sizeof(int));
        const double et = p->eta_mode == NT_ETA_RELATIVE ? E * p->eta + p->eta0 : p->eta;
        const double zh[2] = {E, p->device_eta_factor * et};
        up(c, dz, zh, sizeof(zh));
        zero_dev(c, dst.p, sizeof(int), c->stream);
        DevBuf oLR(c, z * nl * nr), oRL(c, z * nl * nr), oL(c, z * nll), oR(c, z * nrr);
        RgfOutput<T> out;
        out.T_LR = dT.as<double>();
        out.T_RL = dT.as<double>() + 1;
        out.status = dst.as<int>();
        out.G_D_LR = bw(oLR.as<T>(), 0, nl);
        out.G_D_RL = bw(oRL.as<T>(), 0, nr);
        out.G_L = bw(oL.as<T>(), 0, nl);
        out.G_R = bw(oR.as<T>(), 0, nr);
        int bmax = std::max(nl, nr);
        for (int i = 0; i < L; ++i) bmax = std::max(bmax, bsz[i]);
        ensure_ws(c, rgf_scratch<T>(nl, nr, bmax, 1));
        Arena ar;
        ar.base = static_cast<char*>(c->ws);
        rgf_run<T>(c, nl, nr, in, dz.as<T>(), 1, out, ar, 0);
        double hT[2];
        int hs = 0;
        hipcheck(c, hipMemcpyAsync(hT, dT.p, sizeof(hT), hipMemcpyDeviceToHost, c->stream), "D2H");
        hipcheck(c, hipMemcpyAsync(&hs, dst.p, sizeof(int), hipMemcpyDeviceToHost, c->stream), "D2H");
        down(c, G_D_LR, oLR, z * nl * nr);
        down(c, G_D_RL, oRL, z * nl * nr);
        down(c, G_L, oL, z * nll);
        down(c, G_R, oR, z * nrr);
        hipcheck(c, hipStreamSynchronize(c->stream), "sync");
        for (int i = 0; i < 2; ++i) {
            if (hT[i] < 1.0e-60) hT[i] = 0.0;    // lesser.f90:795-800
            if (first_point) hT[i] = 0.0;        // lesser.f90:802-807
        }
        if (transl) *transl = hT[0];
        if (transr) *transr = hT[1];
        rc = hs;
    } catch (...) {
        for (auto* d : owned) delete d;
        throw;
    }
    for (auto* d : owned) delete d;
    return rc;
}

int nt_get_G_full_tdb(nt_ctx* c, const nt_params* p, int first_point, int nl, int nr, int b, int L, double E,
                      const nt_c128* gl, const nt_c128* gr, const nt_c128* gml, const nt_c128* gmr,
                      const nt_c128* sl, const nt_c128* sr, const nt_c128* hamd, const nt_c128* hamdv,
                      const nt_c128* hamld, const nt_c128* hamrd, const nt_c128* haml, const nt_c128* hamr,
                      double* transl, double* transr, nt_c128* G_D_RL, nt_c128* G_D_LR, nt_c128* G_L, nt_c128* G_R)
{
    if (!c) return NT_ERR_ARG;
    if (check_params(p) || nl <= 0 || nr <= 0 || b <= 0 || L <= 0) return NT_ERR_ARG;
    if (!gl || !gr || !gml || !gmr || !sl || !sr || !hamd || (L > 1 && !hamdv) || !hamld || !hamrd || !haml || !hamr)
        return NT_ERR_ARG;
    if (std::max(std::max(nl, nr), b) > inv_max_n()) return NT_ERR_ARG;
    std::vector<int> bsz(L, b);
    std::vector<const nt_c128*> Hd(L), Hdv(std::max(L - 1, 1));
    for (int i = 0; i < L; ++i) Hd[i] = hamd + (size_t)i * b * b;
    for (int i = 0; i + 1 < L; ++i) Hdv[i] = hamdv + (size_t)i * b * b;
    return guarded(c, [&] {
        return rgf_host_common(c, p, first_point, nl, nr, L, bsz.data(), E, gl, gr, gml, gmr, sl, sr, Hd.data(),
                               Hdv.data(), nullptr, hamld, hamrd, nullptr, haml, hamr, transl, transr, G_D_RL, G_D_LR,
                               G_L, G_R);
    });
}

int nt_G_block(nt_ctx* c, const nt_params* p, int nl, int nr, int L, const int* layer_list, double E,
               const nt_c128* gl, const nt_c128* gr, const nt_c128* gml, const nt_c128* gmr, const nt_c128* sl,
               const nt_c128* sr, const nt_c128* const* Hd, const nt_c128* const* Hdv,
               const nt_c128* const* Hdv_back, const nt_c128* Hld, const nt_c128* Hrd, const nt_c128* Hrd_fwd,
               const nt_c128* Haml, const nt_c128* Hamr, double* transl, double* transr, nt_c128* G_D_RL,
               nt_c128* G_D_LR, nt_c128* G_L, nt_c128* G_R)
{
    if (!c) return NT_ERR_ARG;
    if (check_params(p) || nl <= 0 || nr <= 0 || L <= 0 || !layer_list || !Hd || (L > 1 && !Hdv)) return NT_ERR_ARG;
    if (!gl || !gr || !gml || !gmr || !sl || !sr || !Hld || !Hrd || !Haml || !Hamr) return NT_ERR_ARG;
    int mx = std::max(nl, nr);
    for (int i = 0; i < L; ++i) {
        if (layer_list[i] <= 0 || !Hd[i] || (i + 1 < L && !Hdv[i])) return NT_ERR_ARG;
        mx = std::max(mx, layer_list[i]);
    }
    if (mx > inv_max_n()) return NT_ERR_ARG;
    return guarded(c, [&] {
        return rgf_host_common(c, p, 0, nl, nr, L, layer_list, E, gl, gr, gml, gmr, sl, sr, Hd, Hdv, Hdv_back, Hld,
                               Hrd, Hrd_fwd, Haml, Hamr, transl, transr, G_D_RL, G_D_LR, G_L, G_R);
    });
}

int nt_get_ldos(nt_ctx* c, const nt_params* p, int nl, int nr, int L, const int* layer_list, double E,
                const nt_c128* gl, const nt_c128* gr, const nt_c128* const* Hd, const nt_c128* const* Hdv,
                const nt_c128* const* Hdv_back, const nt_c128* Hld, const nt_c128* Hrd, const nt_c128* Hrd_fwd,
                int maxlen, double* ldos, double* ddos)
{
    if (!c) return NT_ERR_ARG;
    if (check_params(p) || nl <= 0 || nr <= 0 || L <= 0 || !layer_list || !Hd || (L > 1 && !Hdv)) return NT_ERR_ARG;
    if (!gl || !gr || !Hld || !Hrd || !ldos) return NT_ERR_ARG;
    int mx = std::max(nl, nr), bl = 0;
    for (int i = 0; i < L; ++i) {
        if (layer_list[i] <= 0 || !Hd[i] || (i + 1 < L && !Hdv[i])) return NT_ERR_ARG;
        mx = std::max(mx, layer_list[i]);
        bl = std::max(bl, layer_list[i]);
    }
    if (mx > inv_max_n() || maxlen < bl) return NT_ERR_ARG;
    return guarded(c, [&] { return ldos_host(c, p, nl, nr, L, layer_list, E, gl, gr, Hd, Hdv, Hdv_back, Hld, Hrd,
                                             Hrd_fwd, maxlen, ldos, ddos); });
}

int nt_autotune_inverse(nt_ctx* c, int n, int batch, int precision, int* best_nb, double* best_ms)
{
    if (!c || n <= 0 || n > inv_max_n() || batch <= 0) return NT_ERR_ARG;
    if (precision != NT_PREC_FP64 && precision != NT_PREC_MIXED) return NT_ERR_ARG;
    return guarded(c, [&] {
        if (precision == NT_PREC_FP64)
            tune_inverse<double>(c, n, batch, best_nb, best_ms);
        else
            tune_inverse<float>(c, n, batch, best_nb, best_ms);
        return 0;
    });
}

int nt_autotune_gemm(nt_ctx* c, int n, int batch, int precision, int* best_tile, double* best_ms)
{
    if (!c || n <= 0 || batch <= 0) return NT_ERR_ARG;
    if (precision != NT_PREC_FP64 && precision != NT_PREC_MIXED) return NT_ERR_ARG;
    return guarded(c, [&] {
        if (precision == NT_PREC_FP64)
            tune_gemm<double>(c, n, batch, best_tile, best_ms);
        else
            tune_gemm<float>(c, n, batch, best_tile, best_ms);
        return 0;
    });
}

int nt_set_gemm_tile(nt_ctx* c, int n, int precision, int tile)
{
    if (!c || n <= 0 || (tile != 0 && tile != 32 && tile != 64)) return NT_ERR_ARG;
    if (precision != NT_PREC_FP64 && precision != NT_PREC_MIXED) return NT_ERR_ARG;
    const int bytes = precision == NT_PREC_FP64 ? 8 : 4;
    if (tile == 0)
        c->gemm_tile.erase({bytes, n});
    else
        c->gemm_tile[{bytes, n}] = tile;
    return NT_OK;
}

int nt_set_gemm_3m(nt_ctx* c, int on)
{
    if (!c || (on != 0 && on != 1)) return NT_ERR_ARG;
    c->gemm_m3 = on == 1;
    return NT_OK;
}

int nt_set_adjoint_checkpoint(nt_ctx* c, int every)
{
    if (!c) return NT_ERR_ARG;
    c->adj_ckpt = every;
    return NT_OK;
}

int nt_adjoint_bytes_per_point(nt_ctx* c, int nl, int nr, int b, int L, int lo, int hi, long long* bytes)
{
    if (!c || !bytes || nl <= 0 || nr <= 0 || b <= 0 || L <= 0 || lo < 0 || hi < lo || hi >= L) return NT_ERR_ARG;
    const long mm = (long)std::max(std::max(nl, nr), b) * std::max(std::max(nl, nr), b);
    *bytes = (long long)adj_slots(lo, hi, L, adj_interval(c->adj_ckpt, lo, L)) * mm * (long long)sizeof(nt_c128);
    return NT_OK;
}

int nt_set_sweeps(nt_ctx* c, int n)
{
    if (!c || (n != 1 && n != 2)) return NT_ERR_ARG;
    c->one_sweep = n == 1;
    return NT_OK;
}

int nt_set_decimation_check(nt_ctx* c, int every)
{
    if (!c || every < 0) return NT_ERR_ARG;
    c->dec_check = every;
    return NT_OK;
}

int nt_set_overlap(nt_ctx* c, int on)
{
    if (!c || (on != 0 && on != 1)) return NT_ERR_ARG;
    c->overlap = on == 1;
    return NT_OK;
}

int nt_set_workspace_poison(nt_ctx* c, int on)
{
    if (!c || (on != 0 && on != 1)) return NT_ERR_ARG;
    c->ws_poison = on == 1;
    return NT_OK;
}

int nt_set_inverse_panel(nt_ctx* c, int n, int precision, int nb)
{
    if (!c || n <= 0 || n > inv_max_n()) return NT_ERR_ARG;
    if (precision != NT_PREC_FP64 && precision != NT_PREC_MIXED) return NT_ERR_ARG;
    const int bytes = precision == NT_PREC_FP64 ? 8 : 4;
    if (nb == 0) {
        c->inv_nb.erase({bytes, n});
        return NT_OK;
    }
    int w[INV_MAX_WIDTHS];
    const int k = precision == NT_PREC_FP64 ? inv_panel_widths<double>(n, w) : inv_panel_widths<float>(n, w);
    for (int i = 0; i < k; ++i)
        if (w[i] == nb) {
            c->inv_nb[{bytes, n}] = nb;
            return NT_OK;
        }
    return NT_ERR_ARG;
}

}  // extern "C"
