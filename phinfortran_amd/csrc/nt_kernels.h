// nt_kernels.h -- launch interfaces of the batched transport kernels.
#pragma once

#include "nt_device.h"

namespace nt {

enum Op { OP_N = 0, OP_C = 1 };

// C[p,g] = alpha * op(A[p,g]) * op(B[p,g]) + beta * Cin[p,g] + diag[p] * I
// C2[p,g] += alpha * op(A) * op(B)          (optional second accumulator)
// maxabs[p] = max(maxabs[p], max |C[p,0]|)  (optional, group 0 only)
template <typename T> struct GemmArgs {
    int M, N, K;
    int opA, opB;
    BMat<T> A, B, Cin;
    BMatW<T> C, C2;
    T alpha_re, alpha_im, beta_re, beta_im;
    const T* diag;          // interleaved complex per item p (stride 2), or null
    const int* active;      // per item p, or null
    unsigned long long* maxabs;  // per item p (bits of a non-negative double), or null
    int batch, groups;
    int m3 = 0;             // 1: three real products per complex product
    int tile = 0;           // output tile edge: 64 (0 = default) or 32
};

template <typename T> hipError_t launch_gemm(const GemmArgs<T>& a, hipStream_t s);

// Batched in-place Gauss-Jordan inverse with partial (row) pivoting.
//   W[p]  n x n work matrix, destroyed.  If src.base is set, the kernel first
//         forms W = z[p] * I - src[p].
//   X[p]  receives inv(W).  Must not alias W.
//   info[p]: left untouched on success; if it is 0 and the inverse hits an exact
//   zero pivot at column k it becomes k (1-based), -1 for NaN input.
template <typename T> struct InvArgs {
    int n;
    BMatW<T> W;
    BMat<T> src;
    const T* z;             // interleaved complex per item, used with src
    BMatW<T> X;
    const int* active;
    int* info;
    int batch;
    int nb = 0;             // panel width (8, 16, 32); 0 or unsupported for n: size rule
};

template <typename T> hipError_t launch_inv(const InvArgs<T>& a, hipStream_t s);
int inv_max_n();
// panel widths launch_inv accepts for an n x n matrix of type T (largest first);
// INV_BLOCKED selects the panel-kernel + batched-GEMM path (last in the list)
constexpr int INV_BLOCKED = 1;
constexpr int INV_MAX_WIDTHS = 6;
template <typename T> int inv_panel_widths(int n, int* out);

// Batched complex Schur / eigen / square-root kernels (nt_eig.hip), FP64.
struct EigArgs {
    int n;
    BMatW<double> A;     // in: matrix; out: Schur form T (upper triangular)
    BMatW<double> Z;     // Schur vectors (A = Z T Z^H)
    BMatW<double> X;     // trevc: upper triangular eigenvectors of T; sqrt_tri: sqrt(T)
    BMatW<double> V;     // ev_normalize: eigenvectors normalised in place
    double* w;           // optional eigenvalues (n complex per item)
    int* status;         // merged: first error of an item wins (-3: QR did not converge)
    int batch;
};
hipError_t launch_schur(const EigArgs& a, hipStream_t s);
hipError_t launch_trevc(const EigArgs& a, hipStream_t s);
hipError_t launch_ev_normalize(const EigArgs& a, hipStream_t s);
hipError_t launch_sqrt_tri(const EigArgs& a, hipStream_t s);
// propagating-mode selection, k_z and Gram-Schmidt (nt_eig.hip)
struct ModesArgs {
    int n;
    const double* w;       // eigenvalues, n complex per item
    BMatW<double> U;       // eigenvectors, propagating columns orthonormalised in place
    double* kz;            // n per item (100 for evanescent modes)
    int* nprop;            // optional count per item
    double lambda_eta, az, sign;
    int batch;
};
hipError_t launch_modes_select(const ModesArgs& a, hipStream_t s);
// M[p] (m x n, leading dim ld, item stride in complex elements) *= sc[p]
hipError_t launch_scale_items(double* M, long stride, int ld, int m, int n, const double* sc, int batch,
                              hipStream_t s);
hipError_t launch_mode_sums(const double* S, long stride, int ld, int m, int n, const double* kz_row,
                            const double* kz_col, double* col_sum, double* total, int batch, hipStream_t s);

// Brillouin-zone unfolding of lead modes (nt_unfold.hip).  Weights of one
// mode set: label bz (1-based k-shift of largest weight, 0 for evanescent
// modes) and unfolded k (column-major n x 3, 100 for evanescent modes).
struct UnfoldDirArgs {
    int n, nuc, nbx, nby;     // modes, orbitals per transverse cell, k-shift grid
    const double* U;          // eigenvectors (complex), item stride ustride, leading dim ldu
    long ustride;
    int ldu;
    const double* w;          // eigenvalues, n complex per item
    const double* kz;         // n per item (>= 100: evanescent)
    const double* ksc;        // 3 per k-point, crystal coordinates of recivec
    const int* kidx;          // k-point of each item, or null (k-point 0)
    double recivec[9], reci_uc[9];   // row-major reciprocal vectors
    const double* pos;        // n x 3 orbital positions
    const double* pos_uc;     // nuc x 3 transverse-cell atom of each first-image orbital
    int* bz;                  // n per item
    double* kxyz;             // 3 n per item
    int batch;
};
// One t or r matrix (outgoing x incident) with the labels and unfolded k of
// its two mode sets; rows: optional (total, specular) per incident mode.
struct UnfoldChain {
    const double* M;
    int mo, mi;
    const int* bz_row;
    const int* bz_col;
    const double* k_row;
    const double* k_col;
    double* rows;
    int slot;                 // split pair index (0..3)
};
struct UnfoldSplitArgs {
    UnfoldChain ch[4];
    double* split;            // 8 per item, or null
    int batch;
};
hipError_t launch_unfold_weights(const UnfoldDirArgs& a, hipStream_t s);
hipError_t launch_unfold_split(const UnfoldSplitArgs& a, hipStream_t s);

// Primitive-cell unfolding (nt_kpc.hip).  Mode m: E[m], k[3m..3m+2]; chunk
// arrays (count, cand) are indexed from the chunk's first mode; outputs kp, v
// (3 per mode) and ncand by mode.
constexpr double KPC_DEVI = 1.0e-5;      // central-difference step (dyn.f90:548)
constexpr int KPC_MAX_SWEEPS = 30;
constexpr int KPC_MAX_N = 48;           // orbitals of the primitive cell
struct KpcArgs {
    int n;                    // 3 x atoms of the primitive cell (orbitals: tight binding)
    int tb;                   // 1: tight-binding conventions (find_k_pc_l, tb/hamiltonian.f90:666-822)
    const double* E;
    const double* k;
    double reci_uc[9], recivec_pc[9];
    int l, maxc;              // candidates k + n . reci_uc, n in [-l, l]^3; maxc = (2l+1)^3
    int* count;               // per chunk mode: candidates, -1 evanescent, -2 the mode's own k
    double* cand;             // per chunk mode: maxc x 3
    double* kp;
    double* v;
    int* ncand;
};
hipError_t launch_kpc_candidates(const KpcArgs& a, int m0, int M, hipStream_t s);
hipError_t launch_kpc_klist(const KpcArgs& a, const int* cm, const int* cs, int c0, int nc, double* kl, double* kc,
                            hipStream_t s);
size_t kpc_eig_smem(int n);
hipError_t launch_kpc_eig(const KpcArgs& a, const int* cm, int m0, int c0, int nc, const double* D, int* found,
                          double* vout, hipStream_t s);
hipError_t launch_kpc_select(const KpcArgs& a, int m0, int M, const int* off, const int* found, const double* vc,
                             const double* kc, hipStream_t s);
// tight binding: H(k) and the three velocity sums of every candidate from an
// element-wise table (slots 0..3 of the 7-slot layout), candidate k into kc
hipError_t launch_kpc_tb_mats(const KpcArgs& a, const int* cm, const int* cs, int c0, int nc, const int* el_off,
                              const double* coef, const double* dr, double* D, double* kc, hipStream_t s);

// H(k) phase sums of the QE force-constant path (nt_hk.hip): out[k] (rows x cols,
// leading dimension ld, k stride kstride complex elements) = sum over the
// entries of each 3x3 segment of coef * exp(i dr . k); zero where no segment.
hipError_t launch_hk_assemble(int nk, const double* kc, int nseg, const int* seg_off, const int* seg_rc,
                              const double* coef, const double* dr, int rows, int cols, double* out, int ld,
                              long kstride, hipStream_t s);
// element -> segment maps of a grid-tiled sub-block table (nt_hk.hip)
struct HkGrid {
    const int *rblk, *rloc, *cblk, *cloc, *rptr, *ckey, *sid;
};
hipError_t launch_hk_assemble_orb(int nk, const double* kc, const int* seg_off, const int* seg_geom, int cstride,
                                  const double* coef, const double* dr, const HkGrid& g, int rows, int cols,
                                  double* out, int ld, long kstride, int hermitize, hipStream_t s);

// Degenerate-subspace rotation of the mode-resolved solve (nt_degen.hip).
constexpr int DEG_MAX = 16;       // modes in one degenerate group
struct DegGroup {
    int p, ndeg;                  // point, group size
    int idx[DEG_MAX];             // 0-based mode indices, ascending
    double lr, li;                // eigenvalue of the group's first mode
};
struct DegArgs {
    int m, G;                     // mode-set size, groups
    BMat<double> H00, V;          // lead blocks of each point
    const DegGroup* g;
    const double* E;              // per point
    double* H;                    // per group m x m (stride hstride): built matrix, then its Schur form
    const double* Z;              // per group Schur vectors (stride hstride)
    long hstride;
    double* U;                    // per point m x m mode eigenvectors (stride ustride), rotated in place
    long ustride;
    const int* gstatus;           // per group Schur status
    int* status;                  // per point
};
hipError_t launch_deg_build(const DegArgs& a, hipStream_t s);
hipError_t launch_deg_solve(const DegArgs& a, hipStream_t s);
// every column with kz < 99.9 normalised (per point m x m, stride ustride, kz m per point)
hipError_t launch_deg_normalize(int m, double* U, long ustride, const double* kz, int P, hipStream_t s);

// Buttiker probes (nt_probe.hip).
// zl[i][p] = zd[p] - sigma_w(i, E[p]) (zeroed sigma when !lifetime), gam / dgam
// = -2 Im sigma_w, -2 Im dsigma_w/dT; layer-major L x P.
hipError_t launch_probe_shifts(const double* E, const double* zd, const double* temp, int L, int P, double B,
                               double C, int lifetime, double* zl, double* gam, double* dgam, hipStream_t s);
// out + 2 (p ostride + g ogstride) = Tr(X^+ Y) (complex) of m x n blocks
template <typename T>
hipError_t launch_frob_dot(int m, int n, BMat<T> X, BMat<T> Y, double* out, long ostride, long ogstride, int batch,
                           int groups, hipStream_t s);
// out + 2 (p ostride + g ogstride) = Tr X (complex) of n x n blocks
template <typename T>
hipError_t launch_trace_diag(int n, BMat<T> X, double* out, long ostride, long ogstride, int batch, int groups,
                             hipStream_t s);
struct ProbeCombineArgs {
    int L, P, p0;             // layers, points in the chunk, chunk offset of the outputs
    const double* E;          // chunk energies
    const double* temp;       // L probe temperatures
    double t_left, t_right;
    const double* gam;        // L x P
    const double* dgam;       // L x P
    const double* acc;        // per point: L^2 + 5L + 6 complex traces
    long acc_stride;          // complex entries per point
    const int* first_mask;    // chunk mask, or null
    double *flux, *jac, *flux_left, *flux_right, *a_sum, *gn_sum;   // indexed from point p0 + p
};
hipError_t launch_probe_combine(const ProbeCombineArgs& a, hipStream_t s);

// Elementwise helpers (nt_elem.hip).
// out[p] = sign * (dagger ? M^+ : M) for an n x n matrix
template <typename T>
hipError_t launch_copy_op(int m, int n, BMat<T> in, BMatW<T> out, int dagger, T sign,
                          const int* active, int batch, hipStream_t s);
// out[p] = sign * (dagger ? M^+ : M) converted between precisions (m x n input)
template <typename Ti, typename To>
hipError_t launch_convert(int m, int n, BMat<Ti> in, BMatW<To> out, int dagger, To sign, int batch, hipStream_t s);
// Gamma[p] = scale * i (S - S^+)
template <typename T>
hipError_t launch_gamma(int n, BMat<T> S, BMatW<T> G, T scale, const int* active, int batch, hipStream_t s);
// out[p] = z[p] I - A[p] - B[p]   (B optional)
template <typename T>
hipError_t launch_shift_sub(int n, const T* z, BMat<T> A, BMat<T> B, BMatW<T> out, const int* active,
                            int batch, hipStream_t s);
// Decimation start: e1 = es = H00, alpha = opA(V), beta = opB(V)
template <typename T>
hipError_t launch_dec_init(int n, BMat<T> H00, BMat<T> V, int alpha_dagger, BMatW<T> e1, BMatW<T> es,
                           BMatW<T> ab, int batch, hipStream_t s);
// T[p] = Re sum_ij X_ij Y_ji over n1 x n2 X and n2 x n1 Y (into out[p] * out_stride)
template <typename T>
hipError_t launch_trace_prod(int n1, int n2, BMat<T> X, BMat<T> Y, double* out, int out_stride,
                             const int* active, int batch, hipStream_t s);
// Im trace of g (for surface DOS), out[p] = coef * sqrt(E[p]) * Im Tr g / pi
template <typename T>
hipError_t launch_trace_im(int n, BMat<T> g, const double* E, double coef, double* out, int batch,
                           hipStream_t s);

// Shifts of one (omega,k) point: z arrays interleaved (re, im), length 2*batch.
//   zl  = E - i eta(E)     (left advanced / right reflected)
//   zr  = E + i eta(E)     (right retarded / left reflected)
//   zd  = E + i f eta(E)   (device and closing lead layers)
hipError_t launch_make_shifts(const double* E, int batch, double eta, double eta0, int eta_mode,
                              double dev_factor, double* zl, double* zr, double* zd, float* zl32,
                              float* zr32, float* zd32, hipStream_t s);
// active[p] &= (maxabs[p] > conv); iters[p] += was_active (iters may be null);
// maxabs[p] = 0; adds the number still active into *count and sets *clear = 0
// (the two words of a two-slot counter: update u adds to slot u % 2 and clears
// the other, so after any update its slot holds that update's count)
hipError_t launch_dec_step_update(int* active, unsigned long long* maxabs, int* iters, double conv, int* count,
                                  int* clear, int batch, hipStream_t s);
hipError_t launch_fill_int(int* p, int v, int n, hipStream_t s);
// Final transmission clean-up: clamp < 1e-60 to 0, zero when flagged, status merge.
hipError_t launch_finish_T(double* T, int n, int zero_all, hipStream_t s);

}  // namespace nt
