/*
 * nt_oracle.h -- CPU FP64 restatement of phinfortran's per-(omega,k) transport solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine in
 * phinfortran_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * All matrices are complex FP64, column-major, leading dimension = rows,
 * exactly like the reference's complex(kind=8) arrays.
 *
 * Pinning of each function (what fixes its result independently of this file):
 *   or_inv             -- A*inv(A) = I and numpy.linalg.inv on random/ill-posed inputs.
 *   or_decimate        -- closed-form surface GF of the 1-D chain (monatomic and
 *                         TB chain), Dyson fixed point g = (z-H00-V^+ g V)^-1 on
 *                         random blocks, numpy brute-force semi-infinite truncation.
 *   or_surface_green   -- as or_decimate plus Gamma Hermitian PSD and the
 *                         identities Sigma = V^+ g V of the reference.
 *   or_rgf             -- dense brute-force inverse of the whole device
 *                         (lead layers + device) for L<=6, b<=12 at uniform eta;
 *                         T_LR = T_RL; T = 1 in-band for the chain; integer channel
 *                         counts for pristine multi-orbital leads; the si-si
 *                         golden T tables of the reference example.
 *   or_circular_g      -- equality with or_decimate on Bloch-periodic leads.
 */
#ifndef NT_ORACLE_H
#define NT_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex ocplx;

/* Lead broadening conventions (negf/surface.f90:479 vs tb/surface.f90:218). */
enum { OR_ETA_RELATIVE = 0, OR_ETA_ABSOLUTE = 1 };

typedef struct {
    double eta;          /* relative (phonon) or absolute (TB) broadening */
    double eta0;         /* absolute part for the relative mode */
    double convergence;  /* decimation stops when max|alpha| <= convergence */
    double device_eta_factor; /* 10 for phonon (lesser.f90:556), 1 for TB */
    int eta_mode;        /* OR_ETA_RELATIVE or OR_ETA_ABSOLUTE */
    int max_iter;        /* decimation iteration cap (reference has none) */
} or_params;

/* Status codes shared with the engine: 0 ok, >0 singular pivot (1-based),
 * -1 NaN input, -2 decimation did not converge within max_iter. */

double or_lead_eta(const or_params* p, double E);

/* inv: ZGETRF+ZGETRI equivalent (negf/util.f90:16-67). */
int or_inv(int n, const ocplx* A, ocplx* Ainv);

/* C = alpha*op(A)*op(B) + beta*C, op in {'N','C'} (conjugate transpose). */
void or_gemm(char opa, char opb, int m, int n, int k, ocplx alpha,
             const ocplx* A, int lda, const ocplx* B, int ldb,
             ocplx beta, ocplx* C, int ldc);

/* One Lopez-Sancho decimation (negf/surface.f90:488-497):
 * returns g = (z - es)^-1, iterations in *iters. */
int or_decimate(int n, const ocplx* H00, const ocplx* alpha0, const ocplx* beta0,
                ocplx z, double conv, int max_iter, ocplx* g, int* iters);

/* surface_green_tdb (negf/surface.f90:385-605) / tb surface_green
 * (tb/surface.f90:129-343).  Lead couplings: Vl = Hamlv (left lead, layer -1 to 0),
 * Vr = Hamrv (right lead, layer N+2 to N+1).  Any output may be NULL except
 * the ones needed to form the others.  iters[4] (may be NULL) receives the
 * iteration counts of gl_adv, gl_plus, gr_ret, gr_minus. */
int or_surface_green(int nl, int nr, const ocplx* Haml, const ocplx* Hamlv,
                     const ocplx* Hamr, const ocplx* Hamrv, double E,
                     const or_params* p, int want_plus_minus,
                     ocplx* gl_adv, ocplx* gr_ret, ocplx* gl_plus, ocplx* gr_minus,
                     ocplx* gml_adv, ocplx* gmr_ret, ocplx* sl_adv, ocplx* sr_ret,
                     int* iters);

/* RGF sweep + Caroli (negf/lesser.f90:475-809 for uniform blocks,
 * negf/device.f90:289-560 and tb/device.f90:291-575 for variable blocks).
 *   bsz[L]       layer sizes
 *   Hd[i]        bsz[i] x bsz[i] diagonal blocks
 *   Hdv[i]       bsz[i] x bsz[i+1] couplings V_{i,i+1} (i < L-1)
 *   Hdv_back[i]  bsz[i+1] x bsz[i] couplings V_{i+1,i} used by the right-to-left
 *                sweep, or NULL (-> Hdv[i]^+, the tdb convention)
 *   Hld          nl x bsz[0]    V_{0,1} (both sweeps use it and its dagger)
 *   Hrd          nr x bsz[L-1]  V_{N+1,N} (right-to-left sweep)
 *   Hrd_fwd      bsz[L-1] x nr  V_{N,N+1} used to close the left-to-right sweep,
 *                or NULL (-> Hrd^+, the tdb convention)
 *   gl, sl, gml  retarded left surface GF, self energy and Gamma (the reference
 *                passes gl_adv^+, sl_adv^+, -gml_adv)
 *   gr, sr, gmr  retarded right surface GF, self energy and Gamma
 * Outputs (any may be NULL): G_D_LR (nl x nr), G_D_RL (nr x nl), G_L (nl x nl),
 * G_R (nr x nr), T_LR, T_RL. */
int or_rgf(int L, const int* bsz, const ocplx* const* Hd, const ocplx* const* Hdv,
           const ocplx* const* Hdv_back, int nl, int nr,
           const ocplx* Hld, const ocplx* Hrd, const ocplx* Hrd_fwd,
           const ocplx* Haml, const ocplx* Hamr,
           const ocplx* gl, const ocplx* gr, const ocplx* gml, const ocplx* gmr,
           const ocplx* sl, const ocplx* sr, double E, const or_params* p,
           int zero_transmission,
           ocplx* G_D_LR, ocplx* G_D_RL, ocplx* G_L, ocplx* G_R,
           double* T_LR, double* T_RL);

/* Whole (omega,k)-solve on uniform tdb blocks: surface GFs (2 decimations, or 4
 * when want_plus_minus), adv->ret conversion of get_G_full_tdb_mpi
 * (negf/lesser.f90:141-154) and the RGF/Caroli of get_G_full_tdb.
 *   Hamd  b x b x L (layer-major), Hamdv b x b x (L-1).
 * Returns T in T_LR/T_RL and the mean decimation iterations in *iters_mean. */
int or_solve_tdb(int nl, int nr, int b, int L,
                 const ocplx* Haml, const ocplx* Hamlv, const ocplx* Hamr, const ocplx* Hamrv,
                 const ocplx* Hamd, const ocplx* Hamdv, const ocplx* Hamld, const ocplx* Hamrd,
                 double E, const or_params* p, int zero_transmission,
                 double* T_LR, double* T_RL, double* iters_mean);

/* Batch of independent solves over energies on the same blocks, OpenMP over
 * points with nthreads threads (<=0: default).  Used for the CPU baseline. */
int or_solve_tdb_batch(int nl, int nr, int b, int L,
                       const ocplx* Haml, const ocplx* Hamlv, const ocplx* Hamr, const ocplx* Hamrv,
                       const ocplx* Hamd, const ocplx* Hamdv, const ocplx* Hamld, const ocplx* Hamrd,
                       int npts, const double* E, const or_params* p, int nthreads,
                       double* T_LR, double* T_RL, int* status);
/* or_solve_tdb_batch plus kappa[i]: the largest 1-norm condition number of
 * the blocks point i's solve inverted (the rounding scale of the result). */
int or_solve_tdb_batch_kappa(int nl, int nr, int b, int L,
                             const ocplx* Haml, const ocplx* Hamlv, const ocplx* Hamr, const ocplx* Hamrv,
                             const ocplx* Hamd, const ocplx* Hamdv, const ocplx* Hamld, const ocplx* Hamrd,
                             int npts, const double* E, const or_params* p, int nthreads,
                             double* T_LR, double* T_RL, int* status, double* kappa);

/* Block-circulant decimation (negf/circular.f90:43-114): the lead supercell is
 * nbx*nby transverse copies; P1/P2 are unitary block DFTs.  Same result as
 * or_decimate for Bloch-periodic leads. */
int or_circular_g(int n, int nby, int nbx, const ocplx* H00, const ocplx* alpha0,
                  const ocplx* beta0, ocplx z, double conv, int max_iter,
                  ocplx* g, int* iters);

/* get_ldos (tb/device.f90:610-855): per-layer LDOS -Im GLR_i(j,j)/pi of the
 * device at z = E + i*factor*eta(E), ldos column-major (L x maxlen), ddos the
 * sum.  Pinned against the diagonal of the dense device Green's function. */
int or_get_ldos(int L, const int* bsz, const ocplx* const* Hd, const ocplx* const* Hdv,
                const ocplx* const* Hdv_back, int nl, int nr,
                const ocplx* Hld, const ocplx* Hrd, const ocplx* Hrd_fwd,
                const ocplx* gl, const ocplx* gr, double E, const or_params* p,
                int maxlen, double* ldos, double* ddos);

#ifdef __cplusplus
}
#endif
#endif
