/*
 * negf_transport.h -- C ABI of the MI355X transport engine.
 *
 * Drop-in boundary for the per-(omega,k) solve of phinfortran.  Every entry
 * point takes plain pointers and sizes; complex matrices are column-major
 * complex(kind=8) (nt_c128 = {re, im}, same bytes as Fortran
 * complex(c_double_complex)), leading dimension = number of rows.
 *
 * Status codes (all entry points):
 *    0  success
 *   >0  exact zero pivot while inverting a block (LAPACK-style column, 1-based);
 *       the reference stops with 'Matrix is numerically singular!'
 *       (negf/util.f90:44-58)
 *   -1  NaN in a block that is inverted (negf/util.f90:29-32)
 *   -2  surface decimation did not converge within max_iter (the reference
 *       loops forever, negf/surface.f90:488)
 *   -3  invalid argument (sizes, null pointers, unsupported block size)
 *   -4  HIP runtime error (see nt_last_error)
 *   -5  out of device memory
 */
#ifndef NEGF_TRANSPORT_H
#define NEGF_TRANSPORT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } nt_c128;
typedef struct nt_ctx nt_ctx;

enum {
    NT_OK = 0,
    NT_ERR_NAN = -1,
    NT_ERR_NOT_CONVERGED = -2,
    NT_ERR_ARG = -3,
    NT_ERR_HIP = -4,
    NT_ERR_NOMEM = -5
};

/* Broadening conventions. */
enum {
    NT_ETA_RELATIVE = 0,   /* eta(E) = E*eta + eta0  (negf/surface.f90:479) */
    NT_ETA_ABSOLUTE = 1    /* eta(E) = eta           (tb/surface.f90:218)   */
};

/* Arithmetic of the RGF sweep. */
enum {
    NT_PREC_FP64 = 0,      /* complex FP64 throughout (the reference's arithmetic) */
    NT_PREC_MIXED = 1      /* FP64 decimation; each sweep direction factorised in complex FP32 on
                              MFMA, its block column refined against FP64 residuals (one or two
                              steps); points failing the error estimate recomputed in FP64 */
};

/* Scalar parameters that the reference keeps as module globals
 * (negf/config.f90:58-68, negf/surface.f90:5-7). */
typedef struct {
    double eta;               /* &convergelist eta  */
    double eta0;              /* &convergelist eta0 */
    double convergence;       /* &convergelist convergence: stop when max|alpha| <= convergence */
    double device_eta_factor; /* 10 in negf (lesser.f90:556, device.f90:348), 1 in tb */
    double lambda_eta;        /* &convergelist lambda_eta (propagating-mode test) */
    int eta_mode;             /* NT_ETA_RELATIVE (negf) or NT_ETA_ABSOLUTE (tb) */
    int max_iter;             /* decimation iteration cap, >= 1 */
    int precision;            /* NT_PREC_FP64 or NT_PREC_MIXED */
    int circular;             /* &unfolding circular_flag: block-circulant lead decimation */
    int n_bloch_x;            /* &unfolding n_bloch_x: lead copies along x (inner) */
    int n_bloch_y;            /* &unfolding n_bloch_y: lead copies along y (outer) */
    int path_mode;            /* &kpath path_mode: 1 adds the degenerate-subspace rotation of
                                 the mode-resolved solve (negf/velocity_mat.f90:1884-1978);
                                 any other value runs as the example's path_mode = 2 */
    int dos_mode;             /* surface DOS convention of the sdos_l / sdos_r outputs:
                                 NT_DOS_FROM_ETA (default) follows eta_mode (relative: negf,
                                 absolute: tb), NT_DOS_NEGF and NT_DOS_TB choose explicitly */
} nt_params;

/* Surface DOS conventions (nt_params.dos_mode). */
enum {
    NT_DOS_FROM_ETA = 0,   /* NT_DOS_NEGF when eta_mode is relative, NT_DOS_TB when absolute */
    NT_DOS_NEGF = 1,       /* +-2 sqrt(E) Im Tr g / pi, E = omega^2   (negf/surface.f90:608-630) */
    NT_DOS_TB = 2          /* +-Im Tr g / pi                          (tb/surface.f90:345-366)   */
};

/* Fills the negf defaults: eta = eta0 = 1e-6 (config.f90:88-89), device x10,
 * convergence 1e-12, lambda_eta 5e-3, max_iter 300, FP64, direct decimation
 * (circular = 0, n_bloch_x = n_bloch_y = 1), path_mode 2, dos_mode NT_DOS_FROM_ETA.
 * With circular = 1 the lead blocks (n x n) must be block-circulant over
 * n_bloch_y x n_bloch_x transverse copies (copy index y outer, x inner); each
 * lead GF is then computed as in circular_g (negf/circular.f90:43-114): a
 * block DFT of the first block row, n_bloch_x*n_bloch_y independent
 * decimations of size n/(n_bloch_x*n_bloch_y), and the back transform. */
void nt_params_default(nt_params* p);

/* NT_OK if p is usable for leads of nl and nr orbitals, else NT_ERR_ARG; every
 * entry point runs the same check first.  No reference counterpart: it stands
 * for the reference's namelist checks (negf/config.f90, tb/config.f90), which
 * stop the program instead of returning. */
int nt_params_validate(const nt_params* p, int nl, int nr);

/* ------------------------------------------------------------ context */
int nt_create(nt_ctx** ctx, int device_ordinal);
int nt_destroy(nt_ctx* ctx);
int nt_sync(nt_ctx* ctx);
/* HIP stream the context launches on (hipStream_t as void*). */
void* nt_stream(nt_ctx* ctx);
/* Use an external stream instead of the context's own one (NULL restores it). */
int nt_set_stream(nt_ctx* ctx, void* stream);
/* Upper bound on (omega,k) points per device batch (default 512). */
int nt_set_batch_limit(nt_ctx* ctx, int max_points);
const char* nt_last_error(nt_ctx* ctx);
const char* nt_status_string(int status);

/* ------------------------------------------- per-point host-buffer ABI
 * These mirror the reference subroutines one-to-one so a Fortran caller can
 * swap a call for a bind(C) call.  Inputs/outputs are host arrays. */

/* surface_green_tdb (negf/surface.f90:385-605) and tb surface_green
 * (tb/surface.f90:129-343, with params->eta_mode = NT_ETA_ABSOLUTE).
 *   Haml, Hamlv  nl x nl: left lead H00 and coupling V_{-1,0}
 *   Hamr, Hamrv  nr x nr: right lead H00 and coupling V_{N+2,N+1}
 * Outputs (nl x nl / nr x nr), any may be NULL except as noted:
 *   gl_adv, gr_ret, gl_plus, gr_minus, gml_adv, gmr_ret, sl_adv, sr_ret.
 *   gl_plus/gr_minus are only computed when non-NULL (they feed the
 *   mode-resolved solve only).
 * iters (int[4], may be NULL): decimation iterations of gl_adv, gl_plus,
 * gr_ret, gr_minus. */
int nt_surface_green_tdb(nt_ctx* ctx, const nt_params* p, int nl, int nr, double E,
                         const nt_c128* Haml, const nt_c128* Hamr,
                         const nt_c128* Hamlv, const nt_c128* Hamrv,
                         nt_c128* gl_adv, nt_c128* gr_ret, nt_c128* gl_plus, nt_c128* gr_minus,
                         nt_c128* gml_adv, nt_c128* gmr_ret, nt_c128* sl_adv, nt_c128* sr_ret,
                         int* iters);

/* get_G_full_tdb (negf/lesser.f90:475-809) as called by get_G_full_tdb_mpi
 * (lesser.f90:141-154): gl, gml, sl are the RETARDED left quantities
 * (gl_adv^+, -gml_adv, sl_adv^+), gr, gmr, sr the right ones.
 *   hamd   b x b x L  (layer i at hamd + i*b*b)
 *   hamdv  b x b x (L-1), V_{i,i+1}
 *   hamld  nl x b (V_{0,1}), hamrd nr x b (V_{N+1,N}), haml nl x nl, hamr nr x nr
 *   first_point != 0 reproduces the reference's T = 0 at ie = 1
 *   (lesser.f90:802-807).
 * Outputs: transl (T_LR), transr (T_RL); G_D_RL (nr x nl), G_D_LR (nl x nr),
 * G_L (nl x nl), G_R (nr x nr) may be NULL.  The Buttiker-probe bookkeeping
 * (GLeft, G_full, G_q_qp, ...) is not produced here; the probe path is
 * nt_buttiker_tdb_batched_dev. */
int nt_get_G_full_tdb(nt_ctx* ctx, const nt_params* p, int first_point,
                      int nl, int nr, int b, int L, double E,
                      const nt_c128* gl, const nt_c128* gr, const nt_c128* gml, const nt_c128* gmr,
                      const nt_c128* sl, const nt_c128* sr,
                      const nt_c128* hamd, const nt_c128* hamdv, const nt_c128* hamld,
                      const nt_c128* hamrd, const nt_c128* haml, const nt_c128* hamr,
                      double* transl, double* transr,
                      nt_c128* G_D_RL, nt_c128* G_D_LR, nt_c128* G_L, nt_c128* G_R);

/* G_block (negf/device.f90:289-560, tb/device.f90:291-575): the same sweep on
 * variable layer sizes layer_list[L] cut from a dense Hamiltonian.
 *   Hd[i]       layer_list[i]^2 diagonal blocks
 *   Hdv[i]      layer_list[i] x layer_list[i+1], V_{i,i+1} (left-to-right sweep)
 *   Hdv_back[i] layer_list[i+1] x layer_list[i], V_{i+1,i} (right-to-left), or NULL
 *               to use Hdv[i]^+
 *   Hld         nl x layer_list[0], V_{0,1}
 *   Hrd         nr x layer_list[L-1], V_{N+1,N}
 *   Hrd_fwd     layer_list[L-1] x nr, V_{N,N+1}, or NULL to use Hrd^+
 *   Haml, Hamr  lead H00 of the lead layers next to the device. */
int nt_G_block(nt_ctx* ctx, const nt_params* p, int nl, int nr, int L, const int* layer_list, double E,
               const nt_c128* gl, const nt_c128* gr, const nt_c128* gml, const nt_c128* gmr,
               const nt_c128* sl, const nt_c128* sr,
               const nt_c128* const* Hd, const nt_c128* const* Hdv, const nt_c128* const* Hdv_back,
               const nt_c128* Hld, const nt_c128* Hrd, const nt_c128* Hrd_fwd,
               const nt_c128* Haml, const nt_c128* Hamr,
               double* transl, double* transr,
               nt_c128* G_D_RL, nt_c128* G_D_LR, nt_c128* G_L, nt_c128* G_R);

/* get_ldos (tb/device.f90:610-855): layer-resolved LDOS of the device.
 * Arguments as nt_G_block (gl = retarded left surface GF, gr = retarded
 * right surface GF, layer blocks with optional dense lower blocks and
 * optional V_{N,N+1}); the device broadening is factor*eta(E) (TB: eta).
 * Outputs ldos (column-major L x maxlen, maxlen >= max layer size, zero
 * beyond each layer's size): -Im GLR_i(j,j)/pi with
 * GLR_i = (I - G^L_i V_{i,i+1} G^R_{i+1} V_{i,i+1}^+)^{-1} G^L_i, and
 * ddos = sum of ldos (may be NULL).  The reference writes ldos with
 * write_ldos and returns ddos. */
int nt_get_ldos(nt_ctx* ctx, const nt_params* p, int nl, int nr, int L, const int* layer_list,
                double E, const nt_c128* gl, const nt_c128* gr,
                const nt_c128* const* Hd, const nt_c128* const* Hdv, const nt_c128* const* Hdv_back,
                const nt_c128* Hld, const nt_c128* Hrd, const nt_c128* Hrd_fwd,
                int maxlen, double* ldos, double* ddos);

/* ----------------------------------------------- batched (omega,k) solve
 * The hot path: for npts points (E[i], k-point kidx[i]) it runs the lead
 * decimations, the adv->ret conversion, both RGF sweeps and Caroli T, all on
 * the GPU.  This replaces the body of the driver's ie loop
 * (negf/phinfortran.f90:1196-1251) for every energy of one or more k-points.
 *
 * Block arrays hold nk k-points back to back:
 *   Haml, Hamlv  nk x (nl x nl)          Hamr, Hamrv  nk x (nr x nr)
 *   Hamd         nk x L x (b x b)        Hamdv        nk x (L-1) x (b x b)
 *   Hamld        nk x (nl x b)           Hamrd        nk x (nr x b)
 * kidx (npts ints, may be NULL when nk == 1) selects the k-point of each
 * point.  first_point_mask (may be NULL): points with a nonzero entry get
 * T = 0 (the reference's ie = 1 quirk).
 * Outputs: T_LR[npts], T_RL[npts]; status[npts] and iters[npts] (mean of the
 * two decimations, may be NULL).  Returns the first nonzero point status. */
typedef struct {
    /* optional per-point outputs (npts x n x n each); NULL to skip */
    nt_c128* gl_adv;  nt_c128* gr_ret;  nt_c128* gl_plus; nt_c128* gr_minus;
    nt_c128* gml_adv; nt_c128* gmr_ret; nt_c128* sl_adv;  nt_c128* sr_ret;
    nt_c128* G_D_LR;  nt_c128* G_D_RL;  nt_c128* G_L;     nt_c128* G_R;
    double* sdos_l;   double* sdos_r;   /* surface_ldos_mpi: +-2 sqrt(E) Im Tr g / pi (negf/surface.f90:608-630);
                                           with eta_mode = NT_ETA_ABSOLUTE (tb) +-Im Tr g / pi (tb/surface.f90:345-366) */
    /* get_ldos (tb/device.f90:610-855): ldos npts x L x b (point, layer,
     * orbital; -Im G_ii(j,j) / pi of the full device G at the device shift),
     * ddos npts (its sum).  Either one makes the solve keep every layer's
     * left-connected G and run both sweeps in FP64 (also with
     * NT_PREC_MIXED); not with the probe entry. */
    double* ldos;     double* ddos;
} nt_solve_extras;

/* Device-pointer variant: every array is device memory (e.g. torch tensors). */
int nt_solve_tdb_batched_dev(nt_ctx* ctx, const nt_params* p, int nl, int nr, int b, int L,
                             int npts, const double* E, const int* kidx, int nk,
                             const nt_c128* Haml, const nt_c128* Hamlv,
                             const nt_c128* Hamr, const nt_c128* Hamrv,
                             const nt_c128* Hamd, const nt_c128* Hamdv,
                             const nt_c128* Hamld, const nt_c128* Hamrd,
                             const int* first_point_mask,
                             double* T_LR, double* T_RL, int* status, int* iters,
                             const nt_solve_extras* extras);

/* nt_solve_tdb_batched_dev without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results; the workspace is a static device
 * array of 256 MiB per GPU (allocated when the library is loaded, held until
 * the process ends) instead of the context's growing workspace, and the
 * points are solved in chunks as large as fits in it (b = 192, L = 40, FP64:
 * about 12 points per chunk).  NT_ERR_ARG when a single point does not fit
 * (nt_last_error gives the size) and for circular decimation (p->circular),
 * whose transforms are allocated.  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams or threads) race. */
int nt_solve_tdb_batched_noalloc_dev(nt_ctx* ctx, const nt_params* p, int nl, int nr, int b, int L,
                                     int npts, const double* E, const int* kidx, int nk,
                                     const nt_c128* Haml, const nt_c128* Hamlv,
                                     const nt_c128* Hamr, const nt_c128* Hamrv,
                                     const nt_c128* Hamd, const nt_c128* Hamdv,
                                     const nt_c128* Hamld, const nt_c128* Hamrd,
                                     const int* first_point_mask,
                                     double* T_LR, double* T_RL, int* status, int* iters,
                                     const nt_solve_extras* extras);

/* Several devices of uneven block size b, lead sizes and layer count L in
 * one call (every pointer of a descriptor is device memory, laid out as for
 * nt_solve_tdb_batched_dev).  Devices whose lead buffers, lead sizes and
 * points (E, k map) are the same form a group whose lead Green's functions
 * are computed once (FP64).  Groups are taken largest first (cost
 * b^3 (L + 2 N_dec) per point) by up to nt_set_multi_workers worker
 * contexts, each with its own stream, workspace and host thread, so that
 * devices whose launches do not fill the GPU run side by side; the work
 * starts after the context's stream has reached the call and the context's
 * stream waits for all of it.  The inverse panel width is autotuned once per
 * distinct block size.  Each descriptor's status array (may be NULL)
 * receives its points' statuses; returns NT_OK or the first argument/HIP
 * error.  No reference counterpart: the reference solves one device per run. */
typedef struct {
    int nl, nr, b, L, nk, npts;
    const nt_c128 *Haml, *Hamlv, *Hamr, *Hamrv, *Hamd, *Hamdv, *Hamld, *Hamrd;
    const double* E;
    const int* kidx;
    const int* first_point_mask;
    double* T_LR;
    double* T_RL;
    int* status;
} nt_device_desc;

int nt_solve_tdb_multi_dev(nt_ctx* ctx, const nt_params* p, int ndev, const nt_device_desc* devs);
/* Worker contexts of nt_solve_tdb_multi_dev, 1..16 (default 4; 1: every
 * group on the context itself, one after the other).  No reference
 * counterpart. */
int nt_set_multi_workers(nt_ctx* ctx, int n);
/* Batched solves (nt_solve_tdb_batched[_dev], FP64 or mixed) of at least
 * n * 256 points run as n parts solved side by side by worker contexts of
 * ctx (own stream, workspace and host thread each, ordered after and before
 * the caller's stream): the latency-bound launches of one part (the
 * inverses, the late decimation steps of few unconverged items) leave
 * compute units to the other's GEMMs.  Part q takes every n-th point from
 * q (inputs, T, statuses and every extra output gathered in the context's
 * workspace and scattered back, so the parts need similar numbers of
 * decimation steps).  Results are bitwise those of one part.  The adjoint solve (nt_solve_tdb_adjoint_batched_dev) splits the
 * same way, interleaved; part 0 accumulates into the caller's gradients,
 * the others into zeroed workspace summed into them at the end (the
 * gradients then differ from one part's in the rounding of that order).
 * n = 2 (default) .. 8; n = 1: every point on the context's own stream.
 * The allocation-free entry and probes always run on one stream.  No
 * counterpart in the reference. */
int nt_set_point_split(nt_ctx* ctx, int n);

/* Host-pointer variant (copies in and out; same layout). */
int nt_solve_tdb_batched(nt_ctx* ctx, const nt_params* p, int nl, int nr, int b, int L,
                         int npts, const double* E, const int* kidx, int nk,
                         const nt_c128* Haml, const nt_c128* Hamlv,
                         const nt_c128* Hamr, const nt_c128* Hamrv,
                         const nt_c128* Hamd, const nt_c128* Hamdv,
                         const nt_c128* Hamld, const nt_c128* Hamrd,
                         const int* first_point_mask,
                         double* T_LR, double* T_RL, int* status, int* iters,
                         const nt_solve_extras* extras);

/* ------------------------------------------------ adjoint (gradient) solve
 * Batched solve that also returns the gradient of W = sum_i weight[i] T_LR[i]
 * with respect to the device blocks of layers lo..hi (0-based, 0 <= lo <= hi
 * < L).  No counterpart in the reference, which has no derivatives; central
 * finite differences of its transl (negf/lesser.f90:787-800) pin it.
 *
 * With G the Green's function of the left-to-right sweep and
 * A = Gamma_R G_{0,N+1}^+ Gamma_L, dT_LR = 2 Re Tr[dH X] where
 * X_{j,i} = G_{j,N+1} A G_{0,i}.  The outputs are ACCUMULATED (+=) per k-point,
 * in the convention of torch's complex autograd (dW/dRe + i dW/dIm):
 *   dW_dHamd   nk x (hi-lo+1) x (b x b):  2 sum_i w_i X_{l,l}^+
 *   dW_dHamdv  nk x nc x (b x b), nc = min(hi, L-2)-lo+1 (may be NULL if nc<=0):
 *              2 sum_i w_i (X_{l+1,l}^+ + X_{l,l+1})  for coupling l -> l+1,
 *              whose dagger is the lower block of the tdb layout.
 * Points with a nonzero status or a nonzero first_point_mask add nothing.
 * The sweep runs in FP64 whatever p->precision says.  Other arguments as
 * nt_solve_tdb_batched_dev (device pointers). */
int nt_solve_tdb_adjoint_batched_dev(nt_ctx* ctx, const nt_params* p, int nl, int nr, int b, int L,
                                     int npts, const double* E, const int* kidx, int nk,
                                     const nt_c128* Haml, const nt_c128* Hamlv,
                                     const nt_c128* Hamr, const nt_c128* Hamrv,
                                     const nt_c128* Hamd, const nt_c128* Hamdv,
                                     const nt_c128* Hamld, const nt_c128* Hamrd,
                                     const int* first_point_mask, const double* weight,
                                     int lo, int hi,
                                     double* T_LR, double* T_RL, int* status, int* iters,
                                     nt_c128* dW_dHamd, nt_c128* dW_dHamdv);

/* ------------------------------------------------ Buttiker probes
 * One probe per device layer with self-energy sigma_w = -2 i sqrt(E) / tau,
 * 1/tau = B E T_i exp(-C / T_i) (set_sigma_bp_w, negf/lesser.f90:12-45;
 * zeroed when lifetime == 0, its T derivative is not), in every device layer
 * of both sweeps (get_G_full_tdb, negf/lesser.f90:475-809).  Replaces, for
 * npts points, the pair get_G_full_tdb_mpi + get_G_lesser_tdb_mpi of the
 * driver's Buttiker loop (negf/phinfortran.f90:1213-1237): T_LR, T_RL and
 * status as nt_solve_tdb_batched_dev, and per point p (row-major by point):
 *   flux[p L + i]            flux_i(i)       probe current of layer i
 *   jacobian[p L^2 + i + j L] jacobian(nn,ie,i,j) = d flux_i / d T_j at fixed
 *                            retarded functions; the diagonal differentiates
 *                            gamma_i only in its A_i term, as the reference
 *   flux_left[p], flux_right[p]  contact currents (times frequency sqrt(E))
 *   a_sum[p L + i], gn_sum[p L + i]  Tr A_i, Tr Gn_i
 * (negf/lesser.f90:1199-1503).  The reference keeps these as complex numbers
 * whose imaginary parts vanish in exact arithmetic; they are returned real.
 * Points with a nonzero first_point_mask get zero probe currents, Jacobian
 * and traces, but their contact currents (negf/lesser.f90:1467-1472).
 * extras (may be NULL): the lead and RGF outputs of nt_solve_tdb_batched_dev,
 * here with the probes in the sweeps -- what vel_mat_tdb receives in the
 * Buttiker loop (negf/phinfortran.f90:1239-1250), so nt_modes_tdb_batched_dev
 * on them gives the mode-resolved transmission with probes.
 * Runs in FP64 whatever p->precision says; needs L >= 2 and nl == nr.
 * temperature: L probe temperatures in device memory. */
typedef struct {
    const double* temperature;   /* temperature_i (L) */
    double t_left, t_right;      /* contact temperatures t_left, t_right */
    double B_probe, T_probe;     /* B and C of the lifetime (config B_probe, T_probe) */
    int lifetime;                /* config lifetime: 0 zeroes sigma_w */
} nt_probe_desc;

typedef struct {
    double* flux;        /* npts x L */
    double* jacobian;    /* npts x L x L (column-major per point) */
    double* flux_left;   /* npts */
    double* flux_right;  /* npts */
    double* a_sum;       /* npts x L */
    double* gn_sum;      /* npts x L */
} nt_probe_out;

int nt_buttiker_tdb_batched_dev(nt_ctx* ctx, const nt_params* p, const nt_probe_desc* probe,
                                int nl, int nr, int b, int L, int npts, const double* E,
                                const int* kidx, int nk,
                                const nt_c128* Haml, const nt_c128* Hamlv,
                                const nt_c128* Hamr, const nt_c128* Hamrv,
                                const nt_c128* Hamd, const nt_c128* Hamdv,
                                const nt_c128* Hamld, const nt_c128* Hamrd,
                                const int* first_point_mask,
                                double* T_LR, double* T_RL, int* status, int* iters,
                                const nt_solve_extras* extras, const nt_probe_out* out);

/* ------------------------------------------------ mode-resolved transmission
 * vel_mat_tdb (negf/velocity_mat.f90:1692-3150; tb/velocity_mat.f90:222-1560)
 * for npts points (the Brillouin-zone unfolding that splits specular and
 * non-specular parts: nt_modes_unfold_tdb_batched_dev below).  Inputs are the outputs of the batched solve for
 * the same points (nt_solve_extras: gl_adv, gr_ret, gl_plus, gr_minus,
 * gml_adv, gmr_ret, G_D_RL, G_D_LR, G_L, G_R; npts x n x n each), the lead
 * blocks per k-point (kidx as in the batched solve) and the lead periods
 * a_z_L, a_z_R.  tb = 0: negf conventions (velocity i a/(2 sqrt E), t prefactor
 * 2 i sqrt(E/(a a'))); tb = 1: tight-binding conventions (no factors).
 * Every output may be NULL; per point:
 *   trl nr x nl (t_RL, left -> right), tlr nl x nr, rrr nr x nr, rll nl x nl;
 *   k_z of the four mode sets (L_adv, L_ret: nl; R_ret, R_adv: nr), 100 for
 *   evanescent modes; n_prop (4 ints: L_adv, L_ret, R_ret, R_adv);
 *   t_mode_L (nl): for each propagating incident left mode, its transmission
 *   summed over propagating outgoing right modes (the r_t column of the
 *   reference's trl tables); t_mode_R (nr) the same for t_LR;
 *   sums (4 doubles): T_RL, T_LR, R_RR, R_LL summed over propagating modes.
 * Points flagged in first_point_mask get zero t and r (velocity_mat.f90:2965). */
typedef struct {
    nt_c128* trl; nt_c128* tlr; nt_c128* rrr; nt_c128* rll;
    double* k_z_L_adv; double* k_z_L_ret; double* k_z_R_ret; double* k_z_R_adv;
    int* n_prop;
    double* t_mode_L; double* t_mode_R;
    double* sums;
} nt_modes_out;

int nt_modes_tdb_batched_dev(nt_ctx* ctx, const nt_params* p, int tb, int nl, int nr, int npts,
                             const double* E, const int* kidx, int nk,
                             const nt_c128* Haml, const nt_c128* Hamlv,
                             const nt_c128* Hamr, const nt_c128* Hamrv,
                             const nt_c128* gl_adv, const nt_c128* gr_ret,
                             const nt_c128* gl_plus, const nt_c128* gr_minus,
                             const nt_c128* gml_adv, const nt_c128* gmr_ret,
                             const nt_c128* G_D_RL, const nt_c128* G_D_LR,
                             const nt_c128* G_L, const nt_c128* G_R,
                             double a_z_L, double a_z_R, const int* first_point_mask,
                             const nt_modes_out* out, int* status);

/* ------------------------------------------------ Brillouin-zone unfolding
 * The specular / non-specular labelling of vel_mat_tdb: every propagating
 * mode gets the in-plane k-shift j (1 .. n_bloch_x n_bloch_y, numbered as
 * k_shift in negf/config.f90:176-183) of largest weight
 *   W(i,j) = sum_o |sum_ir exp(-i(-kc.r + kp_j.(r - r_uc(o)))) U(o + ir nl/nb, i) lambda_i|^2 / nb
 * (negf/velocity_mat.f90:2035-2076, mirrored at :2321-2355, :2578-2614,
 * :2826-2862), and the unfolded k = move_in_ws2((ksc + shift_j) . recivec)
 * (negf/util.f90:675-706) with k_z as third component.  The lead geometry is
 * what transverse_uc_setup (velocity_mat.f90:14-179) derives from the atom
 * positions of the second lead period:
 *   pos_l      nl x 3 (row-major): position of every left-lead orbital (pos_l(:,1:3))
 *   pos_uc_l   (nl / nb) x 3: for orbital o of the first transverse image, the
 *              position of the transverse-cell atom it maps to
 *              (positions(L_list(pidxL(o)), :))
 *   pos_r, pos_uc_r the same for the right lead;
 *   recivec    the device supercell's reciprocal vectors (rows, row-major);
 *   reci_uc_l/r the lead unit cells' (rows, row-major);
 *   ksc        nk x 3: each k-point in crystal coordinates of recivec (k_grid).
 * n_bloch_x * n_bloch_y must divide nl and nr and be at most NT_MAX_BLOCH.
 * All pointers are device memory. */
#define NT_MAX_BLOCH 256
typedef struct {
    int n_bloch_x, n_bloch_y;
    double recivec[9];
    double reci_uc_l[9];
    double reci_uc_r[9];
    const double* ksc;
    const double* pos_l;
    const double* pos_uc_l;
    const double* pos_r;
    const double* pos_uc_r;
} nt_unfold_desc;

/* Outputs (device memory, each may be NULL), per point:
 *   bz_*       n ints: the label (0 for evanescent modes)            (bz_idx_*)
 *   k_xyz_*    n x 3 column-major: unfolded k, 100 for evanescent modes (k_xyz_*)
 *   split      8 doubles: transl_s, transl_ns, transr_s, transr_ns, refll_s,
 *              refll_ns, reflr_s, reflr_ns (velocity_mat.f90:3032-3099)
 *   spec_trl   nl x 2 column-major: for every incident left mode (L_adv) the
 *              sum of |t_RL(j,i)|^2 over propagating outgoing right modes and
 *              its part with equal in-plane k (|dk|^2 < 1e-5): the r_t and r_s
 *              columns of write_specular (negf/write_data.f90:3-76) for 'trl';
 *   spec_tlr   nr x 2: the same for t_LR (incident R_adv, outgoing L_ret),
 *   spec_rrr   nr x 2: r_RR (incident R_adv, outgoing R_ret),
 *   spec_rll   nl x 2: r_LL (incident L_adv, outgoing L_ret); write_md's
 *              columns are spec_rll + spec_trl ('mdl') and spec_rrr + spec_tlr
 *              ('mdr').
 * The primitive-cell columns of those tables (find_k_pc, negf/dyn.f90:770-889)
 * come from nt_unfold_pc_dev on k_xyz_*. */
typedef struct {
    int* bz_L_adv; int* bz_L_ret; int* bz_R_ret; int* bz_R_adv;
    double* k_xyz_L_adv; double* k_xyz_L_ret; double* k_xyz_R_ret; double* k_xyz_R_adv;
    double* split;
    double* spec_trl; double* spec_tlr; double* spec_rrr; double* spec_rll;
} nt_unfold_out;

/* nt_modes_tdb_batched_dev plus the unfolding of every point (uf required). */
int nt_modes_unfold_tdb_batched_dev(nt_ctx* ctx, const nt_params* p, int tb, int nl, int nr, int npts,
                                    const double* E, const int* kidx, int nk,
                                    const nt_c128* Haml, const nt_c128* Hamlv,
                                    const nt_c128* Hamr, const nt_c128* Hamrv,
                                    const nt_c128* gl_adv, const nt_c128* gr_ret,
                                    const nt_c128* gl_plus, const nt_c128* gr_minus,
                                    const nt_c128* gml_adv, const nt_c128* gmr_ret,
                                    const nt_c128* G_D_RL, const nt_c128* G_D_LR,
                                    const nt_c128* G_L, const nt_c128* G_R,
                                    double a_z_L, double a_z_R, const int* first_point_mask,
                                    const nt_modes_out* out, const nt_unfold_desc* uf,
                                    const nt_unfold_out* uo, int* status);

/* find_k_pc (negf/dyn.f90:770-889) for nmodes modes, as vel_mat_tdb calls it
 * for every propagating mode (negf/velocity_mat.f90:2073-2089 and mirrors):
 *   E        nmodes: the point's energy (omega^2) of each mode
 *   k_xyz    nmodes x 3 (row-major): the mode's unfolded k (nt_unfold_out
 *            k_xyz_*, k_z third); k_z >= 99.9 marks an evanescent mode
 *   n_bloch_max  max(n_bloch_x, n_bloch_y): candidates k + n reci_uc, n in [-l, l]^3
 *   reci_uc, recivec_pc  host, 9 doubles each, rows = reciprocal vectors of the
 *            lead unit cell and of the primitive cell (read_pc_qe,
 *            negf/input.f90:688-720)
 *   n_pc     3 x atoms of the primitive cell (<= 48)
 *   nseg ... dr  the primitive cell's dynamical matrix as a phase-sum table in
 *            the format of nt_hk_assemble_dev (gen_dyn_pc_qe_k,
 *            negf/dyn.f90:2537-2644: one segment per 3x3 block (i, j mod
 *            natoms_pc)); the matrix is used as (D + D^H)/2
 * Outputs (device): kp (nmodes x 3): the matching candidate of smallest |k|
 * (k_xyz_*_pc), 0 when none matches, 100 for evanescent modes; v (nmodes x 3):
 * the first matching candidate's group velocity (vel_xyz_*_pc); ncand: the
 * number of matching candidates.  All other pointers are device memory. */
int nt_unfold_pc_dev(nt_ctx* ctx, int nmodes, const double* E, const double* k_xyz, int n_bloch_max,
                     const double* reci_uc, const double* recivec_pc, int n_pc, int nseg, const int* seg_off,
                     const int* seg_rc, int nent, const double* coef, const double* dr, double* kp, double* v,
                     int* ncand);

/* find_k_pc_l / find_k_pc_r of the tight-binding run (tb/hamiltonian.f90:
 * 666-916), as its vel_mat calls them for every propagating mode
 * (tb/velocity_mat.f90:593-605 and mirrors): the candidates of
 * nt_unfold_pc_dev; per candidate k the primitive cell's H(k) = hermitian(
 * sum V exp(i k.r)) (gen_ham_pc_l, :918-989), its eigenpairs and the velocities
 * Re(z^H HF_d z), HF_d = i antihermitian(sum V r_d exp(i k.r)) (gen_hf_pc_l,
 * :1238-1313; group_velocity_pc_l, :1316-1378); a candidate matches when a
 * band lies within 1e-4 of E (compare_evs_l, :1596-1686; within 1e-3 for the
 * mode's own k on the unit cell's zone boundary).
 *   E        nmodes: the point's energy (eV) of each mode
 *   n_pc     orbitals of the primitive cell (<= 48)
 *   el_off   n_pc^2 + 1 ints: the table entries of matrix element
 *            (r, c) are el_off[r + c n_pc] .. el_off[r + c n_pc + 1] - 1
 *   coef, dr nent values V and nent x 3 bond vectors r (Angstrom), one per
 *            (atom pair, cell image within r_cutoff, orbital pair)
 * Other arguments and outputs as nt_unfold_pc_dev.  All pointers except
 * reci_uc and recivec_pc are device memory. */
int nt_unfold_pc_tb_dev(nt_ctx* ctx, int nmodes, const double* E, const double* k_xyz, int n_bloch_max,
                        const double* reci_uc, const double* recivec_pc, int n_pc, const int* el_off, int nent,
                        const double* coef, const double* dr, double* kp, double* v, int* ncand);

/* ------------------------------------------------ H(k) assembly
 * gen_dyn_qe_tdb_k (negf/dyn.f90:1482-2080) on the GPU: one block for nk
 * k-points.  The host tabulates the k-independent part as nseg segments (one
 * per 3x3 atom-pair sub-block at atom row seg_rc[2s], atom column
 * seg_rc[2s+1]) of entries seg_off[s] .. seg_off[s+1]-1, each with a 3x3
 * real coefficient (row-major, coef[9 q + 3 alpha + beta]) = w Phi/sqrt(m m')
 * and its image vector dr[3 q] (Angstrom).  For every k (kcart: nk x 3,
 * 1/Angstrom) out[k] = sum coef exp(i k.dr), a rows x cols complex block with
 * leading dimension ld at out + k * kstride (complex elements); entries of no
 * segment are zero.  All pointers are device memory. */
int nt_hk_assemble_dev(nt_ctx* ctx, int nk, const double* kcart, int nseg, const int* seg_off, const int* seg_rc,
                       int nent, const double* coef, const double* dr, int rows, int cols, nt_c128* out, int ld,
                       long long kstride);

/* gen_ham (tb/hamiltonian.f90:444-519) on the GPU: the same phase sum with
 * sub-blocks of any size (norb_i x norb_j orbitals per atom pair) for nk
 * k-points.  Segment s is the sub-block at orbital row seg_geom[4s], column
 * seg_geom[4s+1] of seg_geom[4s+2] x seg_geom[4s+3] elements; its entries
 * seg_off[s] .. seg_off[s+1]-1 (one per cell image within r_cutoff, in the
 * reference's image order) hold the real coefficients row-major at
 * coef + q * cstride (onsite or V_sk, tb/hamiltonian.f90:57-417) and the
 * bond vector dr[3 q] (Angstrom).  For every k (kcart: nk x 3, 1/Angstrom,
 * = kpoint . recivec) out[k] = sum coef exp(i k.dr); hermitize = 1 then
 * replaces it by (H + H^+)/2 (hermitian, tb/util.f90:709-718; rows = cols).
 * The sub-blocks must tile the block on a grid (as atom-pair blocks do):
 * sub-blocks starting on one row have the same row count, row ranges of
 * different starts do not overlap, likewise for columns, and no two
 * sub-blocks share both ranges; NT_ERR_ARG otherwise.  Every element of
 * out is written once.  All pointers are device memory. */
int nt_hk_assemble_orb_dev(nt_ctx* ctx, int nk, const double* kcart, int nseg, const int* seg_off,
                           const int* seg_geom, int nent, int cstride, const double* coef, const double* dr, int rows,
                           int cols, nt_c128* out, int ld, long long kstride, int hermitize);

/* ------------------------------------------------ autotuning
 * The batched inverse runs blocked Gauss-Jordan in one workgroup per matrix
 * with a panel width of 8, 16 or 32 columns, or, as "width" 1 (n >= 64), with
 * one panel kernel per panel and the trailing updates done by the batched GEMM.
 * Default: width 16 where it fits, width 1 from n = 320.  nt_autotune_inverse
 * times every supported width on `batch` n x n matrices and keeps the fastest
 * for this context (precision NT_PREC_FP64: complex double, NT_PREC_MIXED:
 * complex float); best_nb / best_ms may be NULL.  nt_set_inverse_panel sets a
 * width directly (0 restores the default).  No reference counterpart. */
int nt_autotune_inverse(nt_ctx* ctx, int n, int batch, int precision, int* best_nb, double* best_ms);
int nt_set_inverse_panel(nt_ctx* ctx, int n, int precision, int nb);
/* Output tile of the batched complex GEMM: 64x64 per 256-thread workgroup
 * (default) or 32x32 per one-wave workgroup, which pads less when the matrix
 * size is not a multiple of 64 (24, 80, 96, 216, ...).  nt_autotune_gemm
 * times the n x n x n product of `batch` items with both and keeps the faster
 * for every GEMM of this context whose larger output dimension is n
 * (precision NT_PREC_FP64: complex double, NT_PREC_MIXED: complex float);
 * best_tile / best_ms may be NULL.  nt_set_gemm_tile sets it directly (0
 * restores the default).  No reference counterpart. */
int nt_autotune_gemm(nt_ctx* ctx, int n, int batch, int precision, int* best_tile, double* best_ms);
int nt_set_gemm_tile(nt_ctx* ctx, int n, int precision, int tile);
/* on = 1 (default): the right-to-left RGF sweep runs on a second stream of
 * the context beside the left-to-right one; on = 0: both on the main stream. */
int nt_set_overlap(nt_ctx* ctx, int on);
/* Batched solves: n = 1 (default) computes T_RL, G_D_RL and G_L from the
 * left-to-right sweep and a correction of its corner blocks for the two end
 * blocks where the right-to-left system differs (the broadening of the lead
 * surface layers), two inverses of nl and nr and six GEMMs per point instead
 * of the second sweep; n = 2 runs the reference's second sweep
 * (negf/lesser.f90:694-793).
 * The per-point entries (nt_get_G_full_tdb, nt_G_block), the Buttiker-probe
 * solve and leads given with Hrd_fwd / dense lower couplings always run both
 * sweeps. */
int nt_set_sweeps(nt_ctx* ctx, int n);
/* Surface decimation (negf/surface.f90:484-497, the convergence test of its
 * loop): the host reads the number of unconverged (omega,k) items every
 * `every` iterations (default 4) and stops when it is zero; items that have
 * converged are masked inside every kernel, so the results do not depend on
 * `every`.  every = 0: the host never reads it, each decimation runs max_iter
 * iterations, and the FP64 batched solve (nt_solve_tdb_batched_dev and
 * _noalloc_dev, direct decimation, profiling off) makes no host
 * synchronisation, so a caller can capture it in a HIP graph and replay it
 * (with the allocation-free entry, or after a first call that sized the
 * workspace); set max_iter to the bound the points need (items still
 * unconverged then get status -2, as with the same max_iter eagerly).
 * No reference counterpart. */
int nt_set_decimation_check(nt_ctx* ctx, int every);
/* Storage of nt_solve_tdb_adjoint_batched_dev's backward pass: the left-
 * connected blocks G^L_i, G^L_{0,i} of the layers from the window to the end
 * are kept for every layer (every = 0, default), for every n-th layer
 * (every = n > 0) or for about every sqrt(L - lo)-th layer (every < 0); the
 * others are recomputed from the nearest kept layer during the backward pass
 * by the same launches as the forward sweep (same values, one more
 * left-to-right pass over the recomputed layers).  No reference counterpart
 * (the reference has no derivatives). */
int nt_set_adjoint_checkpoint(nt_ctx* ctx, int every);
/* Device bytes per (omega,k) point that the adjoint solve's stored blocks take
 * with the current checkpoint setting (the chunk of points is capped so that
 * they stay within 24 GiB). */
int nt_adjoint_bytes_per_point(nt_ctx* ctx, int nl, int nr, int b, int L, int lo, int hi, long long* bytes);
/* on = 1: FP64 complex GEMMs form each complex product from three real MFMA
 * products ((Ar+Ai)(Br+Bi) - ArBr - AiBi for the imaginary part) instead of
 * four; on = 0 (default): four. */
int nt_set_gemm_3m(nt_ctx* ctx, int on);
/* Debugging aid, no counterpart in the reference: on = 1 fills the
 * workspace (and, for nt_solve_tdb_batched_noalloc_dev, the static scratch)
 * with all-ones words, a NaN in every double, at the start of each call, so
 * that a read of workspace the call has not written shows up as a NaN
 * status or output; on = 0 (default): no fill. */
int nt_set_workspace_poison(nt_ctx* ctx, int on);

/* ------------------------------------------------ batched building blocks
 * The two kernels every solve is made of, on device memory (column-major,
 * leading dimension = rows, item p at base + p*stride elements; stride 0
 * shares one matrix across the batch).  No counterpart in the reference,
 * whose equivalents are Fortran matmul and inv (negf/util.f90:16-67). */

/* C[p] = alpha op(A[p]) op(B[p]) + beta C[p], op = 'N' or 'C' (conjugate transpose). */
int nt_zgemm_batched_dev(nt_ctx* ctx, char opA, char opB, int m, int n, int k, nt_c128 alpha,
                         const nt_c128* A, int lda, long long strideA,
                         const nt_c128* B, int ldb, long long strideB, nt_c128 beta,
                         nt_c128* C, int ldc, long long strideC, int batch);

/* X[p] = inv(A[p]) by Gauss-Jordan with partial pivoting; A is overwritten.
 * info[p] (may be NULL): 0, k > 0 for an exact zero pivot, -1 for NaN. */
int nt_zinv_batched_dev(nt_ctx* ctx, int n, nt_c128* A, long long strideA, nt_c128* X, long long strideX,
                        int* info, int batch);

/* Eigen-decomposition of general complex matrices, as eigen (negf/util.f90:
 * 226-262, LAPACK ZGEEV with right eigenvectors): A[p] (n x n, preserved)
 * -> eigenvalues w[p] (n) and eigenvectors V[p] (columns, unit 2-norm with
 * the component of largest modulus real and positive).  Hessenberg reduction
 * + single-shift complex QR to the Schur form, then back substitution.
 * status[p] (may be NULL): 0, or -3 when the QR iteration did not converge. */
int nt_zgeev_batched_dev(nt_ctx* ctx, int n, const nt_c128* A, long long strideA, nt_c128* w,
                         nt_c128* V, long long strideV, int* status, int batch);
/* Principal square root of general complex matrices as sqrtm (negf/util.f90:
 * 359-454): complex Schur form A = Z T Z^H, the upper triangular square root
 * of T (zero where the two diagonal roots sum to exactly zero), S = Z U Z^H. */
int nt_zsqrtm_batched_dev(nt_ctx* ctx, int n, const nt_c128* A, long long strideA, nt_c128* S,
                          long long strideS, int* status, int batch);

/* ------------------------------------------------------- measurement */
/* Kernel classes for nt_profile_get. */
enum { NT_K_GEMM = 0, NT_K_INV = 1, NT_K_ELEM = 2, NT_K_EIG = 3, NT_K_COUNT = 4 };
/* While enabled, each launch is bracketed by HIP events on the context's
 * stream; nt_profile_get returns launches, summed milliseconds and the summed
 * algorithmic flops (GEMM 8mnk, inverse 8n^3 per active item) of a class. */
int nt_profile_enable(nt_ctx* ctx, int on);
int nt_profile_get(nt_ctx* ctx, int kind, long long* launches, double* total_ms, double* flops);
int nt_profile_reset(nt_ctx* ctx);
/* Mean decimation iterations of the last batched solve. */
double nt_last_mean_iters(nt_ctx* ctx);
/* Mixed-precision bookkeeping since the last nt_profile_reset: counts[0]
 * points solved with NT_PREC_MIXED, counts[1] second refinement steps (both
 * sweep directions), counts[2] points whose refined result failed the error
 * estimate and were recomputed by the FP64 sweep.  No reference counterpart
 * (the reference computes in FP64 only). */
int nt_mixed_counts(nt_ctx* ctx, long long* counts);

#ifdef __cplusplus
}
#endif
#endif
