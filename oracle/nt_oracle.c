/*
 * nt_oracle.c -- CPU FP64 restatement of phinfortran's per-(omega,k) transport
 * solve.  TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); the
 * product path in phinfortran_amd/ never links or calls this file.
 *
 * Each routine names the reference lines it follows.  Arithmetic is plain IEEE
 * FP64 complex (no fast-math); complex products are written out so that the
 * compiler does not take the C99 NaN-recovery slow path.
 */
#include "nt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RE(z) creal(z)
#define IM(z) cimag(z)
#define MKC(a, b) ((double)(a) + (double)(b) * I)

static inline ocplx cmul(ocplx a, ocplx b)
{
    double ar = RE(a), ai = IM(a), br = RE(b), bi = IM(b);
    return MKC(ar * br - ai * bi, ar * bi + ai * br);
}

/* Smith's division, as Fortran runtimes do for complex division. */
static inline ocplx cdiv(ocplx a, ocplx b)
{
    double ar = RE(a), ai = IM(a), br = RE(b), bi = IM(b);
    if (fabs(br) >= fabs(bi)) {
        double r = bi / br, d = br + bi * r;
        return MKC((ar + ai * r) / d, (ai - ar * r) / d);
    } else {
        double r = br / bi, d = bi + br * r;
        return MKC((ar * r + ai) / d, (ai * r - ar) / d);
    }
}

static inline double cabs1(ocplx z) { return fabs(RE(z)) + fabs(IM(z)); }

double or_lead_eta(const or_params* p, double E)
{
    /* negf/surface.f90:479 (E*eta+eta0) ; tb/surface.f90:218 (eta) */
    return p->eta_mode == OR_ETA_RELATIVE ? E * p->eta + p->eta0 : p->eta;
}

static ocplx* cmalloc(size_t n) { return (ocplx*)malloc(sizeof(ocplx) * (n ? n : 1)); }

/* ------------------------------------------------------------------ GEMM */
void or_gemm(char opa, char opb, int m, int n, int k, ocplx alpha,
             const ocplx* A, int lda, const ocplx* B, int ldb,
             ocplx beta, ocplx* C, int ldc)
{
    /* Materialise op(A) as m x k column-major and op(B) as k x n so the kernel
     * is one cache-friendly j-p-i loop. */
    ocplx* a = cmalloc((size_t)m * k);
    ocplx* bb = cmalloc((size_t)k * n);
    for (int p = 0; p < k; ++p)
        for (int i = 0; i < m; ++i)
            a[i + (size_t)p * m] = (opa == 'N') ? A[i + (size_t)p * lda] : conj(A[p + (size_t)i * lda]);
    for (int j = 0; j < n; ++j)
        for (int p = 0; p < k; ++p)
            bb[p + (size_t)j * k] = (opb == 'N') ? B[p + (size_t)j * ldb] : conj(B[j + (size_t)p * ldb]);
    /* Columns of C are independent: large products (the b = 768 cases) run
     * them over the OpenMP team; inside the per-point parallel loop of
     * or_solve_tdb_batch the region is inactive and runs on one thread. */
    const int par = (double)m * n * k >= 2.0e7 && !omp_in_parallel();
#pragma omp parallel if (par)
    {
        double* accr = (double*)malloc(sizeof(double) * (m ? m : 1));
        double* acci = (double*)malloc(sizeof(double) * (m ? m : 1));
#pragma omp for schedule(static)
        for (int j = 0; j < n; ++j) {
            for (int i = 0; i < m; ++i) { accr[i] = 0.0; acci[i] = 0.0; }
            for (int p = 0; p < k; ++p) {
                double br = RE(bb[p + (size_t)j * k]), bi = IM(bb[p + (size_t)j * k]);
                const ocplx* ac = a + (size_t)p * m;
                for (int i = 0; i < m; ++i) {
                    double ar = RE(ac[i]), ai = IM(ac[i]);
                    accr[i] += ar * br - ai * bi;
                    acci[i] += ar * bi + ai * br;
                }
            }
            for (int i = 0; i < m; ++i) {
                ocplx prod = MKC(accr[i], acci[i]);
                ocplx c = (RE(beta) == 0.0 && IM(beta) == 0.0) ? 0.0 : cmul(beta, C[i + (size_t)j * ldc]);
                C[i + (size_t)j * ldc] = cmul(alpha, prod) + c;
            }
        }
        free(accr); free(acci);
    }
    free(a); free(bb);
}

/* C = op(A) op(B) (fresh output) */
static void mm(char oa, char ob, int m, int n, int k, const ocplx* A, const ocplx* B, ocplx* C)
{
    int lda = (oa == 'N') ? m : k;
    int ldb = (ob == 'N') ? k : n;
    or_gemm(oa, ob, m, n, k, 1.0, A, lda, B, ldb, 0.0, C, m);
}

/* Largest 1-norm condition number ||A||_1 ||A^-1||_1 of the blocks inverted
 * on this thread since the last reset: the rounding scale by which two
 * correct FP64 algorithms may differ (used by the tests' tolerances). */
static _Thread_local double kappa_max = 0.0;

static double norm1(int n, const ocplx* A)
{
    double best = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += cabs(A[i + (size_t)j * n]);
        if (s > best) best = s;
    }
    return best;
}

/* ------------------------------------------------------------- inverse */
/* negf/util.f90:16-67: NaN scan, ZGETRF (partial pivoting on |re|+|im|, first
 * maximum wins as IZAMAX does), ZGETRI.  The inverse is formed by solving
 * LU X = P I column by column, which is the same matrix up to rounding. */
int or_inv(int n, const ocplx* A, ocplx* Ainv)
{
    size_t nn = (size_t)n * n;
    for (size_t i = 0; i < nn; ++i)
        if (isnan(RE(A[i])) || isnan(IM(A[i]))) return -1;
    ocplx* LU = cmalloc(nn);
    int* piv = (int*)malloc(sizeof(int) * (n ? n : 1));
    memcpy(LU, A, sizeof(ocplx) * nn);
    int info = 0;
    /* columns in parallel for the large blocks only (same arithmetic per entry) */
    const int par = n >= 320 && !omp_in_parallel();
    for (int c = 0; c < n; ++c) {
        int p = c;
        double best = cabs1(LU[c + (size_t)c * n]);
        for (int r = c + 1; r < n; ++r) {
            double v = cabs1(LU[r + (size_t)c * n]);
            if (v > best) { best = v; p = r; }
        }
        piv[c] = p;
        if (RE(LU[p + (size_t)c * n]) == 0.0 && IM(LU[p + (size_t)c * n]) == 0.0) {
            if (!info) info = c + 1;
            continue;
        }
        if (p != c)
            for (int j = 0; j < n; ++j) {
                ocplx t = LU[c + (size_t)j * n];
                LU[c + (size_t)j * n] = LU[p + (size_t)j * n];
                LU[p + (size_t)j * n] = t;
            }
        ocplx d = LU[c + (size_t)c * n];
        for (int r = c + 1; r < n; ++r) LU[r + (size_t)c * n] = cdiv(LU[r + (size_t)c * n], d);
#pragma omp parallel for schedule(static) if (par)
        for (int j = c + 1; j < n; ++j) {
            ocplx u = LU[c + (size_t)j * n];
            if (RE(u) == 0.0 && IM(u) == 0.0) continue;
            for (int r = c + 1; r < n; ++r)
                LU[r + (size_t)j * n] -= cmul(LU[r + (size_t)c * n], u);
        }
    }
    if (info) { free(LU); free(piv); return info; }
    /* X = inv(A): solve L U X = P I */
    for (size_t i = 0; i < nn; ++i) Ainv[i] = 0.0;
    for (int j = 0; j < n; ++j) Ainv[j + (size_t)j * n] = 1.0;
    for (int c = 0; c < n; ++c)
        if (piv[c] != c)
            for (int j = 0; j < n; ++j) {
                ocplx t = Ainv[c + (size_t)j * n];
                Ainv[c + (size_t)j * n] = Ainv[piv[c] + (size_t)j * n];
                Ainv[piv[c] + (size_t)j * n] = t;
            }
#pragma omp parallel for schedule(static) if (par)
    for (int j = 0; j < n; ++j) {
        ocplx* x = Ainv + (size_t)j * n;
        for (int c = 0; c < n; ++c) {           /* unit lower */
            ocplx xc = x[c];
            if (RE(xc) == 0.0 && IM(xc) == 0.0) continue;
            for (int r = c + 1; r < n; ++r) x[r] -= cmul(LU[r + (size_t)c * n], xc);
        }
        for (int c = n - 1; c >= 0; --c) {      /* upper */
            x[c] = cdiv(x[c], LU[c + (size_t)c * n]);
            ocplx xc = x[c];
            for (int r = 0; r < c; ++r) x[r] -= cmul(LU[r + (size_t)c * n], xc);
        }
    }
    free(LU); free(piv);
    {
        const double k = norm1(n, A) * norm1(n, Ainv);
        if (k > kappa_max || k != k) kappa_max = k;
    }
    return 0;
}

/* -------------------------------------------------------- decimation */
static double maxabs(size_t nn, const ocplx* A)
{
    double m = 0.0;
    for (size_t i = 0; i < nn; ++i) {
        double v = cabs(A[i]);              /* Fortran abs() of complex */
        if (v > m || isnan(v)) m = isnan(v) ? NAN : v;
        if (isnan(m)) return m;
    }
    return m;
}

/* z*I - X */
static void shift_neg(int n, ocplx z, const ocplx* X, ocplx* out)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            out[i + (size_t)j * n] = (i == j ? z : 0.0) - X[i + (size_t)j * n];
}

/* negf/surface.f90:482-497 (one of the four while-loops) and
 * negf/util.f90:762-798 (`decimation`, used by the circular variant). */
int or_decimate(int n, const ocplx* H00, const ocplx* alpha0, const ocplx* beta0,
                ocplx z, double conv, int max_iter, ocplx* g, int* iters)
{
    size_t nn = (size_t)n * n;
    ocplx *e1 = cmalloc(nn), *es = cmalloc(nn), *al = cmalloc(nn), *be = cmalloc(nn);
    ocplx *ive = cmalloc(nn), *t1 = cmalloc(nn), *t2 = cmalloc(nn), *w = cmalloc(nn), *w2 = cmalloc(nn);
    memcpy(e1, H00, sizeof(ocplx) * nn);
    memcpy(es, H00, sizeof(ocplx) * nn);
    memcpy(al, alpha0, sizeof(ocplx) * nn);
    memcpy(be, beta0, sizeof(ocplx) * nn);
    int status = 0, it = 0;
    double maxdiff = maxabs(nn, al);
    if (isnan(maxdiff)) status = -1;
    while (!status && maxdiff > conv) {
        if (it >= max_iter) { status = -2; break; }
        shift_neg(n, z, e1, w);
        status = or_inv(n, w, ive);
        if (status) break;
        mm('N', 'N', n, n, n, al, ive, t1);   /* alpha*ive */
        mm('N', 'N', n, n, n, be, ive, t2);   /* beta*ive  */
        mm('N', 'N', n, n, n, t1, be, w);     /* alpha*ive*beta */
        mm('N', 'N', n, n, n, t2, al, w2);    /* beta*ive*alpha */
        for (size_t i = 0; i < nn; ++i) { e1[i] = (e1[i] + w[i]) + w2[i]; es[i] = es[i] + w[i]; }
        mm('N', 'N', n, n, n, t1, al, w);     /* alpha*ive*alpha */
        mm('N', 'N', n, n, n, t2, be, w2);    /* beta*ive*beta */
        memcpy(al, w, sizeof(ocplx) * nn);
        memcpy(be, w2, sizeof(ocplx) * nn);
        ++it;
        maxdiff = maxabs(nn, al);
        if (isnan(maxdiff)) status = -1;
    }
    if (!status) {
        shift_neg(n, z, es, w);
        status = or_inv(n, w, g);
    }
    if (iters) *iters = it;
    free(e1); free(es); free(al); free(be); free(ive); free(t1); free(t2); free(w); free(w2);
    return status;
}

static void dagger(int m, int n, const ocplx* A, ocplx* out) /* A is m x n */
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) out[j + (size_t)i * n] = conj(A[i + (size_t)j * m]);
}

/* Gamma = i (S - S^+)  (negf/surface.f90:506, :568) */
static void gamma_of(int n, const ocplx* S, ocplx* G)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            ocplx d = S[i + (size_t)j * n] - conj(S[j + (size_t)i * n]);
            G[i + (size_t)j * n] = MKC(-IM(d), RE(d));
        }
}

/* negf/surface.f90:385-605 (tdb) and tb/surface.f90:129-343. */
int or_surface_green(int nl, int nr, const ocplx* Haml, const ocplx* Hamlv,
                     const ocplx* Hamr, const ocplx* Hamrv, double E,
                     const or_params* p, int want_plus_minus,
                     ocplx* gl_adv, ocplx* gr_ret, ocplx* gl_plus, ocplx* gr_minus,
                     ocplx* gml_adv, ocplx* gmr_ret, ocplx* sl_adv, ocplx* sr_ret,
                     int* iters)
{
    double eta = or_lead_eta(p, E);
    size_t nnl = (size_t)nl * nl, nnr = (size_t)nr * nr;
    ocplx *lio = cmalloc(nnl), *rio = cmalloc(nnr), *g = cmalloc(nnl > nnr ? nnl : nnr);
    ocplx *t = cmalloc(nnl > nnr ? nnl : nnr), *s = cmalloc(nnl > nnr ? nnl : nnr);
    int it[4] = {0, 0, 0, 0}, st = 0;
    /* Ham_l_i_o = Hamlv^+ ; Ham_r_o_i = Hamrv^+ (surface.f90:448-471) */
    dagger(nl, nl, Hamlv, lio);
    dagger(nr, nr, Hamrv, rio);

    /* left, advanced: alpha = Hamlv^+, beta = Hamlv, E - i eta */
    st = or_decimate(nl, Haml, lio, Hamlv, MKC(E, -eta), p->convergence, p->max_iter, g, &it[0]);
    if (st) goto done;
    if (gl_adv) memcpy(gl_adv, g, sizeof(ocplx) * nnl);
    mm('N', 'N', nl, nl, nl, lio, g, t);
    mm('N', 'N', nl, nl, nl, t, Hamlv, s);          /* sigmal_adv */
    if (sl_adv) memcpy(sl_adv, s, sizeof(ocplx) * nnl);
    if (gml_adv) gamma_of(nl, s, gml_adv);

    if (want_plus_minus) {
        /* left, reflected: alpha = Hamlv, beta = Hamlv^+, E + i eta */
        st = or_decimate(nl, Haml, Hamlv, lio, MKC(E, eta), p->convergence, p->max_iter, g, &it[1]);
        if (st) goto done;
        if (gl_plus) memcpy(gl_plus, g, sizeof(ocplx) * nnl);
    }
    /* right, retarded: alpha = Hamrv^+, beta = Hamrv, E + i eta */
    st = or_decimate(nr, Hamr, rio, Hamrv, MKC(E, eta), p->convergence, p->max_iter, g, &it[2]);
    if (st) goto done;
    if (gr_ret) memcpy(gr_ret, g, sizeof(ocplx) * nnr);
    mm('N', 'N', nr, nr, nr, rio, g, t);
    mm('N', 'N', nr, nr, nr, t, Hamrv, s);          /* sigmar_ret */
    if (sr_ret) memcpy(sr_ret, s, sizeof(ocplx) * nnr);
    if (gmr_ret) gamma_of(nr, s, gmr_ret);
    if (want_plus_minus) {
        /* right, reflected: alpha = Hamrv, beta = Hamrv^+, E - i eta */
        st = or_decimate(nr, Hamr, Hamrv, rio, MKC(E, -eta), p->convergence, p->max_iter, g, &it[3]);
        if (st) goto done;
        if (gr_minus) memcpy(gr_minus, g, sizeof(ocplx) * nnr);
    }
done:
    if (iters) memcpy(iters, it, sizeof(it));
    free(lio); free(rio); free(g); free(t); free(s);
    return st;
}

/* ------------------------------------------------------------- RGF */
/* out = (z I - H [- S1]) - M   (all n x n) */
static void dyson_lhs(int n, ocplx z, const ocplx* H, const ocplx* S1, const ocplx* M, ocplx* out)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            size_t k = i + (size_t)j * n;
            ocplx v = (i == j ? z : 0.0) - H[k];
            if (S1) v = v - S1[k];
            out[k] = v - M[k];
        }
}

static double caroli(int n1, int n2, const ocplx* G12, const ocplx* gm2, const ocplx* gm1)
{
    /* Re Tr[(G12 gm2)(G12^+ gm1)]  (negf/lesser.f90:787-793) */
    ocplx* X = cmalloc((size_t)n1 * n2);
    ocplx* Y = cmalloc((size_t)n2 * n1);
    mm('N', 'N', n1, n2, n2, G12, gm2, X);
    mm('C', 'N', n2, n1, n1, G12, gm1, Y);
    double tr = 0.0;
    for (int i = 0; i < n1; ++i) {
        double acc = 0.0;
        for (int j = 0; j < n2; ++j) acc += RE(cmul(X[i + (size_t)j * n1], Y[j + (size_t)i * n2]));
        tr += acc;
    }
    free(X); free(Y);
    return tr;
}

int or_rgf(int L, const int* bsz, const ocplx* const* Hd, const ocplx* const* Hdv,
           const ocplx* const* Hdv_back, int nl, int nr,
           const ocplx* Hld, const ocplx* Hrd, const ocplx* Hrd_fwd,
           const ocplx* Haml, const ocplx* Hamr,
           const ocplx* gl, const ocplx* gr, const ocplx* gml, const ocplx* gmr,
           const ocplx* sl, const ocplx* sr, double E, const or_params* p,
           int zero_transmission,
           ocplx* G_D_LR, ocplx* G_D_RL, ocplx* G_L, ocplx* G_R,
           double* T_LR, double* T_RL)
{
    int bmax = nl > nr ? nl : nr;
    for (int i = 0; i < L; ++i) if (bsz[i] > bmax) bmax = bsz[i];
    size_t big = (size_t)bmax * bmax;
    ocplx *G = cmalloc(big), *Gn = cmalloc(big), *A = cmalloc(big), *T1 = cmalloc(big), *M = cmalloc(big);
    ocplx *G0 = cmalloc(big), *G0n = cmalloc(big), *Vt = cmalloc(big);
    ocplx *GLR = cmalloc((size_t)nl * nr), *GRL = cmalloc((size_t)nr * nl);
    int st = 0;
    /* device eta' = factor * eta(E)  (lesser.f90:556; tb/device.f90:354) */
    ocplx zd = MKC(E, p->device_eta_factor * or_lead_eta(p, E));

    /* ---- left to right (lesser.f90:555-651, device.f90:330-432) ----
     * Every product uses a coupling block V and V^+, where V is the block
     * that couples layer i-1 to layer i (device.f90:347-352, :364-369). */
    {
        int b = bsz[0];
        mm('C', 'N', b, nl, nl, Hld, gl, T1);        /* Hld^+ gl */
        mm('N', 'N', b, b, nl, T1, Hld, M);
        dyson_lhs(b, zd, Hd[0], NULL, M, A);
        if ((st = or_inv(b, A, G))) goto out;
        mm('N', 'N', nl, b, nl, gl, Hld, T1);        /* gl Hld */
        mm('N', 'N', nl, b, b, T1, G, G0);           /* G_{0,1} */
        for (int i = 1; i < L; ++i) {
            int bp = bsz[i - 1], bi = bsz[i];
            const ocplx* V = Hdv[i - 1];              /* V_{i-1,i}: bp x bi */
            mm('C', 'N', bi, bp, bp, V, G, T1);
            mm('N', 'N', bi, bi, bp, T1, V, M);
            dyson_lhs(bi, zd, Hd[i], NULL, M, A);
            if ((st = or_inv(bi, A, Gn))) goto out;
            mm('N', 'N', nl, bi, bp, G0, V, T1);      /* G_{0,n} V */
            mm('N', 'N', nl, bi, bi, T1, Gn, G0n);    /* G_{0,n+1} */
            memcpy(G, Gn, sizeof(ocplx) * bi * bi);
            memcpy(G0, G0n, sizeof(ocplx) * nl * bi);
        }
        int bN = bsz[L - 1];
        /* closing right lead layer: Vc = V_{N,N+1} (bN x nr); tdb uses
         * hamrd^+ (lesser.f90:622), G_block the dense block (device.f90:408) */
        if (Hrd_fwd) memcpy(Vt, Hrd_fwd, sizeof(ocplx) * bN * nr);
        else dagger(nr, bN, Hrd, Vt);
        mm('C', 'N', nr, bN, bN, Vt, G, T1);
        mm('N', 'N', nr, nr, bN, T1, Vt, M);
        dyson_lhs(nr, zd, Hamr, sr, M, A);
        if ((st = or_inv(nr, A, Gn))) goto out;
        mm('N', 'N', nl, nr, bN, G0, Vt, T1);        /* G_{0,N} V_{N,N+1}: nl x nr */
        mm('N', 'N', nl, nr, nr, T1, Gn, GLR);
        if (G_R) memcpy(G_R, Gn, sizeof(ocplx) * nr * nr);
    }
    /* ---- right to left (lesser.f90:684-784, device.f90:434-525) ----
     * Coupling blocks here are V_{i,i-1}: hamdv^+ for tdb, the dense lower
     * block for G_block. */
    {
        int bN = bsz[L - 1];
        mm('C', 'N', bN, nr, nr, Hrd, gr, T1);        /* Hrd^+ gr */
        mm('N', 'N', bN, bN, nr, T1, Hrd, M);
        dyson_lhs(bN, zd, Hd[L - 1], NULL, M, A);
        if ((st = or_inv(bN, A, G))) goto out;
        mm('N', 'N', nr, bN, nr, gr, Hrd, T1);
        mm('N', 'N', nr, bN, bN, T1, G, G0);          /* G_{N+1,N} */
        for (int i = L - 1; i >= 1; --i) {
            int bi = bsz[i], bm = bsz[i - 1];
            /* V = V_{i,i-1} (bi x bm) */
            if (Hdv_back && Hdv_back[i - 1]) memcpy(Vt, Hdv_back[i - 1], sizeof(ocplx) * bi * bm);
            else dagger(bm, bi, Hdv[i - 1], Vt);
            mm('C', 'N', bm, bi, bi, Vt, G, T1);
            mm('N', 'N', bm, bm, bi, T1, Vt, M);
            dyson_lhs(bm, zd, Hd[i - 1], NULL, M, A);
            if ((st = or_inv(bm, A, Gn))) goto out;
            mm('N', 'N', nr, bm, bi, G0, Vt, T1);
            mm('N', 'N', nr, bm, bm, T1, Gn, G0n);
            memcpy(G, Gn, sizeof(ocplx) * bm * bm);
            memcpy(G0, G0n, sizeof(ocplx) * nr * bm);
        }
        int b1 = bsz[0];
        /* V = V_{0,1} = Hld (nl x b1); reference forms (V G) V^+ */
        mm('N', 'N', nl, b1, b1, Hld, G, T1);
        mm('N', 'C', nl, nl, b1, T1, Hld, M);
        dyson_lhs(nl, zd, Haml, sl, M, A);
        if ((st = or_inv(nl, A, Gn))) goto out;
        mm('N', 'C', nr, nl, b1, G0, Hld, T1);
        mm('N', 'N', nr, nl, nl, T1, Gn, GRL);
        if (G_L) memcpy(G_L, Gn, sizeof(ocplx) * nl * nl);
    }
    {
        double tl = caroli(nl, nr, GLR, gmr, gml);
        double tr = caroli(nr, nl, GRL, gml, gmr);
        if (tl < 1.0e-60) tl = 0.0;      /* lesser.f90:795-800 */
        if (tr < 1.0e-60) tr = 0.0;
        if (zero_transmission) { tl = 0.0; tr = 0.0; }   /* lesser.f90:802-807 */
        if (T_LR) *T_LR = tl;
        if (T_RL) *T_RL = tr;
    }
    if (G_D_LR) memcpy(G_D_LR, GLR, sizeof(ocplx) * nl * nr);
    if (G_D_RL) memcpy(G_D_RL, GRL, sizeof(ocplx) * nr * nl);
out:
    free(G); free(Gn); free(A); free(T1); free(M); free(G0); free(G0n); free(Vt);
    free(GLR); free(GRL);
    return st;
}

int or_solve_tdb(int nl, int nr, int b, int L,
                 const ocplx* Haml, const ocplx* Hamlv, const ocplx* Hamr, const ocplx* Hamrv,
                 const ocplx* Hamd, const ocplx* Hamdv, const ocplx* Hamld, const ocplx* Hamrd,
                 double E, const or_params* p, int zero_transmission,
                 double* T_LR, double* T_RL, double* iters_mean)
{
    size_t nnl = (size_t)nl * nl, nnr = (size_t)nr * nr;
    ocplx *gla = cmalloc(nnl), *grr = cmalloc(nnr), *gmla = cmalloc(nnl), *gmrr = cmalloc(nnr);
    ocplx *sla = cmalloc(nnl), *srr = cmalloc(nnr);
    ocplx *gl = cmalloc(nnl), *gml = cmalloc(nnl), *sl = cmalloc(nnl);
    int it[4];
    int st = or_surface_green(nl, nr, Haml, Hamlv, Hamr, Hamrv, E, p, 0,
                              gla, grr, NULL, NULL, gmla, gmrr, sla, srr, it);
    if (!st) {
        /* get_G_full_tdb_mpi passes gl_adv^+, -gml_adv, sl_adv^+ (lesser.f90:141-144) */
        dagger(nl, nl, gla, gl);
        dagger(nl, nl, sla, sl);
        for (size_t i = 0; i < nnl; ++i) gml[i] = -gmla[i];
        const ocplx** Hd = (const ocplx**)malloc(sizeof(ocplx*) * L);
        const ocplx** Hdv = (const ocplx**)malloc(sizeof(ocplx*) * (L > 1 ? L - 1 : 1));
        int* bsz = (int*)malloc(sizeof(int) * L);
        for (int i = 0; i < L; ++i) { Hd[i] = Hamd + (size_t)i * b * b; bsz[i] = b; }
        for (int i = 0; i + 1 < L; ++i) Hdv[i] = Hamdv + (size_t)i * b * b;
        st = or_rgf(L, bsz, Hd, Hdv, NULL, nl, nr, Hamld, Hamrd, NULL, Haml, Hamr,
                    gl, grr, gml, gmrr, sl, srr, E, p, zero_transmission,
                    NULL, NULL, NULL, NULL, T_LR, T_RL);
        free(Hd); free(Hdv); free(bsz);
    }
    if (iters_mean) *iters_mean = 0.5 * (it[0] + it[2]);
    free(gla); free(grr); free(gmla); free(gmrr); free(sla); free(srr); free(gl); free(gml); free(sl);
    return st;
}

int or_solve_tdb_batch(int nl, int nr, int b, int L,
                       const ocplx* Haml, const ocplx* Hamlv, const ocplx* Hamr, const ocplx* Hamrv,
                       const ocplx* Hamd, const ocplx* Hamdv, const ocplx* Hamld, const ocplx* Hamrd,
                       int npts, const double* E, const or_params* p, int nthreads,
                       double* T_LR, double* T_RL, int* status)
{
    int worst = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(min : worst)
#endif
    for (int i = 0; i < npts; ++i) {
        int st = or_solve_tdb(nl, nr, b, L, Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd,
                              E[i], p, 0, &T_LR[i], &T_RL[i], NULL);
        if (status) status[i] = st;
        if (st < worst) worst = st;
        if (st > 0 && worst == 0) worst = st;
    }
    (void)nthreads;
    return worst;
}

/* or_solve_tdb_batch plus, per point, the largest condition number of the
 * blocks its solve inverted (kappa_max above). */
int or_solve_tdb_batch_kappa(int nl, int nr, int b, int L,
                             const ocplx* Haml, const ocplx* Hamlv, const ocplx* Hamr, const ocplx* Hamrv,
                             const ocplx* Hamd, const ocplx* Hamdv, const ocplx* Hamld, const ocplx* Hamrd,
                             int npts, const double* E, const or_params* p, int nthreads,
                             double* T_LR, double* T_RL, int* status, double* kappa)
{
    int worst = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(min : worst)
#endif
    for (int i = 0; i < npts; ++i) {
        kappa_max = 0.0;
        int st = or_solve_tdb(nl, nr, b, L, Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd,
                              E[i], p, 0, &T_LR[i], &T_RL[i], NULL);
        kappa[i] = kappa_max;
        if (status) status[i] = st;
        if (st < worst) worst = st;
        if (st > 0 && worst == 0) worst = st;
    }
    (void)nthreads;
    return worst;
}

/* ------------------------------------------------------- circular */
/* negf/util.f90:732-760: diagonal blocks of the block-DFT of a block-circulant
 * matrix, from its first block row; phase(i,j) = theta*(i)/nb*(j) with theta = G.R */
static void ft_blocks(int n, int nb, double theta, const ocplx* H, int ldh, ocplx* out, int ldo)
{
    int na = n / nb;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) out[i + (size_t)j * ldo] = 0.0;
    for (int bi = 0; bi < nb; ++bi)
        for (int bj = 0; bj < nb; ++bj) {
            double ph = theta * (double)bi / (double)nb * (double)bj;
            ocplx e = MKC(cos(ph), sin(ph));
            for (int c = 0; c < na; ++c)
                for (int r = 0; r < na; ++r)
                    out[(bi * na + r) + (size_t)(bi * na + c) * ldo] +=
                        cmul(H[r + (size_t)(bj * na + c) * ldh], e);
        }
}

/* P of set_pmatrix (negf/circular.f90:9-40) */
static void pmatrix(int n, int nb, double theta, ocplx* P)
{
    int na = n / nb;
    double s = 1.0 / sqrt((double)nb);
    for (size_t i = 0; i < (size_t)n * n; ++i) P[i] = 0.0;
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
            double ph = theta * (double)j / (double)nb * (double)i;
            for (int k = 0; k < na; ++k)
                P[(i * na + k) + (size_t)(j * na + k) * n] = MKC(cos(ph) * s, sin(ph) * s);
        }
}

/* negf/circular.f90:43-114.  nby copies along y (outer), nbx along x (inner).
 * theta_y = reci_uc(2,:).latvec_uc(2,:), theta_x likewise (2*pi for a lattice).
 * The decimation inside uses beta = alpha^+ (util.f90:788). */
static int circular_g_theta(int n, int nby, int nbx, double theta_y, double theta_x,
                            const ocplx* H00, const ocplx* alpha0, ocplx z, double conv,
                            int max_iter, ocplx* g, int* iters)
{
    int na = n / nby, nac = na / nbx;
    size_t nn = (size_t)n * n;
    ocplx *e1 = cmalloc(nn), *al = cmalloc(nn), *g1 = cmalloc(nn), *tmp = cmalloc(nn);
    ocplx *P1 = cmalloc(nn), *P2 = cmalloc((size_t)na * na);
    ocplx *ey = cmalloc((size_t)na * na), *ay = cmalloc((size_t)na * na), *g2 = cmalloc((size_t)na * na);
    ocplx *h = cmalloc((size_t)nac * nac), *a = cmalloc((size_t)nac * nac), *bq = cmalloc((size_t)nac * nac);
    ocplx *gs = cmalloc((size_t)nac * nac), *t2 = cmalloc((size_t)na * na);
    int st = 0, itmax = 0;
    pmatrix(n, nby, theta_y, P1);
    pmatrix(na, nbx, theta_x, P2);
    ft_blocks(n, nby, theta_y, H00, n, e1, n);
    ft_blocks(n, nby, theta_y, alpha0, n, al, n);
    for (size_t i = 0; i < nn; ++i) g1[i] = 0.0;
    for (int i = 0; i < nby && !st; ++i) {
        const ocplx* eb = e1 + (size_t)i * na + (size_t)i * na * n;
        const ocplx* ab = al + (size_t)i * na + (size_t)i * na * n;
        ft_blocks(na, nbx, theta_x, eb, n, ey, na);
        ft_blocks(na, nbx, theta_x, ab, n, ay, na);
        for (size_t q = 0; q < (size_t)na * na; ++q) g2[q] = 0.0;
        for (int j = 0; j < nbx; ++j) {
            for (int c = 0; c < nac; ++c)
                for (int r = 0; r < nac; ++r) {
                    h[r + (size_t)c * nac] = ey[(j * nac + r) + (size_t)(j * nac + c) * na];
                    a[r + (size_t)c * nac] = ay[(j * nac + r) + (size_t)(j * nac + c) * na];
                }
            dagger(nac, nac, a, bq);
            int it = 0;
            st = or_decimate(nac, h, a, bq, z, conv, max_iter, gs, &it);
            if (it > itmax) itmax = it;
            if (st) break;
            for (int c = 0; c < nac; ++c)
                for (int r = 0; r < nac; ++r)
                    g2[(j * nac + r) + (size_t)(j * nac + c) * na] = gs[r + (size_t)c * nac];
        }
        if (st) break;
        mm('N', 'N', na, na, na, P2, g2, t2);
        mm('N', 'C', na, na, na, t2, P2, g2);
        for (int c = 0; c < na; ++c)
            for (int r = 0; r < na; ++r)
                g1[(i * na + r) + (size_t)(i * na + c) * n] = g2[r + (size_t)c * na];
    }
    if (!st) {
        mm('N', 'N', n, n, n, P1, g1, tmp);
        mm('N', 'C', n, n, n, tmp, P1, g);
    }
    if (iters) *iters = itmax;
    free(e1); free(al); free(g1); free(tmp); free(P1); free(P2); free(ey); free(ay); free(g2);
    free(h); free(a); free(bq); free(gs); free(t2);
    return st;
}

int or_circular_g(int n, int nby, int nbx, const ocplx* H00, const ocplx* alpha0,
                  const ocplx* beta0, ocplx z, double conv, int max_iter,
                  ocplx* g, int* iters)
{
    (void)beta0;   /* the reference's `decimation` ignores beta_in (util.f90:788) */
    const double twopi = 6.283185307179586476925286766559;
    return circular_g_theta(n, nby, nbx, twopi, twopi, H00, alpha0, z, conv, max_iter, g, iters);
}

/* ------------------------------------------------------------- layer LDOS */
#define OR_PI 3.14159265358979323846
/* get_ldos (tb/device.f90:610-855): left-connected G^L_i (L->R from gl),
 * right-connected G^R_i (R->L from gr), then per layer
 *   GLR_i = (I - G^L_i V_{i,i+1} G^R_{i+1} V_{i,i+1}^+)^{-1} G^L_i
 * with the right lead closing the last layer (V = Hrd_fwd or Hrd^+, G^R = gr),
 * ldos(i, j) = -Im GLR_i(j, j) / pi (zero beyond the layer size), ddos = sum.
 * The R->L sweep couples layer i to i-1 with the stored lower block
 * (Hdv_back) when given (:768-773), the GLR step with V_{i,i+1}^+ (:813-821). */
int or_get_ldos(int L, const int* bsz, const ocplx* const* Hd, const ocplx* const* Hdv,
                const ocplx* const* Hdv_back, int nl, int nr,
                const ocplx* Hld, const ocplx* Hrd, const ocplx* Hrd_fwd,
                const ocplx* gl, const ocplx* gr, double E, const or_params* p,
                int maxlen, double* ldos, double* ddos)
{
    int bmax = nl > nr ? nl : nr;
    for (int i = 0; i < L; ++i) if (bsz[i] > bmax) bmax = bsz[i];
    size_t big = (size_t)bmax * bmax;
    ocplx** GL = (ocplx**)malloc(sizeof(ocplx*) * L);
    for (int i = 0; i < L; ++i) GL[i] = cmalloc(big);
    ocplx *GR = cmalloc(big), *GRn = cmalloc(big), *A = cmalloc(big), *T1 = cmalloc(big), *M = cmalloc(big);
    ocplx *Vt = cmalloc(big), *X = cmalloc(big);
    int st = 0;
    ocplx zd = MKC(E, p->device_eta_factor * or_lead_eta(p, E));
    for (size_t k = 0; k < (size_t)L * maxlen; ++k) ldos[k] = 0.0;
    /* L -> R */
    mm('C', 'N', bsz[0], nl, nl, Hld, gl, T1);
    mm('N', 'N', bsz[0], bsz[0], nl, T1, Hld, M);
    dyson_lhs(bsz[0], zd, Hd[0], NULL, M, A);
    if ((st = or_inv(bsz[0], A, GL[0]))) goto out;
    for (int i = 1; i < L; ++i) {
        int bp = bsz[i - 1], bi = bsz[i];
        mm('C', 'N', bi, bp, bp, Hdv[i - 1], GL[i - 1], T1);
        mm('N', 'N', bi, bi, bp, T1, Hdv[i - 1], M);
        dyson_lhs(bi, zd, Hd[i], NULL, M, A);
        if ((st = or_inv(bi, A, GL[i]))) goto out;
    }
    /* last layer: V = Vc (bN x nr), partner gr */
    {
        int bN = bsz[L - 1];
        if (Hrd_fwd) memcpy(Vt, Hrd_fwd, sizeof(ocplx) * bN * nr);
        else dagger(nr, bN, Hrd, Vt);
        /* R->L start: G^R_N = (z - H_N - Hrd^+ gr Hrd)^-1, Hrd = V_{N+1,N} */
        mm('C', 'N', bN, nr, nr, Hrd, gr, T1);
        mm('N', 'N', bN, bN, nr, T1, Hrd, M);
        dyson_lhs(bN, zd, Hd[L - 1], NULL, M, A);
        if ((st = or_inv(bN, A, GR))) goto out;
        /* GLR_N with the lead: I - G^L_N Vc gr Vc^+ */
        mm('N', 'N', bN, nr, bN, GL[L - 1], Vt, T1);
        mm('N', 'N', bN, nr, nr, T1, gr, M);
        mm('N', 'C', bN, bN, nr, M, Vt, A);
        for (size_t k = 0; k < (size_t)bN * bN; ++k) A[k] = -A[k];
        for (int j = 0; j < bN; ++j) A[j + (size_t)j * bN] += 1.0;
        if ((st = or_inv(bN, A, M))) goto out;
        mm('N', 'N', bN, bN, bN, M, GL[L - 1], X);
        for (int j = 0; j < bN && j < maxlen; ++j) ldos[(L - 1) + (size_t)j * L] = -IM(X[j + (size_t)j * bN]) / OR_PI;
    }
    for (int i = L - 1; i >= 1; --i) {
        int bi = bsz[i], bm = bsz[i - 1];
        /* V_{i,i-1}: stored lower block or Hdv^+ ; G^R_{i-1} */
        if (Hdv_back && Hdv_back[i - 1]) memcpy(Vt, Hdv_back[i - 1], sizeof(ocplx) * bi * bm);
        else dagger(bm, bi, Hdv[i - 1], Vt);
        /* GLR_{i-1} uses G^R_i (current GR) and V_{i-1,i} = Hdv[i-1] */
        mm('N', 'N', bm, bi, bm, GL[i - 1], Hdv[i - 1], T1);
        mm('N', 'N', bm, bi, bi, T1, GR, M);
        mm('N', 'C', bm, bm, bi, M, Hdv[i - 1], A);
        for (size_t k = 0; k < (size_t)bm * bm; ++k) A[k] = -A[k];
        for (int j = 0; j < bm; ++j) A[j + (size_t)j * bm] += 1.0;
        if ((st = or_inv(bm, A, M))) goto out;
        mm('N', 'N', bm, bm, bm, M, GL[i - 1], X);
        for (int j = 0; j < bm && j < maxlen; ++j) ldos[(i - 1) + (size_t)j * L] = -IM(X[j + (size_t)j * bm]) / OR_PI;
        /* G^R_{i-1} = (z - H_{i-1} - V^+ G^R_i V), V = V_{i,i-1} */
        mm('C', 'N', bm, bi, bi, Vt, GR, T1);
        mm('N', 'N', bm, bm, bi, T1, Vt, M);
        dyson_lhs(bm, zd, Hd[i - 1], NULL, M, A);
        if ((st = or_inv(bm, A, GRn))) goto out;
        ocplx* t = GR; GR = GRn; GRn = t;
    }
    if (ddos) {
        double acc = 0.0;
        for (size_t k = 0; k < (size_t)L * maxlen; ++k) acc += ldos[k];
        *ddos = acc;
    }
out:
    for (int i = 0; i < L; ++i) free(GL[i]);
    free(GL); free(GR); free(GRn); free(A); free(T1); free(M); free(Vt); free(X);
    return st;
}
