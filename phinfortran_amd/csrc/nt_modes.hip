// nt_modes.hip -- mode-resolved transmission on the batched solve's outputs:
// the lead Bloch modes from the eigenproblems of g V (batched Schur +
// back-substitution), the propagating-mode selection and Gram-Schmidt, the
// degenerate-subspace rotation of path_mode = 1, velocity matrices, their
// square roots, the t/r matrices and per-mode sums (vel_mat_tdb,
// negf/velocity_mat.f90:1692-3150; tb vel_mat, tb/velocity_mat.f90:222-1560),
// the Brillouin-zone unfolding of the modes and the primitive-cell k and
// group velocities (find_k_pc), and the batched eigen / sqrtm building blocks
// of the C ABI.
#include "nt_host.h"

// ------------------------------------------------------- eigen / sqrtm
// Scratch: T and Z (and X) per item, n^2 complex each.
size_t eig_scratch(int n, int batch)
{
    Arena a;
    for (int i = 0; i < 3; ++i) a.take(sizeof(double) * 2 * (size_t)n * n * batch);
    return a.off;
}

// w, V from A (A is copied, not modified).  The eigenvector matrix is Z X
// with X the triangular eigenvectors of T.
void eig_run(nt_ctx* c, int n, BMat<double> A, double* w, BMatW<double> V, int* status, int batch, Arena& ar,
             size_t off)
{
    typedef double T;
    const long nn = (long)n * n;
    Arena sa = ar;
    sa.off = off;
    T* Tm = sa.ptr<T>(sa.take(sizeof(T) * 2 * nn * batch));
    T* Zm = sa.ptr<T>(sa.take(sizeof(T) * 2 * nn * batch));
    T* Xm = sa.ptr<T>(sa.take(sizeof(T) * 2 * nn * batch));
    copy_op<T>(c, n, n, A, bw(Tm, nn, n), 0, (T)1, batch);
    EigArgs e;
    e.n = n;
    e.A = bw(Tm, nn, n);
    e.Z = bw(Zm, nn, n);
    e.X = bw(Xm, nn, n);
    e.V = V;
    e.w = w;
    e.status = status;
    e.batch = batch;
    audit<T>(c, e.A, n, n, batch, 1, "schur T");
    audit<T>(c, e.Z, n, n, batch, 1, "schur Z");
    audit<T>(c, e.X, n, n, batch, 1, "trevc X");
    audit<T>(c, V, n, n, batch, 1, "eig V");
    check_range(c, w, sizeof(T) * 2 * (size_t)n * batch, "eig w");
    check_range(c, status, sizeof(int) * batch, "eig status");
    launch(c, NT_K_EIG, 0.0, "schur", [&] { return launch_schur(e, c->stream); });
    launch(c, NT_K_EIG, 0.0, "trevc", [&] { return launch_trevc(e, c->stream); });
    gemm<T>(c, n, n, n, OP_N, ro(e.Z), OP_N, ro(e.X), V, 1.0, none<T>(), 0.0, nullptr, batch);
    launch(c, NT_K_EIG, 0.0, "ev normalize", [&] { return launch_ev_normalize(e, c->stream); });
}

// S = sqrtm(A) (A copied)
void sqrtm_run(nt_ctx* c, int n, BMat<double> A, BMatW<double> S, int* status, int batch, Arena& ar, size_t off)
{
    typedef double T;
    const long nn = (long)n * n;
    Arena sa = ar;
    sa.off = off;
    T* Tm = sa.ptr<T>(sa.take(sizeof(T) * 2 * nn * batch));
    T* Zm = sa.ptr<T>(sa.take(sizeof(T) * 2 * nn * batch));
    T* Xm = sa.ptr<T>(sa.take(sizeof(T) * 2 * nn * batch));
    copy_op<T>(c, n, n, A, bw(Tm, nn, n), 0, (T)1, batch);
    EigArgs e;
    e.n = n;
    e.A = bw(Tm, nn, n);
    e.Z = bw(Zm, nn, n);
    e.X = bw(Xm, nn, n);
    e.V = bw(Xm, nn, n);
    e.w = nullptr;
    e.status = status;
    e.batch = batch;
    audit<T>(c, e.A, n, n, batch, 1, "schur T");
    audit<T>(c, e.Z, n, n, batch, 1, "schur Z");
    audit<T>(c, S, n, n, batch, 1, "sqrtm S");
    check_range(c, status, sizeof(int) * batch, "sqrtm status");
    launch(c, NT_K_EIG, 0.0, "schur", [&] { return launch_schur(e, c->stream); });
    launch(c, NT_K_EIG, 0.0, "sqrt tri", [&] { return launch_sqrt_tri(e, c->stream); });
    // S = (Z U) Z^H ; Z U into the T buffer
    gemm<T>(c, n, n, n, OP_N, ro(e.Z), OP_N, ro(e.X), e.A, 1.0, none<T>(), 0.0, nullptr, batch);
    gemm<T>(c, n, n, n, OP_N, ro(e.A), OP_C, ro(e.Z), S, 1.0, none<T>(), 0.0, nullptr, batch);
}

// ------------------------------------------------------ mode-resolved T
// vel_mat_tdb (negf/velocity_mat.f90:1692-3150, tb/velocity_mat.f90:222-1560)
// for a batch of points, without the Brillouin-zone unfolding.  Per point and
// lead direction: F = g V, eigen-decomposition, propagating selection + k_z +
// Gram-Schmidt, velocity matrix V = s U^+ Gamma' U, its square root, inv(U^+)
// (incident) or inv(U) (outgoing); then the bulk GFs Q_L, Q_R and
//   t_RL = f Vh_Rret iU_Rret G_D_RL iU_Ladv Vh_Ladv,  r_LL = f Vh_Lret iU_Lret (G_L - Q_L^-1) iU_Ladv Vh_Ladv
// and the mirrored t_LR, r_RR (negf/velocity_mat.f90:2945-2963).  negf
// conventions: s = -/+ i a_z / (2 sqrt(E)), f = 2 i sqrt(E / (a a')); tb: s = -/+1,
// f = i (tb/velocity_mat.f90:467-468, 1352-1370).

struct ModeOut {
    double* trl;      // nr x nl per point (scratch when the caller passes none)
    double* tlr;      // nl x nr
    double* rrr;      // nr x nr
    double* rll;      // nl x nl
    double* kz[4];    // L_adv, L_ret, R_ret, R_adv: n per point
    int* nprop[4];    // counts per point
    double* t_mode_L; // nl per point: sum over propagating outgoing right modes of |t_RL(j,i)|^2
    double* t_mode_R; // nr per point: same for t_LR
    double* sums[4];  // T_modes_RL, T_modes_LR, R_modes_R, R_modes_L per point
};

// per-point complex scalars, one array of P per slot (slot q at out + 2 q P):
// 0-3 the V scales of the four directions, 4-7 the prefactors of t_RL, t_LR,
// r_RR, r_LL (zero at the reference's ie = 1 point), 8 the lead shift
// z = E + i eta(E)
__global__ void k_mode_scalars(const double* E, int P, double az_l, double az_r, int tb, double eta, double eta0,
                               int eta_mode, const int* first_mask, double* out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double e = E[p];
    auto put = [&](int q, double re, double im) {
        out[2 * ((long)q * P + p)] = re;
        out[2 * ((long)q * P + p) + 1] = im;
    };
    const double sq = sqrt(e);
    const double vl = tb ? 1.0 : az_l / (2.0 * sq), vr = tb ? 1.0 : az_r / (2.0 * sq);
    // V: L_adv, L_ret: -s ; R_ret, R_adv: +s ; s = i a/(2 sqrt E) (negf) or 1 (tb)
    const double sv[4] = {-vl, -vl, vr, vr};
    for (int q = 0; q < 4; ++q) put(q, tb ? sv[q] : 0.0, tb ? 0.0 : sv[q]);
    const bool zero = first_mask && first_mask[p];
    const double f[4] = {tb ? 1.0 : 2.0 * sqrt(e / (az_r * az_l)), tb ? 1.0 : 2.0 * sqrt(e / (az_l * az_r)),
                         tb ? 1.0 : 2.0 * sqrt(e / (az_r * az_r)), tb ? 1.0 : 2.0 * sqrt(e / (az_l * az_l))};
    for (int q = 0; q < 4; ++q) put(4 + q, 0.0, zero ? 0.0 : f[q]);
    put(8, e, eta_mode == 0 ? e * eta + eta0 : eta);
}

// out[p * rows + q] = in[q * cols + p]   (rows x cols slot-major -> point-major)
template <typename T>
__global__ void k_transpose_small(const T* in, int rows, int cols, T* out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)rows * cols) return;
    const int q = (int)(e / cols), p = (int)(e % cols);
    out[(long)p * rows + q] = in[e];
}

size_t modes_scratch(int nl, int nr, int P)
{
    const int n = std::max(nl, nr);
    const size_t mb = sizeof(double) * 2 * (size_t)n * n * P;
    Arena a;
    for (int i = 0; i < 4; ++i) {
        a.take(mb);                                    // U
        a.take(mb);                                    // Vh
        a.take(mb);                                    // invU
        a.take(sizeof(double) * 2 * (size_t)n * P);   // w
        a.take(sizeof(double) * (size_t)n * P);       // kz (when not output)
    }
    a.take(mb);  // QinvL / scratch
    a.take(mb);  // QinvR
    a.take(mb);  // T1
    a.take(mb);  // T2
    a.take(mb);  // V / W
    for (int i = 0; i < 4; ++i) a.take(mb);   // t/r scratch
    a.take(sizeof(double) * 18 * P);
    for (int i = 0; i < 4; ++i) {
        a.take(sizeof(int) * (size_t)n * P);       // unfolding labels
        a.take(sizeof(double) * 3 * (size_t)n * P);  // unfolded k
    }
    return a.off + eig_scratch(n, P);
}

// Unfolding request of one chunk of points (outputs already offset to it;
// null outputs use scratch).
// find_degenerate (negf/util.f90:513-587) with its 1-based indices: from
// start, the first pair i < j with |l_i - l_j| < 1e-4 and ||l_j| - 1| < 1e-3
// gives the group of every index within 1e-4 of l_j, or {0} if one of them
// was processed before.  Returns whether a pair was found.
bool find_degenerate(const std::vector<std::complex<double>>& lam, int start, const std::vector<int>& deg_all,
                     std::vector<int>& group, int& next_start)
{
    const int ne = (int)lam.size();
    int found = 0, itemp = 0;
    std::complex<double> edeg;
    for (int i = start; i <= ne - 1 && !found; ++i)
        for (int j = i + 1; j <= ne; ++j) {
            if (std::abs(lam[i - 1] - lam[j - 1]) < 1.0e-4 && std::abs(std::abs(lam[j - 1]) - 1.0) < 1.0e-3) {
                found = 1;
                itemp = i;
                edeg = lam[j - 1];
            }
            if (found) break;
        }
    next_start = itemp + 1;
    group.assign(1, 0);
    if (found) {
        group.clear();
        for (int i = 1; i <= ne; ++i)
            if (std::abs(lam[i - 1] - edeg) < 1.0e-4) group.push_back(i);
        for (int g : group)
            for (int d : deg_all)
                if (g == d) {
                    group.assign(1, 0);
                    return true;
                }
    }
    return found != 0;
}

// The path_mode = 1 rotation of one mode set (negf/velocity_mat.f90:1884-1978):
// groups found on the host from the eigenvalues w (m per point), then per
// group the Hermitian H = H00 + V / lam + V^+ lam (upper triangle), its Schur
// (eigen) decomposition, the least-squares coefficients and the rotation of
// U's group columns (nt_degen.hip); finally every propagating column of U is
// normalised.  Uses three scratch slots of mm x P complex from off.
void degenerate_rotation(nt_ctx* c, int m, int P, long mm, const double* E, BMat<double> H00, BMat<double> V,
                         const double* w, double* U, const double* kz, int* status, Arena& ar, size_t off)
{
    typedef double T;
    hipStream_t s = c->stream;
    std::vector<double> hw(2 * (size_t)m * P);
    check_range(c, w, sizeof(T) * hw.size(), "rotation eigenvalues");
    hipcheck(c, hipMemcpyAsync(hw.data(), w, sizeof(T) * hw.size(), hipMemcpyDeviceToHost, s), "D2H");
    hipcheck(c, hipStreamSynchronize(s), "sync");
    std::vector<DegGroup> groups;
    std::vector<std::complex<double>> lam(m);
    std::vector<int> group, deg_all;
    for (int p = 0; p < P; ++p) {
        for (int i = 0; i < m; ++i) lam[i] = {hw[2 * ((size_t)p * m + i)], hw[2 * ((size_t)p * m + i) + 1]};
        deg_all.assign(1, 0);
        bool first = true;
        int start = 1, next = 1;
        while (find_degenerate(lam, start, deg_all, group, next)) {
            if (group.size() > 1) {
                if (group.size() > (size_t)DEG_MAX) {
                    c->err = "more than 16 degenerate modes in one group (path_mode = 1)";
                    throw Fail{NT_ERR_ARG};
                }
                if (first) deg_all.clear();
                first = false;
                deg_all.insert(deg_all.end(), group.begin(), group.end());
                DegGroup g{};
                g.p = p;
                g.ndeg = (int)group.size();
                for (size_t k = 0; k < group.size(); ++k) g.idx[k] = group[k] - 1;
                g.lr = lam[group[0] - 1].real();
                g.li = lam[group[0] - 1].imag();
                groups.push_back(g);
            }
            start = next;
        }
    }
    const int G = (int)groups.size();
    if (G > 0) {
        const size_t need = (sizeof(DegGroup) + sizeof(int)) * (size_t)std::min(G, P);
        if (c->deg_bytes < need) {
            if (c->deg_buf) {
                hipcheck(c, hipStreamSynchronize(s), "sync");
                unreg_region(c, c->deg_buf);
                hipcheck(c, hipFree(c->deg_buf), "hipFree");
                c->deg_buf = nullptr;
                c->deg_bytes = 0;
            }
            hipcheck(c, hipMalloc(&c->deg_buf, need), "hipMalloc rotation groups");
            c->deg_bytes = need;
            reg_region(c, c->deg_buf, need);
        }
        Arena sa = ar;
        sa.off = off;
        T* Hb = sa.ptr<T>(sa.take(sizeof(T) * 2 * mm * P));
        T* Zb = sa.ptr<T>(sa.take(sizeof(T) * 2 * mm * P));
        T* Xb = sa.ptr<T>(sa.take(sizeof(T) * 2 * mm * P));
        DegGroup* gd = static_cast<DegGroup*>(c->deg_buf);
        int* gst = reinterpret_cast<int*>(gd + std::min(G, P));
        for (int g0 = 0; g0 < G; g0 += P) {
            const int Gc = std::min(P, G - g0);
            hipcheck(c, hipMemcpyAsync(gd, groups.data() + g0, sizeof(DegGroup) * Gc, hipMemcpyHostToDevice, s),
                     "H2D groups");
            hipcheck(c, launch_fill_int(gst, 0, Gc, s), "zero group status");
            DegArgs da;
            da.m = m;
            da.G = Gc;
            da.H00 = H00;
            da.V = V;
            da.g = gd;
            da.E = E;
            da.H = Hb;
            da.Z = Zb;
            da.hstride = mm;
            da.U = U;
            da.ustride = mm;
            da.gstatus = gst;
            da.status = status;
            audit<T>(c, H00, m, m, P, 1, "rotation H00");
            audit<T>(c, V, m, m, P, 1, "rotation V");
            audit<T>(c, bw(Hb, mm, m), m, m, Gc, 1, "rotation H");
            audit<T>(c, bw(U, mm, m), m, m, P, 1, "rotation U");
            check_range(c, E, sizeof(T) * P, "rotation E");
            check_range(c, status, sizeof(int) * P, "rotation status");
            launch(c, NT_K_EIG, 0.0, "rotation build", [&] { return launch_deg_build(da, s); });
            EigArgs e;
            e.n = m;
            e.A = bw(Hb, mm, m);
            e.Z = bw(Zb, mm, m);
            e.X = bw(Xb, mm, m);
            e.V = nonew<T>();
            e.w = nullptr;
            e.status = gst;
            e.batch = Gc;
            launch(c, NT_K_EIG, 0.0, "rotation schur", [&] { return launch_schur(e, s); });
            launch(c, NT_K_EIG, 0.0, "rotation solve", [&] { return launch_deg_solve(da, s); });
            hipcheck(c, hipStreamSynchronize(s), "sync");     // groups vector reused by the next chunk
        }
    }
    audit<T>(c, bw(U, mm, m), m, m, P, 1, "rotation U");
    check_range(c, kz, sizeof(T) * (size_t)m * P, "rotation kz");
    launch(c, NT_K_ELEM, 0.0, "rotation normalize", [&] { return launch_deg_normalize(m, U, mm, kz, P, s); });
}

struct UnfoldRun {
    const nt_unfold_desc* d;
    int* bz[4];          // L_adv, L_ret, R_ret, R_adv
    double* kxyz[4];
    double* split;       // 8 per point
    double* rows[4];     // t_RL, t_LR, r_RR, r_LL: (total, specular) per incident mode
};

void modes_run(nt_ctx* c, const nt_params* prm, int tb, int nl, int nr, int P, const double* E, const int* kidx,
               int nk, BMat<double> Haml, BMat<double> Hamlv, BMat<double> Hamr, BMat<double> Hamrv,
               BMat<double> gl_adv, BMat<double> gr_ret, BMat<double> gl_plus, BMat<double> gr_minus,
               BMat<double> gml_adv, BMat<double> gmr_ret, BMat<double> G_D_RL, BMat<double> G_D_LR,
               BMat<double> G_L, BMat<double> G_R, double az_l, double az_r, const int* first_mask, ModeOut o,
               int* status, Arena& ar, size_t off, const UnfoldRun* uf = nullptr)
{
    typedef double T;
    const int n = std::max(nl, nr);
    const long mm = (long)n * n;
    const size_t mb = sizeof(T) * 2 * (size_t)mm * P;
    Arena sa = ar;
    sa.off = off;
    T *U[4], *Vh[4], *iU[4], *w[4], *kzs[4];
    for (int i = 0; i < 4; ++i) {
        U[i] = sa.ptr<T>(sa.take(mb));
        Vh[i] = sa.ptr<T>(sa.take(mb));
        iU[i] = sa.ptr<T>(sa.take(mb));
        w[i] = sa.ptr<T>(sa.take(sizeof(T) * 2 * (size_t)n * P));
        kzs[i] = sa.ptr<T>(sa.take(sizeof(T) * (size_t)n * P));
    }
    T* QL = sa.ptr<T>(sa.take(mb));
    T* QR = sa.ptr<T>(sa.take(mb));
    T* T1 = sa.ptr<T>(sa.take(mb));
    T* T2 = sa.ptr<T>(sa.take(mb));
    T* Vw = sa.ptr<T>(sa.take(mb));
    T* tr[4];
    for (int i = 0; i < 4; ++i) tr[i] = sa.ptr<T>(sa.take(mb));
    T* sc = sa.ptr<T>(sa.take(sizeof(T) * 18 * P));
    int* bzs[4];
    double* kxs[4];
    for (int i = 0; i < 4; ++i) {
        bzs[i] = sa.ptr<int>(sa.take(sizeof(int) * (size_t)n * P));
        kxs[i] = sa.ptr<double>(sa.take(sizeof(double) * 3 * (size_t)n * P));
    }
    const size_t eig_off = sa.off;
    hipStream_t s = c->stream;
    auto W = [&](T* q, int ld) { return bw(q, mm, ld); };
    auto R = [&](T* q, int ld) { return bm<T>(q, mm, ld); };
    check_range(c, E, sizeof(double) * P, "modes E");
    check_range(c, first_mask, sizeof(int) * P, "modes first mask");
    hipLaunchKernelGGL(k_mode_scalars, dim3(nblk(P)), dim3(256), 0, s, E, P, az_l, az_r, tb, prm->eta, prm->eta0,
                       prm->eta_mode, first_mask, sc);
    hipcheck(c, hipGetLastError(), "mode scalars");
    // the four directions: (n, g, op(g), V lead coupling, Gamma', op(Gamma'), k_z sign, incident?)
    struct Dir {
        int n;
        BMat<T> g;
        int opg;
        BMat<T> V;
        BMat<T> gam;
        int opgam;
        double sign;
        bool incident;
        double az;
    };
    const Dir dirs[4] = {{nl, gl_adv, OP_N, Hamlv, gml_adv, OP_N, -1.0, true, az_l},
                         {nl, gl_adv, OP_C, Hamlv, gml_adv, OP_C, -1.0, false, az_l},
                         {nr, gr_ret, OP_N, Hamrv, gmr_ret, OP_N, 1.0, false, az_r},
                         {nr, gr_ret, OP_C, Hamrv, gmr_ret, OP_C, 1.0, true, az_r}};
    for (int d = 0; d < 4; ++d) {
        const Dir& D = dirs[d];
        const int m = D.n;
        double* kz = o.kz[d] ? o.kz[d] : kzs[d];
        // F = op(g) V  -> U (eigen input, overwritten by the eigenvectors)
        gemm<T>(c, m, m, m, D.opg, D.g, OP_N, D.V, W(U[d], m), 1.0, none<T>(), 0.0, nullptr, P);
        eig_run(c, m, R(U[d], m), w[d], W(U[d], m), status, P, ar, eig_off);
        ModesArgs ma;
        ma.n = m;
        ma.w = w[d];
        ma.U = W(U[d], m);
        ma.kz = kz;
        ma.nprop = o.nprop[d];
        ma.lambda_eta = prm->lambda_eta;
        ma.az = D.az;
        ma.sign = D.sign;
        ma.batch = P;
        audit<T>(c, ma.U, m, m, P, 1, "modes U");
        check_range(c, kz, sizeof(double) * (size_t)m * P, "modes kz");
        check_range(c, o.nprop[d], sizeof(int) * P, "modes nprop");
        launch(c, NT_K_EIG, 0.0, "modes select", [&] { return launch_modes_select(ma, s); });
        if (prm->path_mode == 1)
            degenerate_rotation(c, m, P, mm, E, d < 2 ? Haml : Hamr, D.V, w[d], U[d], kz, status, ar, eig_off);
        // V = s (U^+ Gamma' U)
        gemm<T>(c, m, m, m, OP_C, R(U[d], m), D.opgam, D.gam, W(T1, m), 1.0, none<T>(), 0.0, nullptr, P);
        gemm<T>(c, m, m, m, OP_N, R(T1, m), OP_N, R(U[d], m), W(Vw, m), 1.0, none<T>(), 0.0, nullptr, P);
        launch(c, NT_K_ELEM, 0.0, "scale V",
               [&] { return launch_scale_items(Vw, mm, m, m, m, sc + 2 * (long)d * P, P, s); });
        sqrtm_run(c, m, R(Vw, m), W(Vh[d], m), status, P, ar, eig_off);
        // inv(U^+) for incident directions, inv(U) for outgoing ones
        copy_op<T>(c, m, m, R(U[d], m), W(T2, m), D.incident ? 1 : 0, (T)1, P);
        inverse<T>(c, m, W(T2, m), none<T>(), nullptr, W(iU[d], m), status, P);
    }
    // bulk GFs: Q = z - H00 - V^+ op(g) V - V op(g') V^+, inverted
    {
        // left: z - Haml - Hamlv^+ gl_adv^+ Hamlv - Hamlv gl_plus Hamlv^+
        gemm<T>(c, nl, nl, nl, OP_C, Hamlv, OP_C, gl_adv, W(T1, nl), 1.0, none<T>(), 0.0, nullptr, P);
        gemm<T>(c, nl, nl, nl, OP_N, R(T1, nl), OP_N, Hamlv, W(T2, nl), -1.0, Haml, -1.0, sc + 16 * (long)P, P);
        gemm<T>(c, nl, nl, nl, OP_N, Hamlv, OP_N, gl_plus, W(T1, nl), 1.0, none<T>(), 0.0, nullptr, P);
        gemm<T>(c, nl, nl, nl, OP_N, R(T1, nl), OP_C, Hamlv, W(T2, nl), -1.0, R(T2, nl), 1.0, nullptr, P);
        inverse<T>(c, nl, W(T2, nl), none<T>(), nullptr, W(QL, nl), status, P);
        // right: z - Hamr - Hamrv^+ gr_ret Hamrv - Hamrv gr_minus^+ Hamrv^+
        gemm<T>(c, nr, nr, nr, OP_C, Hamrv, OP_N, gr_ret, W(T1, nr), 1.0, none<T>(), 0.0, nullptr, P);
        gemm<T>(c, nr, nr, nr, OP_N, R(T1, nr), OP_N, Hamrv, W(T2, nr), -1.0, Hamr, -1.0, sc + 16 * (long)P, P);
        gemm<T>(c, nr, nr, nr, OP_N, Hamrv, OP_C, gr_minus, W(T1, nr), 1.0, none<T>(), 0.0, nullptr, P);
        gemm<T>(c, nr, nr, nr, OP_N, R(T1, nr), OP_C, Hamrv, W(T2, nr), -1.0, R(T2, nr), 1.0, nullptr, P);
        inverse<T>(c, nr, W(T2, nr), none<T>(), nullptr, W(QR, nr), status, P);
    }
    // t and r: out = f Vh_o iU_o M iU_i Vh_i   (o: outgoing, i: incident)
    struct Chain {
        int o, i;        // direction indices
        BMat<T> M;       // middle matrix (G block)
        T* Qinv;         // subtracted bulk inverse (reflections) or null
        double* out;
        int q;           // prefactor index
    };
    const Chain ch[4] = {{2, 0, G_D_RL, nullptr, o.trl ? o.trl : tr[0], 0},
                         {1, 3, G_D_LR, nullptr, o.tlr ? o.tlr : tr[1], 1},
                         {2, 3, G_R, QR, o.rrr ? o.rrr : tr[2], 2},
                         {1, 0, G_L, QL, o.rll ? o.rll : tr[3], 3}};
    for (const Chain& C : ch) {
        const int mo = dirs[C.o].n, mi = dirs[C.i].n;
        const long so = (long)mo * mi;
        BMatW<T> outw = bw(C.out, so, mo);
        audit<T>(c, outw, mo, mi, P, 1, "modes t/r");
        gemm<T>(c, mi, mi, mi, OP_N, R(iU[C.i], mi), OP_N, R(Vh[C.i], mi), W(T1, mi), 1.0, none<T>(), 0.0, nullptr,
                P);
        if (C.Qinv)   // M - Q^-1 into T2 (square: mo == mi)
            copy_op<T>(c, mo, mi, R(C.Qinv, mo), W(T2, mo), 0, (T)-1, P);
        gemm<T>(c, mo, mi, mi, OP_N, C.M, OP_N, R(T1, mi), W(Vw, mo), 1.0, none<T>(), 0.0, nullptr, P);
        if (C.Qinv) {
            // (M - Q^-1) T1 = M T1 + (-Q^-1) T1
            gemm<T>(c, mo, mi, mi, OP_N, R(T2, mo), OP_N, R(T1, mi), W(Vw, mo), 1.0, R(Vw, mo), 1.0, nullptr, P);
        }
        gemm<T>(c, mo, mi, mo, OP_N, R(iU[C.o], mo), OP_N, R(Vw, mo), W(T1, mo), 1.0, none<T>(), 0.0, nullptr, P);
        gemm<T>(c, mo, mi, mo, OP_N, R(Vh[C.o], mo), OP_N, R(T1, mo), outw, 1.0, none<T>(), 0.0, nullptr, P);
        launch(c, NT_K_ELEM, 0.0, "scale t",
               [&] { return launch_scale_items(C.out, so, mo, mo, mi, sc + 2 * (long)(4 + C.q) * P, P, s); });
        double* kzo = o.kz[C.o] ? o.kz[C.o] : kzs[C.o];
        double* kzi = o.kz[C.i] ? o.kz[C.i] : kzs[C.i];
        double* colsum = C.q == 0 ? o.t_mode_L : (C.q == 1 ? o.t_mode_R : nullptr);
        check_range(c, colsum, sizeof(double) * (size_t)mi * P, "modes t_mode");
        check_range(c, o.sums[C.q], sizeof(double) * P, "modes sums");
        launch(c, NT_K_ELEM, 0.0, "mode sums", [&] {
            return launch_mode_sums(C.out, so, mo, mo, mi, kzo, kzi, colsum, o.sums[C.q], P, s);
        });
    }
    if (!uf) return;
    // Brillouin-zone unfolding (negf/velocity_mat.f90:2020-2130 and mirrors, :3032-3099)
    const nt_unfold_desc& D = *uf->d;
    const int nb = D.n_bloch_x * D.n_bloch_y;
    for (int d = 0; d < 4; ++d) {
        const int m = dirs[d].n;
        const bool left = d < 2;
        UnfoldDirArgs ua;
        ua.n = m;
        ua.nuc = m / nb;
        ua.nbx = D.n_bloch_x;
        ua.nby = D.n_bloch_y;
        ua.U = U[d];
        ua.ustride = mm;
        ua.ldu = m;
        ua.w = w[d];
        ua.kz = o.kz[d] ? o.kz[d] : kzs[d];
        ua.ksc = D.ksc;
        ua.kidx = kidx;
        for (int q = 0; q < 9; ++q) {
            ua.recivec[q] = D.recivec[q];
            ua.reci_uc[q] = left ? D.reci_uc_l[q] : D.reci_uc_r[q];
        }
        ua.pos = left ? D.pos_l : D.pos_r;
        ua.pos_uc = left ? D.pos_uc_l : D.pos_uc_r;
        ua.bz = uf->bz[d] ? uf->bz[d] : bzs[d];
        ua.kxyz = uf->kxyz[d] ? uf->kxyz[d] : kxs[d];
        ua.batch = P;
        bzs[d] = ua.bz;
        kxs[d] = ua.kxyz;
        check_range(c, ua.bz, sizeof(int) * (size_t)m * P, "unfold bz");
        check_range(c, ua.kxyz, sizeof(double) * 3 * (size_t)m * P, "unfold k");
        launch(c, NT_K_ELEM, 0.0, "unfold weights", [&] { return launch_unfold_weights(ua, s); });
    }
    UnfoldSplitArgs sp;
    const int slot[4] = {1, 0, 3, 2};   // t_RL -> transr, t_LR -> transl, r_RR -> reflr, r_LL -> refll
    for (int q = 0; q < 4; ++q) {
        const Chain& C = ch[q];
        UnfoldChain& u = sp.ch[q];
        u.M = C.out;
        u.mo = dirs[C.o].n;
        u.mi = dirs[C.i].n;
        u.bz_row = bzs[C.o];
        u.bz_col = bzs[C.i];
        u.k_row = kxs[C.o];
        u.k_col = kxs[C.i];
        u.rows = uf->rows[q];
        u.slot = slot[q];
        check_range(c, u.rows, sizeof(double) * 2 * (size_t)u.mi * P, "unfold rows");
    }
    sp.split = uf->split;
    sp.batch = P;
    check_range(c, sp.split, sizeof(double) * 8 * (size_t)P, "unfold split");
    launch(c, NT_K_ELEM, 0.0, "unfold split", [&] { return launch_unfold_split(sp, s); });
}

// find_k_pc for nmodes modes (negf/dyn.f90:770-889; tb/hamiltonian.f90:666-916):
// candidates per mode, the matrices of each candidate (mats: the dynamical
// matrices at k and k +- 1e-5 e_d, or the tight-binding H and HF_d), their
// eigenpairs and the selection.  Runs inside guarded().
template <typename Mats>
int unfold_pc_run(nt_ctx* c, KpcArgs& a, int nmodes, const double* k_xyz, Mats mats)
{
    const int n_pc = a.n;
    // modes per chunk; candidates per matrix batch (7 matrices each, kept under
    // 512 MB and 65536 phase-sum launches rows)
    const int Mmax = 8192;
    const size_t per_cand = 7 * (size_t)n_pc * n_pc * 2 * sizeof(double) + 7 * 3 * sizeof(double);
    const int NCsub = (int)std::max<size_t>(1, std::min<size_t>(9000, (size_t)(512u << 20) / per_cand));
    Arena pr;
    const size_t o_cnt = pr.take(sizeof(int) * Mmax);
    const size_t o_cand = pr.take(sizeof(double) * 3 * (size_t)Mmax * a.maxc);
    const size_t o_off = pr.take(sizeof(int) * Mmax);
    const size_t o_cm = pr.take(sizeof(int) * (size_t)Mmax * a.maxc);
    const size_t o_cs = pr.take(sizeof(int) * (size_t)Mmax * a.maxc);
    const size_t o_kc = pr.take(sizeof(double) * 3 * (size_t)Mmax * a.maxc);
    const size_t o_found = pr.take(sizeof(int) * (size_t)Mmax * a.maxc);
    const size_t o_vc = pr.take(sizeof(double) * 3 * (size_t)Mmax * a.maxc);
    const size_t o_kl = pr.take(sizeof(double) * 21 * (size_t)NCsub);
    const size_t o_D = pr.take(sizeof(double) * 14 * (size_t)NCsub * n_pc * n_pc);
    ensure_ws(c, pr.off);
    Arena ar;
    ar.base = static_cast<char*>(c->ws);
    a.count = ar.ptr<int>(o_cnt);
    a.cand = ar.ptr<double>(o_cand);
    int* doff = ar.ptr<int>(o_off);
    int* dcm = ar.ptr<int>(o_cm);
    int* dcs = ar.ptr<int>(o_cs);
    double* dkc = ar.ptr<double>(o_kc);
    int* dfound = ar.ptr<int>(o_found);
    double* dvc = ar.ptr<double>(o_vc);
    double* dkl = ar.ptr<double>(o_kl);
    double* dD = ar.ptr<double>(o_D);
    std::vector<int> cnt, off, cm, cs;
    for (int m0 = 0; m0 < nmodes; m0 += Mmax) {
        const int M = std::min(Mmax, nmodes - m0);
        check_range(c, k_xyz + 3L * m0, sizeof(double) * 3 * (size_t)M, "unfold_pc k");
        launch(c, NT_K_ELEM, 0.0, "kpc candidates", [&] { return launch_kpc_candidates(a, m0, M, c->stream); });
        cnt.resize(M);
        hipcheck(c, hipMemcpyAsync(cnt.data(), a.count, sizeof(int) * M, hipMemcpyDeviceToHost, c->stream), "D2H");
        hipcheck(c, hipStreamSynchronize(c->stream), "sync");
        off.assign(M, 0);
        cm.clear();
        cs.clear();
        for (int i = 0; i < M; ++i) {
            off[i] = (int)cm.size();
            const int nci = cnt[i] == -2 ? 1 : std::max(0, cnt[i]);
            if (nci > a.maxc) {
                c->err = "unfold_pc: candidate count out of range";
                return (int)NT_ERR_HIP;
            }
            for (int q = 0; q < nci; ++q) {
                cm.push_back(i);
                cs.push_back(q);
            }
        }
        const int nc = (int)cm.size();
        hipcheck(c, hipMemcpyAsync(doff, off.data(), sizeof(int) * M, hipMemcpyHostToDevice, c->stream), "H2D");
        if (nc > 0) {
            hipcheck(c, hipMemcpyAsync(dcm, cm.data(), sizeof(int) * nc, hipMemcpyHostToDevice, c->stream), "H2D");
            hipcheck(c, hipMemcpyAsync(dcs, cs.data(), sizeof(int) * nc, hipMemcpyHostToDevice, c->stream), "H2D");
        }
        for (int c0 = 0; c0 < nc; c0 += NCsub) {
            const int ncs = std::min(NCsub, nc - c0);
            mats(c0, ncs, dcm, dcs, dkl, dkc, dD);
            launch(c, NT_K_EIG, 0.0, "kpc eigenpairs",
                   [&] { return launch_kpc_eig(a, dcm, m0, c0, ncs, dD, dfound, dvc, c->stream); });
        }
        launch(c, NT_K_ELEM, 0.0, "kpc select",
               [&] { return launch_kpc_select(a, m0, M, doff, dfound, dvc, dkc, c->stream); });
        // the host vectors are reused by the next chunk: its copies must have landed
        hipcheck(c, hipStreamSynchronize(c->stream), "sync");
    }
    return 0;
}

// the mode-level arguments every find_k_pc entry shares
inline KpcArgs kpc_args(int tb, int n_pc, const double* E, const double* k_xyz, int n_bloch_max,
                        const double* reci_uc, const double* recivec_pc, double* kp, double* v, int* ncand)
{
    KpcArgs a;
    a.n = n_pc;
    a.tb = tb;
    a.E = E;
    a.k = k_xyz;
    for (int q = 0; q < 9; ++q) {
        a.reci_uc[q] = reci_uc[q];
        a.recivec_pc[q] = recivec_pc[q];
    }
    a.l = n_bloch_max;
    a.maxc = (2 * a.l + 1) * (2 * a.l + 1) * (2 * a.l + 1);
    a.kp = kp;
    a.v = v;
    a.ncand = ncand;
    return a;
}

extern "C" {

static int modes_entry(nt_ctx* c, const nt_params* p, int tb, int nl, int nr, int npts, const double* E,
                       const int* kidx, int nk, const nt_c128* Haml, const nt_c128* Hamlv, const nt_c128* Hamr,
                       const nt_c128* Hamrv, const nt_c128* gl_adv, const nt_c128* gr_ret, const nt_c128* gl_plus,
                       const nt_c128* gr_minus, const nt_c128* gml_adv, const nt_c128* gmr_ret,
                       const nt_c128* G_D_RL, const nt_c128* G_D_LR, const nt_c128* G_L, const nt_c128* G_R,
                       double a_z_L, double a_z_R, const int* first_point_mask, const nt_modes_out* out,
                       const nt_unfold_desc* uf, const nt_unfold_out* uo, int* status)
{
    if (!c || check_params(p) || (tb != 0 && tb != 1)) return NT_ERR_ARG;
    if (nl <= 0 || nr <= 0 || npts < 0 || nk <= 0 || (nk > 1 && !kidx) || !(a_z_L > 0.0) || !(a_z_R > 0.0))
        return NT_ERR_ARG;
    if (!E || !Haml || !Hamlv || !Hamr || !Hamrv || !gl_adv || !gr_ret || !gl_plus || !gr_minus || !gml_adv ||
        !gmr_ret || !G_D_RL || !G_D_LR || !G_L || !G_R)
        return NT_ERR_ARG;
    if (std::max(nl, nr) > inv_max_n()) return NT_ERR_ARG;
    if (uf) {
        const int nb = uf->n_bloch_x * uf->n_bloch_y;
        if (uf->n_bloch_x <= 0 || uf->n_bloch_y <= 0 || nb > NT_MAX_BLOCH || nl % nb || nr % nb || !uf->ksc ||
            !uf->pos_l || !uf->pos_uc_l || !uf->pos_r || !uf->pos_uc_r)
            return NT_ERR_ARG;
    }
    if (npts == 0) return NT_OK;
    return guarded(c, [&] {
        const size_t z = sizeof(nt_c128);
        const long nll = (long)nl * nl, nrr = (long)nr * nr, nlr = (long)nl * nr;
        nt_unfold_out u = uo ? *uo : nt_unfold_out{};
        if (uf) {
            const int nb = uf->n_bloch_x * uf->n_bloch_y;
            reg_region(c, uf->ksc, sizeof(double) * 3 * nk);
            reg_region(c, uf->pos_l, sizeof(double) * 3 * nl);
            reg_region(c, uf->pos_r, sizeof(double) * 3 * nr);
            reg_region(c, uf->pos_uc_l, sizeof(double) * 3 * (nl / nb));
            reg_region(c, uf->pos_uc_r, sizeof(double) * 3 * (nr / nb));
            const int nn[4] = {nl, nl, nr, nr};
            int* bz[4] = {u.bz_L_adv, u.bz_L_ret, u.bz_R_ret, u.bz_R_adv};
            double* kx[4] = {u.k_xyz_L_adv, u.k_xyz_L_ret, u.k_xyz_R_ret, u.k_xyz_R_adv};
            for (int d = 0; d < 4; ++d) {
                reg_region(c, bz[d], sizeof(int) * (size_t)nn[d] * npts);
                reg_region(c, kx[d], sizeof(double) * 3 * (size_t)nn[d] * npts);
            }
            reg_region(c, u.split, sizeof(double) * 8 * npts);
            reg_region(c, u.spec_trl, sizeof(double) * 2 * (size_t)nl * npts);
            reg_region(c, u.spec_tlr, sizeof(double) * 2 * (size_t)nr * npts);
            reg_region(c, u.spec_rrr, sizeof(double) * 2 * (size_t)nr * npts);
            reg_region(c, u.spec_rll, sizeof(double) * 2 * (size_t)nl * npts);
        }
        reg_region(c, E, sizeof(double) * npts);
        reg_region(c, kidx, sizeof(int) * npts);
        reg_region(c, first_point_mask, sizeof(int) * npts);
        reg_region(c, status, sizeof(int) * npts);
        reg_region(c, Haml, z * nll * nk);
        reg_region(c, Hamlv, z * nll * nk);
        reg_region(c, Hamr, z * nrr * nk);
        reg_region(c, Hamrv, z * nrr * nk);
        const nt_c128* pl[] = {gl_adv, gl_plus, gml_adv, G_L};
        const nt_c128* pr[] = {gr_ret, gr_minus, gmr_ret, G_R};
        for (auto q : pl) reg_region(c, q, z * nll * npts);
        for (auto q : pr) reg_region(c, q, z * nrr * npts);
        reg_region(c, G_D_RL, z * nlr * npts);
        reg_region(c, G_D_LR, z * nlr * npts);
        nt_modes_out o = out ? *out : nt_modes_out{};
        reg_region(c, o.trl, z * nlr * npts);
        reg_region(c, o.tlr, z * nlr * npts);
        reg_region(c, o.rrr, z * nrr * npts);
        reg_region(c, o.rll, z * nll * npts);
        reg_region(c, o.k_z_L_adv, sizeof(double) * nl * npts);
        reg_region(c, o.k_z_L_ret, sizeof(double) * nl * npts);
        reg_region(c, o.k_z_R_ret, sizeof(double) * nr * npts);
        reg_region(c, o.k_z_R_adv, sizeof(double) * nr * npts);
        reg_region(c, o.n_prop, sizeof(int) * 4 * npts);
        reg_region(c, o.t_mode_L, sizeof(double) * nl * npts);
        reg_region(c, o.t_mode_R, sizeof(double) * nr * npts);
        reg_region(c, o.sums, sizeof(double) * 4 * npts);
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
        // per-point scratch for the counts and sums of a chunk (outputs are point-major)
        const int n = std::max(nl, nr);
        const double per_pt = (double)modes_scratch(nl, nr, 1) + 64.0 * n;
        const int Pmax = std::max(1, std::min(std::min(npts, c->batch_limit), (int)std::min(1.0e9, 16.0e9 / per_pt)));
        Arena probe;
        const size_t o_np = probe.take(sizeof(int) * 4 * Pmax);
        const size_t o_sum = probe.take(sizeof(double) * 4 * Pmax);
        const size_t o_st = probe.take(sizeof(int) * Pmax);
        const size_t o_run = probe.off;
        ensure_ws(c, o_run + modes_scratch(nl, nr, Pmax));
        Arena ar;
        ar.base = static_cast<char*>(c->ws);
        int* np4 = ar.ptr<int>(o_np);
        double* sum4 = ar.ptr<double>(o_sum);
        int* stc = ar.ptr<int>(o_st);
        typedef double T;
        auto kst = [&](long per_k) { return kidx ? per_k : 0L; };
        for (int p0 = 0; p0 < npts; p0 += Pmax) {
            const int P = std::min(Pmax, npts - p0);
            const int* kk = kidx ? kidx + p0 : nullptr;
            auto lead = [&](const nt_c128* b, long per_k, int ld) {
                return bm<T>(reinterpret_cast<const T*>(b), kst(per_k), ld, kk, 0, nk);
            };
            auto pt = [&](const nt_c128* b, long per, int ld) {
                return bm<T>(reinterpret_cast<const T*>(b) + 2 * (long)p0 * per, per, ld);
            };
            auto ptw = [&](nt_c128* b, long per) {
                return b ? reinterpret_cast<double*>(b) + 2 * (long)p0 * per : nullptr;
            };
            auto ptd = [&](double* b, long per) { return b ? b + (long)p0 * per : nullptr; };
            ModeOut mo;
            mo.trl = ptw(o.trl, nlr);
            mo.tlr = ptw(o.tlr, nlr);
            mo.rrr = ptw(o.rrr, nrr);
            mo.rll = ptw(o.rll, nll);
            mo.kz[0] = ptd(o.k_z_L_adv, nl);
            mo.kz[1] = ptd(o.k_z_L_ret, nl);
            mo.kz[2] = ptd(o.k_z_R_ret, nr);
            mo.kz[3] = ptd(o.k_z_R_adv, nr);
            for (int d = 0; d < 4; ++d) mo.nprop[d] = o.n_prop ? np4 + d * P : nullptr;
            mo.t_mode_L = ptd(o.t_mode_L, nl);
            mo.t_mode_R = ptd(o.t_mode_R, nr);
            for (int q = 0; q < 4; ++q) mo.sums[q] = o.sums ? sum4 + q * P : nullptr;
            zero_dev(c, stc, sizeof(int) * P, c->stream);
            UnfoldRun ur;
            if (uf) {
                auto pti = [&](int* b, long per) { return b ? b + (long)p0 * per : nullptr; };
                ur.d = uf;
                ur.bz[0] = pti(u.bz_L_adv, nl);
                ur.bz[1] = pti(u.bz_L_ret, nl);
                ur.bz[2] = pti(u.bz_R_ret, nr);
                ur.bz[3] = pti(u.bz_R_adv, nr);
                ur.kxyz[0] = ptd(u.k_xyz_L_adv, 3L * nl);
                ur.kxyz[1] = ptd(u.k_xyz_L_ret, 3L * nl);
                ur.kxyz[2] = ptd(u.k_xyz_R_ret, 3L * nr);
                ur.kxyz[3] = ptd(u.k_xyz_R_adv, 3L * nr);
                ur.split = ptd(u.split, 8);
                ur.rows[0] = ptd(u.spec_trl, 2L * nl);
                ur.rows[1] = ptd(u.spec_tlr, 2L * nr);
                ur.rows[2] = ptd(u.spec_rrr, 2L * nr);
                ur.rows[3] = ptd(u.spec_rll, 2L * nl);
            }
            modes_run(c, p, tb, nl, nr, P, E + p0, kk, nk, lead(Haml, nll, nl), lead(Hamlv, nll, nl),
                      lead(Hamr, nrr, nr), lead(Hamrv, nrr, nr), pt(gl_adv, nll, nl), pt(gr_ret, nrr, nr),
                      pt(gl_plus, nll, nl), pt(gr_minus, nrr, nr), pt(gml_adv, nll, nl), pt(gmr_ret, nrr, nr),
                      pt(G_D_RL, nlr, nr), pt(G_D_LR, nlr, nl), pt(G_L, nll, nl), pt(G_R, nrr, nr), a_z_L, a_z_R,
                      first_point_mask ? first_point_mask + p0 : nullptr, mo, stc, ar, o_run, uf ? &ur : nullptr);
            // counts and sums: slot-major scratch -> point-major outputs
            if (o.n_prop) {
                hipLaunchKernelGGL(k_transpose_small<int>, dim3(nblk(4L * P)), dim3(256), 0, c->stream, np4, 4, P,
                                   o.n_prop + 4L * p0);
            }
            if (o.sums) {
                hipLaunchKernelGGL(k_transpose_small<double>, dim3(nblk(4L * P)), dim3(256), 0, c->stream, sum4, 4,
                                   P, o.sums + 4L * p0);
            }
            if (status)
                hipcheck(c, hipMemcpyAsync(status + p0, stc, sizeof(int) * P, hipMemcpyDeviceToDevice, c->stream),
                         "status copy");
            hipcheck(c, hipGetLastError(), "modes outputs");
        }
        hipcheck(c, hipStreamSynchronize(c->stream), "sync");
        return 0;
    });
}

int nt_modes_tdb_batched_dev(nt_ctx* c, const nt_params* p, int tb, int nl, int nr, int npts, const double* E,
                             const int* kidx, int nk, const nt_c128* Haml, const nt_c128* Hamlv,
                             const nt_c128* Hamr, const nt_c128* Hamrv, const nt_c128* gl_adv,
                             const nt_c128* gr_ret, const nt_c128* gl_plus, const nt_c128* gr_minus,
                             const nt_c128* gml_adv, const nt_c128* gmr_ret, const nt_c128* G_D_RL,
                             const nt_c128* G_D_LR, const nt_c128* G_L, const nt_c128* G_R, double a_z_L,
                             double a_z_R, const int* first_point_mask, const nt_modes_out* out, int* status)
{
    return modes_entry(c, p, tb, nl, nr, npts, E, kidx, nk, Haml, Hamlv, Hamr, Hamrv, gl_adv, gr_ret, gl_plus,
                       gr_minus, gml_adv, gmr_ret, G_D_RL, G_D_LR, G_L, G_R, a_z_L, a_z_R, first_point_mask, out,
                       nullptr, nullptr, status);
}

int nt_modes_unfold_tdb_batched_dev(nt_ctx* c, const nt_params* p, int tb, int nl, int nr, int npts,
                                    const double* E, const int* kidx, int nk, const nt_c128* Haml,
                                    const nt_c128* Hamlv, const nt_c128* Hamr, const nt_c128* Hamrv,
                                    const nt_c128* gl_adv, const nt_c128* gr_ret, const nt_c128* gl_plus,
                                    const nt_c128* gr_minus, const nt_c128* gml_adv, const nt_c128* gmr_ret,
                                    const nt_c128* G_D_RL, const nt_c128* G_D_LR, const nt_c128* G_L,
                                    const nt_c128* G_R, double a_z_L, double a_z_R, const int* first_point_mask,
                                    const nt_modes_out* out, const nt_unfold_desc* uf, const nt_unfold_out* uo,
                                    int* status)
{
    if (!uf) return NT_ERR_ARG;
    return modes_entry(c, p, tb, nl, nr, npts, E, kidx, nk, Haml, Hamlv, Hamr, Hamrv, gl_adv, gr_ret, gl_plus,
                       gr_minus, gml_adv, gmr_ret, G_D_RL, G_D_LR, G_L, G_R, a_z_L, a_z_R, first_point_mask, out,
                       uf, uo, status);
}

int nt_unfold_pc_dev(nt_ctx* c, int nmodes, const double* E, const double* k_xyz, int n_bloch_max,
                     const double* reci_uc, const double* recivec_pc, int n_pc, int nseg, const int* seg_off,
                     const int* seg_rc, int nent, const double* coef, const double* dr, double* kp, double* v,
                     int* ncand)
{
    if (!c || nmodes < 0 || n_bloch_max < 0 || n_bloch_max > 4 || !reci_uc || !recivec_pc || n_pc <= 0 ||
        n_pc % 3 || n_pc > KPC_MAX_N || nseg < 0 || nent < 0 || (nseg > 0 && (!seg_off || !seg_rc || !coef || !dr)))
        return NT_ERR_ARG;
    if (nmodes == 0) return NT_OK;
    if (!E || !k_xyz || !kp || !v || !ncand) return NT_ERR_ARG;
    return guarded(c, [&] {
        if (nseg > 0) {
            std::vector<int> off(nseg + 1), rc(2 * (size_t)nseg);
            hipcheck(c, hipMemcpy(off.data(), seg_off, sizeof(int) * (nseg + 1), hipMemcpyDeviceToHost), "D2H");
            hipcheck(c, hipMemcpy(rc.data(), seg_rc, sizeof(int) * 2 * nseg, hipMemcpyDeviceToHost), "D2H");
            if (off[0] != 0 || off[nseg] != nent) {
                c->err = "unfold_pc: segment offsets must run from 0 to nent";
                return (int)NT_ERR_ARG;
            }
            for (int q = 0; q < nseg; ++q)
                if (off[q + 1] < off[q] || rc[2 * q] < 0 || 3 * rc[2 * q] + 3 > n_pc || rc[2 * q + 1] < 0 ||
                    3 * rc[2 * q + 1] + 3 > n_pc) {
                    c->err = "unfold_pc: segment outside the matrix";
                    return (int)NT_ERR_ARG;
                }
        }
        reg_region(c, E, sizeof(double) * nmodes);
        reg_region(c, k_xyz, sizeof(double) * 3 * (size_t)nmodes);
        reg_region(c, kp, sizeof(double) * 3 * (size_t)nmodes);
        reg_region(c, v, sizeof(double) * 3 * (size_t)nmodes);
        reg_region(c, ncand, sizeof(int) * (size_t)nmodes);
        reg_region(c, seg_off, sizeof(int) * (nseg + 1));
        reg_region(c, seg_rc, sizeof(int) * 2 * nseg);
        reg_region(c, coef, sizeof(double) * 9 * (size_t)nent);
        reg_region(c, dr, sizeof(double) * 3 * (size_t)nent);
        KpcArgs a = kpc_args(0, n_pc, E, k_xyz, n_bloch_max, reci_uc, recivec_pc, kp, v, ncand);
        return unfold_pc_run(c, a, nmodes, k_xyz,
                             [&](int c0, int ncs, const int* dcm, const int* dcs, double* dkl, double* dkc,
                                 double* dD) {
                                 launch(c, NT_K_ELEM, 0.0, "kpc klist", [&] {
                                     return launch_kpc_klist(a, dcm, dcs, c0, ncs, dkl, dkc, c->stream);
                                 });
                                 launch(c, NT_K_ELEM, 0.0, "kpc dynamical matrices", [&] {
                                     return launch_hk_assemble(7 * ncs, dkl, nseg, seg_off, seg_rc, coef, dr, n_pc,
                                                               n_pc, dD, n_pc, (long)n_pc * n_pc, c->stream);
                                 });
                             });
    });
}

int nt_unfold_pc_tb_dev(nt_ctx* c, int nmodes, const double* E, const double* k_xyz, int n_bloch_max,
                        const double* reci_uc, const double* recivec_pc, int n_pc, const int* el_off, int nent,
                        const double* coef, const double* dr, double* kp, double* v, int* ncand)
{
    if (!c || nmodes < 0 || n_bloch_max < 0 || n_bloch_max > 4 || !reci_uc || !recivec_pc || n_pc <= 0 ||
        n_pc > KPC_MAX_N || nent < 0 || !el_off || (nent > 0 && (!coef || !dr)))
        return NT_ERR_ARG;
    if (nmodes == 0) return NT_OK;
    if (!E || !k_xyz || !kp || !v || !ncand) return NT_ERR_ARG;
    return guarded(c, [&] {
        // host check of the element table: offsets monotone from 0 to nent
        const int ne = n_pc * n_pc;
        std::vector<int> off(ne + 1);
        hipcheck(c, hipMemcpy(off.data(), el_off, sizeof(int) * (ne + 1), hipMemcpyDeviceToHost), "D2H");
        if (off[0] != 0 || off[ne] != nent) {
            c->err = "unfold_pc_tb: element offsets must run from 0 to nent";
            return (int)NT_ERR_ARG;
        }
        for (int e = 0; e < ne; ++e)
            if (off[e + 1] < off[e]) {
                c->err = "unfold_pc_tb: element offsets must not decrease";
                return (int)NT_ERR_ARG;
            }
        reg_region(c, E, sizeof(double) * nmodes);
        reg_region(c, k_xyz, sizeof(double) * 3 * (size_t)nmodes);
        reg_region(c, kp, sizeof(double) * 3 * (size_t)nmodes);
        reg_region(c, v, sizeof(double) * 3 * (size_t)nmodes);
        reg_region(c, ncand, sizeof(int) * (size_t)nmodes);
        reg_region(c, el_off, sizeof(int) * (ne + 1));
        reg_region(c, coef, sizeof(double) * (size_t)nent);
        reg_region(c, dr, sizeof(double) * 3 * (size_t)nent);
        check_range(c, el_off, sizeof(int) * (ne + 1), "unfold_pc_tb offsets");
        if (nent > 0) {
            check_range(c, coef, sizeof(double) * (size_t)nent, "unfold_pc_tb coef");
            check_range(c, dr, sizeof(double) * 3 * (size_t)nent, "unfold_pc_tb dr");
        }
        KpcArgs a = kpc_args(1, n_pc, E, k_xyz, n_bloch_max, reci_uc, recivec_pc, kp, v, ncand);
        return unfold_pc_run(c, a, nmodes, k_xyz,
                             [&](int c0, int ncs, const int* dcm, const int* dcs, double*, double* dkc, double* dD) {
                                 launch(c, NT_K_ELEM, 0.0, "kpc tight-binding matrices", [&] {
                                     return launch_kpc_tb_mats(a, dcm, dcs, c0, ncs, el_off, coef, dr, dD, dkc,
                                                               c->stream);
                                 });
                             });
    });
}

int nt_zgeev_batched_dev(nt_ctx* c, int n, const nt_c128* A, long long strideA, nt_c128* w, nt_c128* V,
                         long long strideV, int* status, int batch)
{
    if (!c || n <= 0 || batch < 0 || !A || !w || !V || strideA < (long long)n * n || strideV < (long long)n * n)
        return NT_ERR_ARG;
    if (batch == 0) return NT_OK;
    return guarded(c, [&] {
        const size_t z = sizeof(nt_c128);
        reg_region(c, A, z * ((size_t)strideA * (batch - 1) + (size_t)n * n));
        reg_region(c, V, z * ((size_t)strideV * (batch - 1) + (size_t)n * n));
        reg_region(c, w, z * (size_t)n * batch);
        reg_region(c, status, sizeof(int) * batch);
        ensure_ws(c, eig_scratch(n, batch));
        Arena ar;
        ar.base = static_cast<char*>(c->ws);
        eig_run(c, n, bm<double>(reinterpret_cast<const double*>(A), strideA, n), reinterpret_cast<double*>(w),
                bw(reinterpret_cast<double*>(V), strideV, n), status, batch, ar, 0);
        return 0;
    });
}

int nt_zsqrtm_batched_dev(nt_ctx* c, int n, const nt_c128* A, long long strideA, nt_c128* S, long long strideS,
                          int* status, int batch)
{
    if (!c || n <= 0 || batch < 0 || !A || !S || strideA < (long long)n * n || strideS < (long long)n * n)
        return NT_ERR_ARG;
    if (batch == 0) return NT_OK;
    return guarded(c, [&] {
        const size_t z = sizeof(nt_c128);
        reg_region(c, A, z * ((size_t)strideA * (batch - 1) + (size_t)n * n));
        reg_region(c, S, z * ((size_t)strideS * (batch - 1) + (size_t)n * n));
        reg_region(c, status, sizeof(int) * batch);
        ensure_ws(c, eig_scratch(n, batch));
        Arena ar;
        ar.base = static_cast<char*>(c->ws);
        sqrtm_run(c, n, bm<double>(reinterpret_cast<const double*>(A), strideA, n),
                  bw(reinterpret_cast<double*>(S), strideS, n), status, batch, ar, 0);
        return 0;
    });
}

}  // extern "C"
