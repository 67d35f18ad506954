// nt_blocks.hip -- C-ABI entries that wrap one kernel family each: the H(k)
// assembly of the QE force-constant tables and of the tight-binding
// Slater-Koster tables (kernels in nt_hk.hip), and the batched complex GEMM
// and inverse building blocks (nt_gemm.hip, nt_inv.hip).
#include <algorithm>
#include <vector>

#include "nt_host.h"

extern "C" {

int nt_hk_assemble_dev(nt_ctx* c, int nk, const double* kcart, int nseg, const int* seg_off, const int* seg_rc,
                       int nent, const double* coef, const double* dr, int rows, int cols, nt_c128* out, int ld,
                       long long kstride)
{
    if (!c || nk < 0 || nseg < 0 || nent < 0 || rows <= 0 || cols <= 0 || rows % 3 || cols % 3 || ld < rows ||
        !out || (nk > 1 && kstride < (long long)ld * cols) || (nseg > 0 && (!seg_off || !seg_rc || !coef || !dr)) ||
        (nk > 0 && !kcart))
        return NT_ERR_ARG;
    if (nk == 0) return NT_OK;
    return guarded(c, [&] {
        // host check of the segment table: offsets monotone within [0, nent],
        // atom indices inside the block
        if (nseg > 0) {
            std::vector<int> off(nseg + 1), rc(2 * (size_t)nseg);
            hipcheck(c, hipMemcpy(off.data(), seg_off, sizeof(int) * (nseg + 1), hipMemcpyDeviceToHost), "D2H");
            hipcheck(c, hipMemcpy(rc.data(), seg_rc, sizeof(int) * 2 * nseg, hipMemcpyDeviceToHost), "D2H");
            if (off[0] != 0 || off[nseg] != nent) {
                c->err = "hk: segment offsets must run from 0 to nent";
                return (int)NT_ERR_ARG;
            }
            for (int q = 0; q < nseg; ++q)
                if (off[q + 1] < off[q] || rc[2 * q] < 0 || 3 * rc[2 * q] + 3 > rows || rc[2 * q + 1] < 0 ||
                    3 * rc[2 * q + 1] + 3 > cols) {
                    c->err = "hk: segment outside the block";
                    return (int)NT_ERR_ARG;
                }
        }
        const size_t z = sizeof(nt_c128);
        reg_region(c, kcart, sizeof(double) * 3 * nk);
        reg_region(c, seg_off, sizeof(int) * (nseg + 1));
        reg_region(c, seg_rc, sizeof(int) * 2 * nseg);
        reg_region(c, coef, sizeof(double) * 9 * (size_t)nent);
        reg_region(c, dr, sizeof(double) * 3 * (size_t)nent);
        const size_t span = z * ((size_t)(nk - 1) * kstride + (size_t)ld * (cols - 1) + rows);
        reg_region(c, out, span);
        check_range(c, kcart, sizeof(double) * 3 * nk, "hk k");
        launch(c, NT_K_ELEM, 0.0, "hk assemble", [&] {
            return launch_hk_assemble(nk, kcart, nseg, seg_off, seg_rc, coef, dr, rows, cols,
                                      reinterpret_cast<double*>(out), ld, (long)kstride, c->stream);
        });
        return 0;
    });
}

// Interval partition of one axis from the segments' (start, length) pairs:
// segments with the same start must have the same length and intervals of
// different starts must not overlap.  Fills blk / loc (per row or column, -1
// outside every interval) and seg_blk (each segment's interval index);
// returns the number of intervals or -1.
static int hk_axis(const std::vector<int>& g, int nseg, int off, int n, std::vector<int>& blk, std::vector<int>& loc,
                   std::vector<int>& seg_blk)
{
    std::vector<std::pair<int, int>> iv;                 // (start, length)
    for (int q = 0; q < nseg; ++q) iv.push_back({g[4 * q + off], g[4 * q + off + 2]});
    std::sort(iv.begin(), iv.end());
    iv.erase(std::unique(iv.begin(), iv.end()), iv.end());
    for (size_t t = 1; t < iv.size(); ++t)
        if (iv[t].first < iv[t - 1].first + iv[t - 1].second) return -1;     // overlap or two lengths
    blk.assign(n, -1);
    loc.assign(n, 0);
    for (size_t t = 0; t < iv.size(); ++t)
        for (int x = 0; x < iv[t].second; ++x) {
            blk[iv[t].first + x] = (int)t;
            loc[iv[t].first + x] = x;
        }
    seg_blk.resize(nseg);
    for (int q = 0; q < nseg; ++q) seg_blk[q] = blk[g[4 * q + off]];
    return (int)iv.size();
}

int nt_hk_assemble_orb_dev(nt_ctx* c, int nk, const double* kcart, int nseg, const int* seg_off,
                           const int* seg_geom, int nent, int cstride, const double* coef, const double* dr, int rows,
                           int cols, nt_c128* out, int ld, long long kstride, int hermitize)
{
    if (!c || nk < 0 || nseg < 0 || nent < 0 || cstride <= 0 || rows <= 0 || cols <= 0 || ld < rows || !out ||
        (hermitize && rows != cols) || (nk > 1 && kstride < (long long)ld * cols) ||
        (nseg > 0 && (!seg_off || !seg_geom || !coef || !dr)) || (nk > 0 && !kcart))
        return NT_ERR_ARG;
    if (nk == 0) return NT_OK;
    return guarded(c, [&] {
        // host check of the segment table (offsets monotone within [0, nent],
        // every sub-block inside the block and no larger than cstride, the
        // sub-blocks on a grid with no two in one cell) and the element ->
        // segment maps the kernel reads
        std::vector<int> off(nseg + 1, 0), g(4 * (size_t)nseg);
        if (nseg > 0) {
            hipcheck(c, hipMemcpy(off.data(), seg_off, sizeof(int) * (nseg + 1), hipMemcpyDeviceToHost), "D2H");
            hipcheck(c, hipMemcpy(g.data(), seg_geom, sizeof(int) * 4 * nseg, hipMemcpyDeviceToHost), "D2H");
        }
        if (off[0] != 0 || off[nseg] != nent) {
            c->err = "hk: segment offsets must run from 0 to nent";
            return (int)NT_ERR_ARG;
        }
        for (int q = 0; q < nseg; ++q) {
            const int r0 = g[4 * q], c0 = g[4 * q + 1], ni = g[4 * q + 2], nj = g[4 * q + 3];
            if (off[q + 1] < off[q] || r0 < 0 || c0 < 0 || ni < 1 || nj < 1 || (long)ni * nj > cstride ||
                r0 > rows - ni || c0 > cols - nj) {
                c->err = "hk: segment outside the block";
                return (int)NT_ERR_ARG;
            }
        }
        std::vector<int> rblk, rloc, cblk, cloc, srb, scb;
        const int nrb = hk_axis(g, nseg, 0, rows, rblk, rloc, srb);
        const int ncb = hk_axis(g, nseg, 1, cols, cblk, cloc, scb);
        if (nrb < 0 || ncb < 0) {
            c->err = "hk: segments must tile the block on a grid of row and column intervals";
            return (int)NT_ERR_ARG;
        }
        std::vector<int> order(nseg);
        for (int q = 0; q < nseg; ++q) order[q] = q;
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            return srb[a] != srb[b] ? srb[a] < srb[b] : scb[a] < scb[b];
        });
        std::vector<int> rptr(nrb + 1, 0), ckey(nseg), sid(nseg);
        for (int t = 0; t < nseg; ++t) {
            const int q = order[t];
            if (t > 0 && srb[q] == srb[order[t - 1]] && scb[q] == scb[order[t - 1]]) {
                c->err = "hk: two segments in one sub-block";
                return (int)NT_ERR_ARG;
            }
            ckey[t] = scb[q];
            sid[t] = q;
            ++rptr[srb[q] + 1];
        }
        for (int b = 0; b < nrb; ++b) rptr[b + 1] += rptr[b];
        // maps packed into the workspace: rblk rloc cblk cloc rptr ckey sid
        std::vector<int> pack;
        pack.reserve(2 * (size_t)rows + 2 * (size_t)cols + nrb + 1 + 2 * (size_t)nseg + 1);
        size_t o_rblk = pack.size();
        pack.insert(pack.end(), rblk.begin(), rblk.end());
        size_t o_rloc = pack.size();
        pack.insert(pack.end(), rloc.begin(), rloc.end());
        size_t o_cblk = pack.size();
        pack.insert(pack.end(), cblk.begin(), cblk.end());
        size_t o_cloc = pack.size();
        pack.insert(pack.end(), cloc.begin(), cloc.end());
        size_t o_rptr = pack.size();
        pack.insert(pack.end(), rptr.begin(), rptr.end());
        size_t o_ckey = pack.size();
        pack.insert(pack.end(), ckey.begin(), ckey.end());
        size_t o_sid = pack.size();
        pack.insert(pack.end(), sid.begin(), sid.end());
        pack.push_back(0);                                  // keeps the last array non-empty
        ensure_ws(c, sizeof(int) * pack.size());
        int* w = static_cast<int*>(c->ws);
        hipcheck(c, hipMemcpyAsync(w, pack.data(), sizeof(int) * pack.size(), hipMemcpyHostToDevice, c->stream),
                 "H2D");
        const HkGrid hg{w + o_rblk, w + o_rloc, w + o_cblk, w + o_cloc, w + o_rptr, w + o_ckey, w + o_sid};
        const size_t z = sizeof(nt_c128);
        reg_region(c, kcart, sizeof(double) * 3 * nk);
        reg_region(c, seg_off, sizeof(int) * (nseg + 1));
        reg_region(c, seg_geom, sizeof(int) * 4 * nseg);
        reg_region(c, coef, sizeof(double) * (size_t)cstride * nent);
        reg_region(c, dr, sizeof(double) * 3 * (size_t)nent);
        const size_t span = z * ((size_t)(nk - 1) * kstride + (size_t)ld * (cols - 1) + rows);
        reg_region(c, out, span);
        check_range(c, kcart, sizeof(double) * 3 * nk, "hk k");
        launch(c, NT_K_ELEM, 0.0, "hk assemble orb", [&] {
            return launch_hk_assemble_orb(nk, kcart, seg_off, seg_geom, cstride, coef, dr, hg, rows, cols,
                                          reinterpret_cast<double*>(out), ld, (long)kstride, hermitize, c->stream);
        });
        // the maps were copied from pageable host memory that ends with this call
        hipcheck(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        return 0;
    });
}

int nt_zgemm_batched_dev(nt_ctx* c, char opA, char opB, int m, int n, int k, nt_c128 alpha, const nt_c128* A,
                         int lda, long long strideA, const nt_c128* B, int ldb, long long strideB, nt_c128 beta,
                         nt_c128* C, int ldc, long long strideC, int batch)
{
    if (!c) return NT_ERR_ARG;
    auto opv = [](char o) { return o == 'N' || o == 'n' ? OP_N : (o == 'C' || o == 'c' ? OP_C : -1); };
    const int oa = opv(opA), ob = opv(opB);
    if (oa < 0 || ob < 0 || m < 0 || n < 0 || k < 0 || batch < 0 || !C || (k > 0 && (!A || !B))) return NT_ERR_ARG;
    if (ldc < std::max(1, m) || lda < 1 || ldb < 1) return NT_ERR_ARG;
    return guarded(c, [&] {
        typedef double T;
        // the caller vouches for the ranges its arguments describe
        auto span = [](long long stride, int ld, int rows, int cols, int nb) {
            return (size_t)((nb - 1) * std::max(0LL, stride) + (long long)(cols - 1) * ld + rows) * sizeof(nt_c128);
        };
        if (k > 0) {
            reg_region(c, A, span(strideA, lda, oa == OP_N ? m : k, oa == OP_N ? k : m, batch));
            reg_region(c, B, span(strideB, ldb, ob == OP_N ? k : n, ob == OP_N ? n : k, batch));
        }
        if (m > 0 && n > 0) reg_region(c, C, span(strideC, ldc, m, n, batch));
        GemmArgs<T> a;
        a.M = m;
        a.N = n;
        a.K = k;
        a.opA = oa;
        a.opB = ob;
        a.A = bm<T>(reinterpret_cast<const T*>(A), strideA, lda);
        a.B = bm<T>(reinterpret_cast<const T*>(B), strideB, ldb);
        a.C = bw<T>(reinterpret_cast<T*>(C), strideC, ldc);
        a.Cin = (beta.re != 0.0 || beta.im != 0.0) ? bm<T>(reinterpret_cast<const T*>(C), strideC, ldc) : none<T>();
        a.C2 = nonew<T>();
        a.alpha_re = alpha.re;
        a.alpha_im = alpha.im;
        a.beta_re = beta.re;
        a.beta_im = beta.im;
        a.diag = nullptr;
        a.active = nullptr;
        a.maxabs = nullptr;
        a.batch = batch;
        a.groups = 1;
        a.m3 = c->gemm_m3 ? 1 : 0;
        launch(c, NT_K_GEMM, 8.0 * m * n * k * (double)batch, "gemm", [&] { return launch_gemm<T>(a, c->stream); });
        return 0;
    });
}

int nt_zinv_batched_dev(nt_ctx* c, int n, nt_c128* A, long long strideA, nt_c128* X, long long strideX, int* info,
                        int batch)
{
    if (!c) return NT_ERR_ARG;
    if (n <= 0 || n > inv_max_n() || batch < 0 || !A || !X || A == X) return NT_ERR_ARG;
    return guarded(c, [&] {
        typedef double T;
        const size_t span = (size_t)((batch - 1) * std::max(0LL, strideA) + (long long)n * n) * sizeof(nt_c128);
        const size_t spanx = (size_t)((batch - 1) * std::max(0LL, strideX) + (long long)n * n) * sizeof(nt_c128);
        reg_region(c, A, span);
        reg_region(c, X, spanx);
        reg_region(c, info, sizeof(int) * batch);
        inverse<T>(c, n, bw<T>(reinterpret_cast<T*>(A), strideA, n), none<T>(), nullptr,
                   bw<T>(reinterpret_cast<T*>(X), strideX, n), info, batch);
        return 0;
    });
}

}  // extern "C"
