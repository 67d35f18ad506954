// Worker contexts: a context's own contexts on the same device, each with its
// own stream, workspace and host thread.  They run the parts of a batched or
// adjoint solve of many points side by side (nt_set_point_split) and the
// block-size groups of nt_solve_tdb_multi_dev (nt_set_multi_workers).
#include <algorithm>
#include <atomic>
#include <functional>
#include <mutex>
#include <thread>
#include "nt_host.h"
#include "nt_kernels.h"
#include "nt_solve.h"

// out[i] += in[i] (n doubles): the gradient of one part of a split adjoint solve
__global__ void k_add_doubles(double* __restrict__ out, const double* __restrict__ in, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += in[i];
}

// Worker contexts of c (created once, kept until nt_destroy) given c's
// settings and each a workspace cap of its share of the device memory left
// (its own workspace from earlier calls counts as available to it).
static int prepare_workers(nt_ctx* c, int W)
{
    while ((int)c->subs.size() < W) {
        nt_ctx* s = nullptr;
        const int st = nt_create(&s, c->device);
        if (st) return st;
        c->subs.push_back(s);
    }
    size_t free_b = 0, total_b = 0;
    if (hipSetDevice(c->device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess) return NT_ERR_HIP;
    size_t avail = free_b;
    for (int w = 0; w < W; ++w) avail += c->subs[w]->ws_bytes;
    const size_t share = std::max<size_t>(size_t(2) << 30, (size_t)(0.6 * (double)avail / W));
    for (int w = 0; w < W; ++w) {
        nt_ctx* s = c->subs[w];
        if (s->ws && s->ws_bytes > share + share / 8) {
            // a workspace grown by an earlier call beyond this call's share
            // (and the eighth ensure_ws adds) would leave the other workers
            // less than theirs: release it
            if (hipSetDevice(s->device) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess ||
                hipFree(s->ws) != hipSuccess)
                return NT_ERR_HIP;
            s->ws = nullptr;
            s->ws_bytes = 0;
        }
        s->ws_cap = share;
        s->overlap = c->overlap;
        s->gemm_m3 = c->gemm_m3;
        s->batch_limit = c->batch_limit;
        s->one_sweep = c->one_sweep;
        s->dec_check = c->dec_check;
        s->adj_ckpt = c->adj_ckpt;
        s->inv_nb = c->inv_nb;
        s->gemm_tile = c->gemm_tile;
        s->prof = c->prof;
        s->ws_poison = c->ws_poison;
        s->err.clear();
    }
    return NT_OK;
}

// Solve the jobs on W prepared worker contexts, each on its own stream with
// its own workspace from its own host thread, each worker taking the next
// job when it is done.  The workers start after c's stream has reached this
// call and c's stream waits for all of them; their profile counts, mixed
// counters and mean decimation iterations go to c.  Runs inside guarded().
static void run_on_workers(nt_ctx* c, const nt_params* p, int W, const std::vector<const TdbArgs*>& jobs)
{
    const int nj = (int)jobs.size();
    hipcheck(c, hipEventRecord(c->fork, c->stream), "event fork");
    for (int w = 0; w < W; ++w) hipcheck(c, hipStreamWaitEvent(c->subs[w]->stream, c->fork, 0), "wait fork");
    std::atomic<int> next{0};
    std::mutex mu;
    int fail = 0;
    std::string fail_err;
    std::vector<double> iters_w(W, 0.0), pts_w(W, 0.0);
    auto work = [&](int w) {
        nt_ctx* s = c->subs[w];
        try {
            hipcheck(s, hipSetDevice(s->device), "hipSetDevice");
            s->regions = c->regions;                       // the caller's buffers
            if (s->ws) reg_region(s, s->ws, s->ws_bytes);
            for (;;) {
                {
                    std::lock_guard<std::mutex> lk(mu);
                    if (fail) break;
                }
                const int j = next.fetch_add(1);
                if (j >= nj) break;
                solve_tdb_dev(s, p, *jobs[j]);
                if (s->last_mean_iters >= 0.0) {
                    iters_w[w] += s->last_mean_iters * jobs[j]->npts;
                    pts_w[w] += jobs[j]->npts;
                }
            }
            if (s->prof) harvest_profile(s);
        } catch (const Fail& e) {
            std::lock_guard<std::mutex> lk(mu);
            if (!fail) {
                fail = e.code;
                fail_err = s->err;
            }
        } catch (const std::bad_alloc&) {
            std::lock_guard<std::mutex> lk(mu);
            if (!fail) {
                fail = NT_ERR_NOMEM;
                fail_err = "host allocation failed";
            }
        } catch (...) {
            // nothing may leave a worker thread
            std::lock_guard<std::mutex> lk(mu);
            if (!fail) {
                fail = NT_ERR_HIP;
                fail_err = "unexpected error in a worker context";
            }
        }
    };
    std::vector<std::thread> threads;
    threads.reserve(W);
    for (int w = 0; w < W; ++w) threads.emplace_back(work, w);
    for (auto& t : threads) t.join();
    double it_sum = 0.0, pt_sum = 0.0;
    for (int w = 0; w < W; ++w) {
        nt_ctx* s = c->subs[w];
        hipcheck(c, hipEventRecord(s->join, s->stream), "event join");
        hipcheck(c, hipStreamWaitEvent(c->stream, s->join, 0), "wait join");
        for (int k = 0; k < NT_K_COUNT; ++k) {
            c->launches[k] += s->launches[k];
            c->ms[k] += s->ms[k];
            c->flops[k] += s->flops[k];
            s->launches[k] = 0;
            s->ms[k] = 0;
            s->flops[k] = 0;
        }
        for (int k = 0; k < 3; ++k) {
            c->mixed_counts[k] += s->mixed_counts[k];
            s->mixed_counts[k] = 0;
        }
        it_sum += iters_w[w];
        pt_sum += pts_w[w];
    }
    c->last_mean_iters = pt_sum > 0.0 ? it_sum / pt_sum : -1.0;
    if (fail) {
        c->err = fail_err;
        throw Fail{fail};
    }
}

// the part [p0, p0 + n) of a batched solve's points
static TdbArgs point_part(const TdbArgs& a, int p0, int n)
{
    auto o = [&](auto* ptr, long per) { return ptr ? ptr + (long)p0 * per : ptr; };
    const long nll = (long)a.nl * a.nl, nrr = (long)a.nr * a.nr, nlr = (long)a.nl * a.nr;
    TdbArgs t = a;
    t.npts = n;
    t.E = a.E + p0;
    t.kidx = o(a.kidx, 1);
    t.first_mask = o(a.first_mask, 1);
    t.T_LR = a.T_LR + p0;
    t.T_RL = a.T_RL + p0;
    t.status = o(a.status, 1);
    t.iters = o(a.iters, 1);
    nt_solve_extras& x = t.ex;
    x.gl_adv = o(x.gl_adv, nll);
    x.gr_ret = o(x.gr_ret, nrr);
    x.gl_plus = o(x.gl_plus, nll);
    x.gr_minus = o(x.gr_minus, nrr);
    x.gml_adv = o(x.gml_adv, nll);
    x.gmr_ret = o(x.gmr_ret, nrr);
    x.sl_adv = o(x.sl_adv, nll);
    x.sr_ret = o(x.sr_ret, nrr);
    x.G_D_LR = o(x.G_D_LR, nlr);
    x.G_D_RL = o(x.G_D_RL, nlr);
    x.G_L = o(x.G_L, nll);
    x.G_R = o(x.G_R, nrr);
    x.sdos_l = o(x.sdos_l, 1);
    x.sdos_r = o(x.sdos_r, 1);
    x.ldos = o(x.ldos, (long)a.L * a.b);
    x.ddos = o(x.ddos, 1);
    return t;
}

// Part q takes the points q, q + parts, ... (gathered in c's
// workspace by strided copies, the results scattered back): neighbouring
// energies need similar numbers of decimation steps, so interleaved parts
// take equal times; the extra outputs are gathered and scattered the same
// way.  An adjoint solve accumulates part 0's gradients into the
// caller's and the other parts' into zeroed workspace, summed into the
// caller's after the join.
void solve_split(nt_ctx* c, const nt_params* p, const TdbArgs& a, int parts)
{
    if (const int st = prepare_workers(c, parts)) {
        c->err = "could not set up the worker contexts of a split solve";
        throw Fail{st};
    }
    const int npts = a.npts;
    std::vector<TdbArgs> part(parts);
    std::vector<const TdbArgs*> jobs(parts);
    std::vector<int> cnt(parts), off(parts);
    for (int q = 0, p0 = 0; q < parts; ++q) {
        cnt[q] = npts / parts + (q < npts % parts ? 1 : 0);
        off[q] = p0;
        p0 += cnt[q];
    }
    const nt_solve_extras& x = a.ex;
    const bool extras_any = x.gl_adv || x.gr_ret || x.gl_plus || x.gr_minus || x.gml_adv || x.gmr_ret || x.sl_adv ||
                            x.sr_ret || x.G_D_LR || x.G_D_RL || x.G_L || x.G_R || x.sdos_l || x.sdos_r || x.ldos ||
                            x.ddos;
    const bool adj = a.adj_w != nullptr;
    if (a.probe || !a.siblings.empty()) {
        c->err = "a split solve takes no probes or sibling devices";
        throw Fail{NT_ERR_ARG};
    }
    if (extras_any && adj) {
        c->err = "a split adjoint solve takes no extras";
        throw Fail{NT_ERR_ARG};
    }
    const size_t nd = sizeof(double) * npts, ni = sizeof(int) * npts;
    const size_t z = sizeof(nt_c128), bb = (size_t)a.b * a.b;
    const int nc = std::min(a.adj_hi, a.L - 2) - a.adj_lo + 1;
    const size_t gd = adj ? z * bb * (a.adj_hi - a.adj_lo + 1) * a.nk : 0;
    const size_t gv = adj && nc > 0 && a.dHdv ? z * bb * nc * a.nk : 0;
    Arena pa;
    const size_t oE = pa.take(nd), oW = pa.take(adj ? nd : 0), oTl = pa.take(nd), oTr = pa.take(nd),
                 oK = pa.take(ni), oM = pa.take(ni), oS = pa.take(ni), oI = pa.take(ni),
                 oG = pa.take(gd * (parts - 1)), oGv = pa.take(gv * (parts - 1));
    // the extra outputs, gathered the same way: (the gathered copy's pointer
    // in g, the caller's buffer, bytes per point, workspace offset)
    TdbArgs g = a;
    struct ExOut {
        void* gath;
        std::function<void(char*)> set;
        const void* user;
        size_t w, off;
    };
    std::vector<ExOut> exo;
    auto ex = [&](auto*& gp, size_t w) {
        if (!gp) return;
        exo.push_back({nullptr, [&gp](char* b) { gp = reinterpret_cast<std::remove_reference_t<decltype(gp)>>(b); },
                       gp, w, pa.take(w * npts)});
    };
    {
        const size_t nll = (size_t)a.nl * a.nl, nrr = (size_t)a.nr * a.nr, nlr = (size_t)a.nl * a.nr;
        nt_solve_extras& e = g.ex;
        ex(e.gl_adv, z * nll), ex(e.gr_ret, z * nrr), ex(e.gl_plus, z * nll), ex(e.gr_minus, z * nrr);
        ex(e.gml_adv, z * nll), ex(e.gmr_ret, z * nrr), ex(e.sl_adv, z * nll), ex(e.sr_ret, z * nrr);
        ex(e.G_D_LR, z * nlr), ex(e.G_D_RL, z * nlr), ex(e.G_L, z * nll), ex(e.G_R, z * nrr);
        ex(e.sdos_l, sizeof(double)), ex(e.sdos_r, sizeof(double));
        ex(e.ldos, sizeof(double) * a.L * a.b), ex(e.ddos, sizeof(double));
    }
    ensure_ws(c, pa.off);
    pa.base = static_cast<char*>(c->ws);
    for (ExOut& o : exo) {
        o.gath = pa.ptr<char>(o.off);
        o.set(pa.ptr<char>(o.off));
    }
    g.E = pa.ptr<double>(oE);
    g.adj_w = adj ? pa.ptr<double>(oW) : nullptr;
    g.T_LR = pa.ptr<double>(oTl);
    g.T_RL = pa.ptr<double>(oTr);
    g.kidx = a.kidx ? pa.ptr<int>(oK) : nullptr;
    g.first_mask = a.first_mask ? pa.ptr<int>(oM) : nullptr;
    g.status = a.status ? pa.ptr<int>(oS) : nullptr;
    g.iters = a.iters ? pa.ptr<int>(oI) : nullptr;
    // element q + parts i of the caller's array <-> element off[q] + i of the gathered one
    auto move = [&](const void* gath, const void* user, size_t w, bool in) {
        for (int q = 0; q < parts; ++q) {
            char* gq = static_cast<char*>(const_cast<void*>(gath)) + w * off[q];
            char* uq = static_cast<char*>(const_cast<void*>(user)) + w * q;
            hipcheck(c, in ? hipMemcpy2DAsync(gq, w, uq, w * parts, w, cnt[q], hipMemcpyDeviceToDevice, c->stream)
                           : hipMemcpy2DAsync(uq, w * parts, gq, w, w, cnt[q], hipMemcpyDeviceToDevice, c->stream),
                     "point gather");
        }
    };
    move(g.E, a.E, sizeof(double), true);
    if (adj) move(g.adj_w, a.adj_w, sizeof(double), true);
    if (a.kidx) move(g.kidx, a.kidx, sizeof(int), true);
    if (a.first_mask) move(g.first_mask, a.first_mask, sizeof(int), true);
    if (gd) zero_dev(c, pa.ptr<char>(oG), gd * (parts - 1), c->stream);
    if (gv) zero_dev(c, pa.ptr<char>(oGv), gv * (parts - 1), c->stream);
    for (int q = 0; q < parts; ++q) {
        part[q] = point_part(g, off[q], cnt[q]);
        if (adj) {
            part[q].adj_w = g.adj_w + off[q];
            if (q > 0) {
                part[q].dHd = reinterpret_cast<nt_c128*>(pa.ptr<char>(oG) + gd * (q - 1));
                part[q].dHdv = gv ? reinterpret_cast<nt_c128*>(pa.ptr<char>(oGv) + gv * (q - 1)) : nullptr;
            }
        }
        jobs[q] = &part[q];
    }
    run_on_workers(c, p, parts, jobs);
    move(g.T_LR, a.T_LR, sizeof(double), false);
    move(g.T_RL, a.T_RL, sizeof(double), false);
    if (a.status) move(g.status, a.status, sizeof(int), false);
    if (a.iters) move(g.iters, a.iters, sizeof(int), false);
    for (const ExOut& o : exo) move(o.gath, o.user, o.w, false);
    for (int q = 1; q < parts && adj; ++q) {
        const long n1 = (long)(gd / sizeof(double)), n2 = (long)(gv / sizeof(double));
        hipLaunchKernelGGL(k_add_doubles, dim3(nblk(n1)), dim3(256), 0, c->stream, reinterpret_cast<double*>(a.dHd),
                           pa.ptr<double>(oG + gd * (q - 1)), n1);
        if (gv)
            hipLaunchKernelGGL(k_add_doubles, dim3(nblk(n2)), dim3(256), 0, c->stream,
                               reinterpret_cast<double*>(a.dHdv), pa.ptr<double>(oGv + gv * (q - 1)), n2);
        hipcheck(c, hipGetLastError(), "gradient sum");
    }
}

int nt_solve_tdb_multi_dev(nt_ctx* c, const nt_params* p, int ndev, const nt_device_desc* devs)
{
    if (!c || ndev < 0 || (ndev > 0 && !devs) || check_params(p)) return NT_ERR_ARG;
    for (int d = 0; d < ndev; ++d) {
        const nt_device_desc& D = devs[d];
        if (check_circular(p, D.nl, D.nr)) return NT_ERR_ARG;
        if (D.nl <= 0 || D.nr <= 0 || D.b <= 0 || D.L <= 0 || D.npts < 0 || D.nk <= 0 || (D.nk > 1 && !D.kidx))
            return NT_ERR_ARG;
        if (std::max(std::max(D.nl, D.nr), D.b) > inv_max_n()) return NT_ERR_ARG;
        if (D.npts > 0 && (!D.E || !D.Haml || !D.Hamlv || !D.Hamr || !D.Hamrv || !D.Hamd || (D.L > 1 && !D.Hamdv) ||
                           !D.Hamld || !D.Hamrd || !D.T_LR || !D.T_RL))
            return NT_ERR_ARG;
    }
    // largest first: b^3 (L + 2 N_dec) per point, N_dec ~ 15
    std::vector<int> order(ndev);
    for (int d = 0; d < ndev; ++d) order[d] = d;
    auto cost = [&](int d) {
        const double b = devs[d].b;
        return b * b * b * (devs[d].L + 30.0) * devs[d].npts;
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
    // Devices with the same lead buffers, lead sizes and points (E and k
    // map) form a group whose lead Green's functions are computed once (FP64
    // solves; a mixed-precision device is a group of its own).
    auto same_leads = [&](const nt_device_desc& x, const nt_device_desc& y) {
        return x.nl == y.nl && x.nr == y.nr && x.nk == y.nk && x.npts == y.npts && x.Haml == y.Haml &&
               x.Hamlv == y.Hamlv && x.Hamr == y.Hamr && x.Hamrv == y.Hamrv && x.E == y.E && x.kidx == y.kidx;
    };
    std::vector<std::vector<int>> groups;
    for (int d : order) {
        if (devs[d].npts == 0) continue;
        bool placed = false;
        if (p->precision == NT_PREC_FP64)
            for (auto& g : groups)
                if (same_leads(devs[g[0]], devs[d])) {
                    g.push_back(d);
                    placed = true;
                    break;
                }
        if (!placed) groups.push_back({d});
    }
    std::vector<std::pair<int, int>> tuned;   // (bytes, n) already tuned this call
    const int bytes = p->precision == NT_PREC_FP64 ? 8 : 4;
    for (const auto& g : groups)
        for (int d : g) {
            const nt_device_desc& D = devs[d];
            for (int n : {D.b, std::max(D.nl, D.nr)}) {
                if (c->inv_nb.count({bytes, n}) ||
                    std::find(tuned.begin(), tuned.end(), std::make_pair(bytes, n)) != tuned.end())
                    continue;
                const int tb = std::min(D.npts, c->batch_limit);
                const int st = nt_autotune_inverse(c, n, std::max(1, std::min(tb, 256)), p->precision, nullptr,
                                                   nullptr);
                if (st) return st;
                tuned.push_back({bytes, n});
            }
        }
    const int ng = (int)groups.size();
    const int W = std::min(std::max(1, c->multi_workers), ng);
    if (W > 1) {
        if (const int st = prepare_workers(c, W)) return st;
    }
    return guarded(c, [&] {
        const size_t z = sizeof(nt_c128);
        std::vector<std::vector<TdbArgs>> args(ng);
        for (int gi = 0; gi < ng; ++gi) {
            auto& a = args[gi];
            a.reserve(groups[gi].size());
            for (int d : groups[gi]) {
                const nt_device_desc& D = devs[d];
                const size_t nll = (size_t)D.nl * D.nl, nrr = (size_t)D.nr * D.nr, bb = (size_t)D.b * D.b;
                reg_region(c, D.E, sizeof(double) * D.npts);
                reg_region(c, D.kidx, sizeof(int) * D.npts);
                reg_region(c, D.Haml, z * nll * D.nk);
                reg_region(c, D.Hamlv, z * nll * D.nk);
                reg_region(c, D.Hamr, z * nrr * D.nk);
                reg_region(c, D.Hamrv, z * nrr * D.nk);
                reg_region(c, D.Hamd, z * bb * D.L * D.nk);
                reg_region(c, D.Hamdv, z * bb * std::max(D.L - 1, 1) * D.nk);
                reg_region(c, D.Hamld, z * D.nl * D.b * D.nk);
                reg_region(c, D.Hamrd, z * D.nr * D.b * D.nk);
                reg_region(c, D.first_point_mask, sizeof(int) * D.npts);
                reg_region(c, D.T_LR, sizeof(double) * D.npts);
                reg_region(c, D.T_RL, sizeof(double) * D.npts);
                reg_region(c, D.status, sizeof(int) * D.npts);
                a.push_back(TdbArgs{D.nl, D.nr, D.b, D.L, D.npts, D.nk, D.E, D.kidx, D.Haml, D.Hamlv, D.Hamr,
                                    D.Hamrv, D.Hamd, D.Hamdv, D.Hamld, D.Hamrd, D.first_point_mask, D.T_LR, D.T_RL,
                                    D.status, nullptr, nt_solve_extras{}});
            }
            for (size_t i = 1; i < a.size(); ++i) a[0].siblings.push_back(&a[i]);
            // every k index must name one of the group's nk block sets
            const nt_device_desc& D0 = devs[groups[gi][0]];
            if (D0.kidx) {
                std::vector<int> hk(D0.npts);
                hipcheck(c, hipMemcpyAsync(hk.data(), D0.kidx, sizeof(int) * D0.npts, hipMemcpyDeviceToHost,
                                           c->stream), "D2H");
                hipcheck(c, hipStreamSynchronize(c->stream), "sync");
                for (int v : hk)
                    if (v < 0 || v >= D0.nk) {
                        c->err = "kidx out of range [0, nk)";
                        return (int)NT_ERR_ARG;
                    }
            }
        }
        if (W <= 1) {
            for (int gi = 0; gi < ng; ++gi) solve_tdb_dev(c, p, args[gi][0]);
            return 0;
        }
        // Groups go to W worker contexts, largest first: small devices, whose
        // launches do not fill the GPU, run side by side.
        std::vector<const TdbArgs*> jobs(ng);
        for (int gi = 0; gi < ng; ++gi) jobs[gi] = &args[gi][0];
        run_on_workers(c, p, W, jobs);
        return 0;
    });
}

int nt_set_multi_workers(nt_ctx* c, int n)
{
    if (!c || n < 1 || n > 16) return NT_ERR_ARG;
    c->multi_workers = n;
    return NT_OK;
}

int nt_set_point_split(nt_ctx* c, int n)
{
    if (!c || n < 1 || n > 8) return NT_ERR_ARG;
    c->point_split = n;
    return NT_OK;
}
