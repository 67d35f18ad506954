// The batched solve's argument block and the entries shared by the
// translation units of the library: solve_tdb_dev (nt_engine.hip) solves one
// batch of points on a context; solve_split (nt_workers.hip) solves them as
// parts side by side on worker contexts.
#pragma once
#include <vector>
#include "nt_host.h"

// the probe path's inputs and outputs (device pointers; outputs per point:
// flux, a_sum, gn_sum L doubles, jac L x L, flux_left, flux_right one)
struct ProbeSpec {
    const double* temp;       // L probe temperatures
    double t_left, t_right, B, C;
    int lifetime;
    double *flux, *jac, *flux_left, *flux_right, *a_sum, *gn_sum;
};

struct TdbArgs {
    int nl, nr, b, L, npts, nk;
    const double* E;
    const int* kidx;
    const nt_c128 *Haml, *Hamlv, *Hamr, *Hamrv, *Hamd, *Hamdv, *Hamld, *Hamrd;
    const int* first_mask;
    double *T_LR, *T_RL;
    int *status, *iters;
    nt_solve_extras ex;
    // adjoint of sum_p w_p T_LR(p) (w null: no adjoint)
    const double* adj_w = nullptr;
    int adj_lo = 0, adj_hi = -1;
    nt_c128* dHd = nullptr;    // (nk, hi - lo + 1, b, b), accumulated
    nt_c128* dHdv = nullptr;   // (nk, min(hi, L-2) - lo + 1, b, b), accumulated
    const ProbeSpec* probe = nullptr;
    // workspace in the static scratch (no device allocation); the chunk of
    // points shrinks until the plan fits
    bool noalloc = false;
    // devices with the same leads, points (E, k map) and lead sizes: their
    // RGF sweeps reuse this call's lead Green's functions (FP64, no extras)
    std::vector<const TdbArgs*> siblings;
};

// Solve a's points on c (its stream and workspace, chunked as they fit).
// Throws Fail; runs inside guarded().
void solve_tdb_dev(nt_ctx* c, const nt_params* prm, const TdbArgs& a);

// Solve a's points (FP64 sweep, no probes, no siblings) as `parts` parts side
// by side on worker contexts of c, each on its own stream with its own
// workspace, ordered after and before c's stream; results bitwise those of
// solve_tdb_dev on one stream (an adjoint's gradients to the rounding of the
// sum over parts).  Throws Fail; runs inside guarded().
void solve_split(nt_ctx* c, const nt_params* p, const TdbArgs& a, int parts);
