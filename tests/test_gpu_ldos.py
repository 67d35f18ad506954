"""get_ldos on the GPU against the oracle (itself pinned to the diagonal of
the dense device Green's function), tb and negf conventions, uniform and
variable layers with dense lower blocks."""
import numpy as np
import pytest

import oracle as O
from phinfortran_amd import dense as D
from phinfortran_amd import lattice as lt

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from phinfortran_amd import transport
    return transport


def oparams(p):
    return O.OrParams(p.eta, p.eta0, p.convergence, p.device_eta_factor, p.eta_mode, p.max_iter)


@pytest.mark.parametrize("kind,nl,b,L", [("tb", 6, 6, 5), ("phonon", 12, 12, 3), ("tb", 40, 40, 6)])
def test_ldos_tdb(engine, T, kind, nl, b, L):
    blk = lt.random_blocks(nl, b, L, seed=b + L)
    p = T.params(kind, eta=0.01 if kind == "tb" else 1e-3, eta0=0.0 if kind == "tb" else 1e-4)
    op = oparams(p)
    E = 0.29
    sg = O.surface_green(blk.Haml, blk.Hamlv, blk.Hamr, blk.Hamrv, E, op, want_plus_minus=False)
    gl, gr = sg["gl_adv"].conj().T, sg["gr_ret"]
    print(f"ldos {kind} nl={nl} b={b} L={L}", flush=True)
    ld, dd = engine.get_ldos(gl, gr, E, list(blk.Hamd), list(blk.Hamdv), blk.Hamld, blk.Hamrd, p)
    lo, do = O.get_ldos(list(blk.Hamd), list(blk.Hamdv), blk.Hamld, blk.Hamrd, gl, gr, E, op)
    assert np.abs(ld - lo).max() < 1e-9 * np.abs(lo).max()
    assert abs(dd - do) < 1e-9 * abs(do)


def test_ldos_variable_layers_dense(engine, T):
    blk = lt.random_blocks(4, 4, 6, seed=5)
    H = D.assemble(blk)
    layers = [4, 8, 4, 8]
    s = D.slices(H, layers, 4, 4, variant="tb")
    p = T.params("tb", eta=0.02)
    op = oparams(p)
    E = -0.2
    sg = O.surface_green(s.Haml, s.Hamlv, s.Hamr, s.Hamrv, E, op, want_plus_minus=False)
    gl, gr = sg["gl_adv"].conj().T, sg["gr_ret"]
    ld, dd = engine.get_ldos(gl, gr, E, s.Hd, s.Hdv, s.Hld, s.Hrd, p, Hdv_back=s.Hdv_back, Hrd_fwd=s.Hrd_fwd)
    lo, do = O.get_ldos(s.Hd, s.Hdv, s.Hld, s.Hrd, gl, gr, E, op, Hdv_back=s.Hdv_back, Hrd_fwd=s.Hrd_fwd)
    assert ld.shape == (4, 8)
    assert np.abs(ld - lo).max() < 1e-9 * np.abs(lo).max()
    assert np.all(ld[0, 4:] == 0) and np.all(ld[2, 4:] == 0)


@pytest.mark.parametrize("kind,nl,b,L,nk", [("tb", 6, 6, 5, 2), ("phonon", 12, 12, 3, 1), ("tb", 24, 24, 1, 2),
                                            ("tb", 40, 40, 7, 1)])
@pytest.mark.parametrize("entry", ["alloc", "noalloc"])
def test_ldos_batched(engine, T, kind, nl, b, L, nk, entry):
    """the ldos / ddos outputs of the batched solve (diagonal blocks of the
    full device G by the backward recursion over the kept left-connected
    blocks) against the oracle's get_ldos at every point; T unchanged"""
    blocks = [lt.random_blocks(nl, b, L, seed=3 * b + L + i) for i in range(nk)]
    p = T.params(kind, eta=0.01 if kind == "tb" else 1e-3, eta0=0.0 if kind == "tb" else 1e-4)
    op = oparams(p)
    E = np.array([-0.41, 0.07, 0.29, 0.63, 1.1])
    E = np.tile(E if kind == "tb" else np.abs(E), nk)
    kidx = np.repeat(np.arange(nk, dtype=np.int32), 5)
    d = engine.upload_blocks(blocks)
    print(f"batched ldos {kind} nl={nl} b={b} L={L} nk={nk} {entry}", flush=True)
    r = engine.solve_tdb(d, E, p, kidx=kidx, extras=("ldos", "ddos"), noalloc=entry == "noalloc")
    r0 = engine.solve_tdb(d, E, p, kidx=kidx)
    assert (r["status"].cpu().numpy() == 0).all()
    ld, dd = r["ldos"].cpu().numpy(), r["ddos"].cpu().numpy()
    assert ld.shape == (len(E), L, b)
    for i, e in enumerate(E):
        blk = blocks[kidx[i]]
        sg = O.surface_green(blk.Haml, blk.Hamlv, blk.Hamr, blk.Hamrv, e, op, want_plus_minus=False)
        lo, do = O.get_ldos(list(blk.Hamd), list(blk.Hamdv), blk.Hamld, blk.Hamrd, sg["gl_adv"].conj().T,
                            sg["gr_ret"], e, op)
        assert np.abs(ld[i] - lo).max() < 1e-9 * np.abs(lo).max(), (i, e)
        assert abs(dd[i] - do) < 1e-9 * abs(do), (i, e)
        assert abs(dd[i] - ld[i].sum()) < 1e-12 * abs(do)
    for k in ("T_LR", "T_RL"):
        a, b0 = r[k].cpu().numpy(), r0[k].cpu().numpy()
        assert np.abs(a - b0).max() <= 1e-9 * max(1.0, np.abs(b0).max()), k


def test_ldos_batched_mixed_runs_fp64(engine, T):
    """with NT_PREC_MIXED the device DOS solve runs in FP64: the same values"""
    blk = lt.random_blocks(16, 16, 4, seed=9)
    d = engine.upload_blocks(blk)
    E = np.array([0.1, 0.35, 0.8])
    a = engine.solve_tdb(d, E, T.params("tb", eta=0.01), extras=("ldos", "ddos"))
    b = engine.solve_tdb(d, E, T.params("tb", eta=0.01, precision="mixed"), extras=("ldos", "ddos"))
    for k in ("ldos", "ddos", "T_LR"):
        assert np.array_equal(a[k].cpu().numpy(), b[k].cpu().numpy()), k
