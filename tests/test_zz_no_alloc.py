"""The allocation-free batched solve (nt_solve_tdb_batched_noalloc_dev, its
workspace in the library's static scratch) against the allocating entry point
(nt_solve_tdb_batched_dev) on the same inputs: two calls in a row, the second
with a smaller input, in FP64 and in mixed precision, through the C ABI."""
import numpy as np
import pytest

from phinfortran_amd import lattice as lt

pytestmark = pytest.mark.gpu

KEYS = ("T_LR", "T_RL", "status", "iters", "gl_adv", "G_D_LR", "sdos_l")


def _same(a, b, what):
    for k in KEYS:
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        if x.dtype.kind in "iu":
            assert (x == y).all(), (what, k)
        else:
            scale = max(1.0, float(np.abs(x).max()))
            assert np.abs(x - y).max() <= 1e-12 * scale, (what, k, float(np.abs(x - y).max()))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_allocating_entry(engine, precision):
    from phinfortran_amd import transport as T
    ks = lt.mp_grid(8, 8)
    # b = 192, L = 40: the scratch holds about a dozen points, so 40 points
    # run in several chunks; then a smaller, different input
    big = [lt.si_ge_interface_film(nx=4, ny=2, L=40, kfrac=ks[k]) for k in (3, 20)]
    small = lt.si_ge_interface_film(nx=2, ny=1, L=9, kfrac=ks[7])
    p = T.params("phonon", precision=precision)
    cases = [(engine.upload_blocks(big), (16.0 * np.arange(1, 41) / 40) ** 2, np.arange(40, dtype=np.int32) % 2),
             (engine.upload_blocks(small), (16.0 * np.arange(1, 8) / 7.3) ** 2, None)]
    for i, (d, E, kidx) in enumerate(cases):
        print(f"noalloc {precision} call {i + 1}: b={d['b']} L={d['L']} points={len(E)}", flush=True)
        ref = engine.solve_tdb(d, E, p, kidx=kidx, extras=("gl_adv", "G_D_LR", "sdos_l"))
        got = engine.solve_tdb(d, E, p, kidx=kidx, extras=("gl_adv", "G_D_LR", "sdos_l"), noalloc=True)
        assert (got["status"].cpu().numpy() == 0).all()
        _same(ref, got, f"{precision} call {i + 1}")


def test_noalloc_rejects_what_does_not_fit(engine):
    """one b = 768, L = 100 point needs more than the scratch: a clear error,
    no launch"""
    from phinfortran_amd import _native as N
    from phinfortran_amd import transport as T
    blk = lt.rough_interface_film(nx=8, ny=4, L=100, kfrac=(0.1, 0.2))
    d = engine.upload_blocks(blk)
    with pytest.raises(N.TransportError) as e:
        engine.solve_tdb(d, np.array([4.0]), T.params("phonon"), noalloc=True)
    assert e.value.status == N.NT_ERR_ARG
    assert "static scratch" in str(e.value)
