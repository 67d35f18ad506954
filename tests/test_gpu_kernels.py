"""The two batched kernels (complex GEMM, Gauss-Jordan inverse) against numpy,
through the building-block entry points of the C ABI, over the block sizes of
the workloads (24..768) and the awkward ones in between."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    return torch, torch.device("cuda", 0)


def to_dev(torch, dev, mats):
    """list of (m, n) numpy -> device buffer of column-major matrices"""
    a = np.ascontiguousarray(np.swapaxes(np.stack(mats), -1, -2))
    return torch.from_numpy(a).to(dev)


def from_dev(t):
    return np.swapaxes(t.cpu().numpy(), -1, -2)


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 5, 5), (16, 16, 16), (17, 9, 33), (37, 37, 37), (64, 64, 16),
                                   (65, 33, 47), (100, 37, 200), (192, 192, 192), (250, 130, 70)])
@pytest.mark.parametrize("ops", ["NN", "NC", "CN", "CC"])
def test_zgemm_batched(engine, torch_dev, m, n, k, ops):
    _zgemm_vs_numpy(engine, torch_dev, m, n, k, ops)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (24, 24, 24), (17, 9, 33), (33, 65, 40), (80, 80, 80), (96, 96, 96),
                                   (100, 37, 200), (216, 216, 216)])
@pytest.mark.parametrize("ops", ["NN", "NC", "CN", "CC"])
def test_zgemm_tile32(engine, torch_dev, m, n, k, ops):
    """the one-wave 32x32 output tile (selected per size) against numpy"""
    engine.set_gemm_tile(max(m, n), 32)
    try:
        _zgemm_vs_numpy(engine, torch_dev, m, n, k, ops)
    finally:
        engine.set_gemm_tile(max(m, n), 0)


def test_gemm_autotune(engine):
    for n in (24, 96, 192):
        tile, ms = engine.autotune_gemm(n, 256)
        print(f"gemm tile for n={n}: {tile} ({ms:.3f} ms for 256 items)", flush=True)
        assert tile in (32, 64) and ms > 0
        engine.set_gemm_tile(n, 0)


def _zgemm_vs_numpy(engine, torch_dev, m, n, k, ops):
    from phinfortran_amd import _native as N
    torch, dev = torch_dev
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    batch = 3
    A = [crand(rng, *((m, k) if ops[0] == "N" else (k, m))) for _ in range(batch)]
    B = [crand(rng, *((k, n) if ops[1] == "N" else (n, k))) for _ in range(batch)]
    C0 = [crand(rng, m, n) for _ in range(batch)]
    alpha, beta = 0.7 - 0.2j, -0.4 + 1.1j
    dA, dB, dC = to_dev(torch, dev, A), to_dev(torch, dev, B), to_dev(torch, dev, C0)
    lda = A[0].shape[0]
    ldb = B[0].shape[0]
    print(f"zgemm {m}x{n}x{k} {ops}", flush=True)
    st = N.lib().nt_zgemm_batched_dev(engine.ctx, ops[0].encode(), ops[1].encode(), m, n, k,
                                      N.NtC128(alpha.real, alpha.imag), C.c_void_p(dA.data_ptr()), lda,
                                      A[0].size, C.c_void_p(dB.data_ptr()), ldb, B[0].size,
                                      N.NtC128(beta.real, beta.imag), C.c_void_p(dC.data_ptr()), m, m * n, batch)
    assert st == 0
    engine.sync()
    got = from_dev(dC)
    for i in range(batch):
        a = A[i] if ops[0] == "N" else A[i].conj().T
        b = B[i] if ops[1] == "N" else B[i].conj().T
        ref = alpha * a @ b + beta * C0[i]
        assert np.abs(got[i] - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()) * max(1, k) ** 0.5


def test_zgemm_shared_operand(engine, torch_dev):
    """stride 0 shares B across the batch (H(k) blocks across energies)"""
    from phinfortran_amd import _native as N
    torch, dev = torch_dev
    rng = np.random.default_rng(3)
    m = n = k = 40
    A = [crand(rng, m, k) for _ in range(4)]
    B = crand(rng, k, n)
    dA, dB = to_dev(torch, dev, A), to_dev(torch, dev, [B])
    dC = torch.zeros((4, n, m), dtype=torch.complex128, device=dev)
    st = N.lib().nt_zgemm_batched_dev(engine.ctx, b"N", b"N", m, n, k, N.NtC128(1, 0), C.c_void_p(dA.data_ptr()), m,
                                      m * k, C.c_void_p(dB.data_ptr()), k, 0, N.NtC128(0, 0),
                                      C.c_void_p(dC.data_ptr()), m, m * n, 4)
    assert st == 0
    engine.sync()
    got = from_dev(dC)
    for i in range(4):
        assert np.abs(got[i] - A[i] @ B).max() < 1e-11


@pytest.mark.parametrize("n", [1, 2, 5, 16, 24, 31, 32, 33, 37, 48, 64, 80, 96, 150, 192, 193, 216, 256, 300,
                               384, 500, 768])
def test_zinv_batched(engine, torch_dev, n):
    from phinfortran_amd import _native as N
    torch, dev = torch_dev
    rng = np.random.default_rng(n)
    batch = 3
    A = [crand(rng, n, n) + (0.5 * n ** 0.5) * np.eye(n) * (i % 2) for i in range(batch)]
    if n > 1:
        A[0][0, 0] = 0.0          # first pivot must come from another row
    dA = to_dev(torch, dev, A)
    dX = torch.zeros_like(dA)
    info = torch.full((batch,), 0, dtype=torch.int32, device=dev)
    print(f"zinv n={n}", flush=True)
    st = N.lib().nt_zinv_batched_dev(engine.ctx, n, C.c_void_p(dA.data_ptr()), n * n, C.c_void_p(dX.data_ptr()),
                                     n * n, C.c_void_p(info.data_ptr()), batch)
    assert st == 0
    engine.sync()
    assert (info.cpu().numpy() == 0).all()
    X = from_dev(dX)
    for i in range(batch):
        ref = np.linalg.inv(A[i])
        cond = np.linalg.cond(A[i])
        assert np.abs(X[i] - ref).max() < 1e-14 * cond * max(1.0, np.abs(ref).max()) * n, (i, cond)


def test_zinv_singular_and_nan(engine, torch_dev):
    from phinfortran_amd import _native as N
    torch, dev = torch_dev
    n = 40
    rng = np.random.default_rng(0)
    S = crand(rng, n, n)
    S[:, 7] = 0.0                  # exact zero column -> zero pivot at column 8
    Nn = crand(rng, n, n)
    Nn[3, 5] = np.nan
    dA = to_dev(torch, dev, [S, Nn])
    dX = torch.zeros_like(dA)
    info = torch.zeros(2, dtype=torch.int32, device=dev)
    st = N.lib().nt_zinv_batched_dev(engine.ctx, n, C.c_void_p(dA.data_ptr()), n * n, C.c_void_p(dX.data_ptr()),
                                     n * n, C.c_void_p(info.data_ptr()), 2)
    assert st == 0
    engine.sync()
    inf = info.cpu().numpy()
    assert inf[0] == 8
    assert inf[1] == -1


@pytest.mark.parametrize("n", [24, 96, 192, 200, 384, 500])
def test_zinv_every_panel_width(engine, torch_dev, n):
    """Each panel width the autotuner may pick inverts to the same accuracy,
    and the autotuner returns one of them."""
    from phinfortran_amd import _native as N
    torch, dev = torch_dev
    rng = np.random.default_rng(100 + n)
    batch = 2
    A = [crand(rng, n, n) + 0.5 * n ** 0.5 * np.eye(n) for _ in range(batch)]
    A[0][0, 0] = 0.0
    widths = [w for w, lim in ((32, 192), (24, 256), (20, 256), (16, 384), (8, 1024)) if n <= lim]
    if n >= 64:
        widths.append(1)        # panel kernels + batched GEMM
    print(f"zinv panels n={n} widths={widths}", flush=True)
    for w in widths:
        engine.set_inverse_panel(n, w)
        dA = to_dev(torch, dev, A)
        dX = torch.zeros_like(dA)
        info = torch.zeros(batch, dtype=torch.int32, device=dev)
        st = N.lib().nt_zinv_batched_dev(engine.ctx, n, C.c_void_p(dA.data_ptr()), n * n,
                                         C.c_void_p(dX.data_ptr()), n * n, C.c_void_p(info.data_ptr()), batch)
        assert st == 0
        engine.sync()
        X = from_dev(dX)
        for i in range(batch):
            ref = np.linalg.inv(A[i])
            assert np.abs(X[i] - ref).max() < 1e-14 * np.linalg.cond(A[i]) * np.abs(ref).max() * n, (w, i)
    engine.set_inverse_panel(n, 0)
    nb, ms = engine.autotune_inverse(n, 64)
    assert nb in widths and ms > 0
    engine.set_inverse_panel(n, 0)


@pytest.mark.parametrize("m,n,k", [(5, 5, 5), (37, 37, 37), (192, 192, 192), (250, 130, 70)])
@pytest.mark.parametrize("ops", ["NN", "CC", "NC"])
def test_zgemm_three_products(engine, torch_dev, m, n, k, ops):
    """the 3-multiplication complex product agrees with numpy to FP64 rounding"""
    from phinfortran_amd import _native as N
    torch, dev = torch_dev
    rng = np.random.default_rng(m + 7 * n + 13 * k)
    batch = 2
    A = [crand(rng, *((m, k) if ops[0] == "N" else (k, m))) for _ in range(batch)]
    B = [crand(rng, *((k, n) if ops[1] == "N" else (n, k))) for _ in range(batch)]
    C0 = [crand(rng, m, n) for _ in range(batch)]
    alpha, beta = 0.7 - 0.2j, -0.4 + 1.1j
    dA, dB, dC = to_dev(torch, dev, A), to_dev(torch, dev, B), to_dev(torch, dev, C0)
    engine.set_gemm_3m(True)
    try:
        st = N.lib().nt_zgemm_batched_dev(engine.ctx, ops[0].encode(), ops[1].encode(), m, n, k,
                                          N.NtC128(alpha.real, alpha.imag), C.c_void_p(dA.data_ptr()),
                                          A[0].shape[0], A[0].size, C.c_void_p(dB.data_ptr()), B[0].shape[0],
                                          B[0].size, N.NtC128(beta.real, beta.imag), C.c_void_p(dC.data_ptr()), m,
                                          m * n, batch)
        engine.sync()
    finally:
        engine.set_gemm_3m(False)
    assert st == 0
    got = from_dev(dC)
    for i in range(batch):
        a = A[i] if ops[0] == "N" else A[i].conj().T
        b = B[i] if ops[1] == "N" else B[i].conj().T
        ref = alpha * a @ b + beta * C0[i]
        assert np.abs(got[i] - ref).max() < 2e-12 * max(1.0, np.abs(ref).max()) * max(1, k) ** 0.5


@pytest.mark.parametrize("batch,groups_like", [(8, 1), (16, 1), (24, 1)])
def test_zgemm_batch_multiple_of_xcds(engine, torch_dev, batch, groups_like):
    """batches divisible by the XCD count take the renumbered tile order"""
    from phinfortran_amd import _native as N
    torch, dev = torch_dev
    rng = np.random.default_rng(batch)
    m, n, k = 130, 150, 70
    A = [crand(rng, m, k) for _ in range(batch)]
    B = [crand(rng, k, n) for _ in range(batch)]
    C0 = [crand(rng, m, n) for _ in range(batch)]
    dA, dB, dC = to_dev(torch, dev, A), to_dev(torch, dev, B), to_dev(torch, dev, C0)
    for m3 in (False, True):
        engine.set_gemm_3m(m3)
        dC.copy_(to_dev(torch, dev, C0))
        st = N.lib().nt_zgemm_batched_dev(engine.ctx, b"N", b"N", m, n, k, N.NtC128(1.0, 0.0),
                                          C.c_void_p(dA.data_ptr()), m, m * k, C.c_void_p(dB.data_ptr()), k, k * n,
                                          N.NtC128(1.0, 0.0), C.c_void_p(dC.data_ptr()), m, m * n, batch)
        engine.sync()
        engine.set_gemm_3m(False)
        assert st == 0
        got = from_dev(dC)
        for i in range(batch):
            ref = A[i] @ B[i] + C0[i]
            assert np.abs(got[i] - ref).max() < 2e-12 * np.abs(ref).max() * k ** 0.5, (m3, i)


@pytest.mark.parametrize("width", [16, 1])
def test_zinv_singular_and_nan_n96(engine, torch_dev, width):
    """status codes of the one-workgroup and the panel + GEMM inverse: the
    first exact zero pivot (1-based column), -1 for NaN input"""
    from phinfortran_amd import _native as N
    torch, dev = torch_dev
    n = 96
    rng = np.random.default_rng(1)
    S = crand(rng, n, n)
    S[:, 40] = 0.0                 # exact zero column -> zero pivot at column 41
    Nn = crand(rng, n, n)
    Nn[70, 50] = np.nan
    dA = to_dev(torch, dev, [S, Nn])
    dX = torch.zeros_like(dA)
    info = torch.zeros(2, dtype=torch.int32, device=dev)
    engine.set_inverse_panel(n, width)
    st = N.lib().nt_zinv_batched_dev(engine.ctx, n, C.c_void_p(dA.data_ptr()), n * n, C.c_void_p(dX.data_ptr()),
                                     n * n, C.c_void_p(info.data_ptr()), 2)
    engine.set_inverse_panel(n, 0)
    assert st == 0
    engine.sync()
    inf = info.cpu().numpy()
    assert inf[0] == 41
    assert inf[1] == -1
