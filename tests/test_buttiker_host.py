"""Host steps of the Buttiker-probe loop (phinfortran_amd/buttiker.py) against
the probe oracle and closed forms: the starting temperatures, Simpson's
weights, the Newton update, the phonon temperature, the frequency integral
and the unformatted record layout of the output files."""
import os

import numpy as np
import pytest

from oracle import buttiker as BT
from phinfortran_amd import buttiker as PB


def test_initial_temperatures():
    assert np.allclose(PB.initial_temperatures(5, 300.0, 200.0), [275.0, 262.5, 250.0, 237.5, 225.0])
    assert PB.initial_temperatures(1, 310.0, 200.0).tolist() == [310.0]


def test_simpson_weights_match_the_rule():
    x = np.sqrt(np.linspace(0.0, 4.0, 13) ** 2 + 1e-6)
    x = np.linspace(0.1, 2.5, 13)
    rng = np.random.default_rng(0)
    y = rng.normal(size=13) + 1j * rng.normal(size=13)
    assert abs(PB.simpson_weights(x) @ y - BT.simpson(x, y)) < 1e-13 * np.abs(y).sum()
    with pytest.raises(ValueError):
        PB.simpson_weights(np.arange(6.0))


def test_newton_update_matches_oracle():
    rng = np.random.default_rng(1)
    J = rng.normal(size=(4, 4)) + 4 * np.eye(4) + 1e-14j * rng.normal(size=(4, 4))
    f = rng.normal(size=4) + 0j
    T = np.array([300.0, 310.0, 290.0, 305.0])
    a, da = PB.update_temperature(T, J, f)
    b, db = BT.update_temperature(T, J, f)
    assert np.allclose(a, b, rtol=1e-13) and abs(da - db) < 1e-12 * max(1.0, db)
    # a negative temperature resets every probe to 1 K
    c, _ = PB.update_temperature([1.0, 2.0], np.eye(2), np.array([5.0, 0.0]))
    assert c.tolist() == [1.0, 1.0]


def test_phonon_temperature_matches_oracle():
    E = np.linspace(0.02, 3.0, 31) ** 2
    rng = np.random.default_rng(2)
    A = np.abs(rng.normal(size=(31, 3))) + 0.1 + 0j
    Gn = np.stack([BT.bose(np.sqrt(E), t) for t in (240.0, 300.0, 350.0)], axis=1) * A * rng.uniform(0.9, 1.1, (31, 3))
    a = PB.get_ph_temperature(A, Gn, E, [300.0, 300.0, 300.0])
    b = BT.get_ph_temperature(A, Gn, E, [300.0, 300.0, 300.0])
    assert np.allclose(a, b, rtol=1e-9)


def test_integrate_jacobian_matches_oracle():
    import torch
    nk, ne, L = 3, 9, 4
    E = np.linspace(0.05, 2.0, ne) ** 2
    rng = np.random.default_rng(3)
    jac = rng.normal(size=(nk, ne, L, L))
    flux = rng.normal(size=(nk, ne, L))
    a_j, a_f = PB.integrate_jacobian(torch.from_numpy(jac.reshape(nk * ne, L, L)),
                                     torch.from_numpy(flux.reshape(nk * ne, L)), E, nk)
    b_j, b_f = BT.integrate_jacobian(jac + 0j, flux + 0j, E)
    assert np.allclose(a_j, b_j.real, rtol=1e-12, atol=1e-14) and np.allclose(a_f, b_f.real, rtol=1e-12, atol=1e-14)


def test_output_records(tmp_path):
    nk, ne, L = 2, 5, 3
    rng = np.random.default_rng(4)
    fl, gn, A = (rng.normal(size=(nk, ne, L)) for _ in range(3))
    left, right = rng.normal(size=(nk, ne)), rng.normal(size=(nk, ne))
    PB.write_flux(str(tmp_path), fl, gn, A, left, right)
    PB.write_temperature(str(tmp_path), np.array([1.0, 2.0, 3.0]), np.array([4.0, 5.0, 6.0]))
    raw = open(os.path.join(tmp_path, "flux.dat"), "rb").read()
    assert len(raw) == 8 * nk * ne * L + 8
    assert int(np.frombuffer(raw, "<i4", 1, 0)[0]) == 8 * nk * ne * L
    # Fortran order: k fastest, then energy, then layer
    rec = PB.read_fortran_records(os.path.join(tmp_path, "flux.dat"))
    assert len(rec) == 1 and np.array_equal(rec[0].reshape((nk, ne, L), order="F"), fl)
    assert np.array_equal(PB.read_fortran_records(os.path.join(tmp_path, "flux_right.dat"))[0].reshape(
        (nk, ne), order="F"), right)
    assert PB.read_fortran_records(os.path.join(tmp_path, "temperature_ph.dat"))[0].tolist() == [4.0, 5.0, 6.0]
