"""Tecplot ASCII writer (host formatter of the library) and reader round trip, and pins of
the output-variable restatement the GPU writer test compares with."""
import numpy as np
import pytest

from trans_amd import tecplot

import tecplot_fields as tf


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    ni, nj, nk = 5, 4, 3
    x, y, z = (rng.normal(size=(ni, nj, nk)) for _ in range(3))
    cv = {"rho": rng.uniform(0.5, 2.0, (ni - 1, nj - 1, nk - 1)), "p q": -rng.lognormal(size=(ni - 1, nj - 1, nk - 1)) * 1e7,
          "tiny": rng.normal(size=(ni - 1, nj - 1, nk - 1)) * 1e-300}
    path = tmp_path / "f.dat"
    tecplot.write_ascii(path, x, y, z, cv, title="t", zone="block 1")
    head = path.read_text().splitlines()[:3]
    assert head[0] == 'TITLE = "t"'
    assert head[1] == 'VARIABLES = "X" "Y" "Z" "rho" "p q" "tiny"'
    assert head[2] == ('ZONE T="block 1", I=5, J=4, K=3, ZONETYPE=ORDERED, DATAPACKING=BLOCK, '
                       'VARLOCATION=([4-6]=CELLCENTERED)')
    r = tecplot.read_ascii(path)
    assert r["dims"] == (ni, nj, nk) and r["zone"] == "block 1"
    assert list(r["variables"]) == ["X", "Y", "Z", "rho", "p q", "tiny"]
    for n, a in [("X", x), ("Y", y), ("Z", z)] + list(cv.items()):
        np.testing.assert_allclose(r["variables"][n], a, rtol=6e-11, atol=0)
    # five values to a line, i fastest: the first line of X holds x[0:5, 0, 0]
    first = np.array(path.read_text().splitlines()[3].split(), dtype=float)
    np.testing.assert_allclose(first, x[:, 0, 0], rtol=6e-11)


def test_single_cell_variable_and_rejects_quotes(tmp_path):
    x = np.zeros((2, 2, 2))
    tecplot.write_ascii(tmp_path / "a.dat", x, x, x, {"one": np.ones((1, 1, 1))})
    assert "VARLOCATION=([4]=CELLCENTERED)" in (tmp_path / "a.dat").read_text()
    with pytest.raises(OSError):
        tecplot.write_ascii(tmp_path / "b.dat", x, x, x, {'a"b': np.ones((1, 1, 1))})


def _gradients(G, shape=(2, 2, 2)):
    dq = np.zeros((18,) + shape)
    for a in range(3):
        for e in range(3):
            dq[3 * a + e] = G[a][e]
    return dq


def _state(shape=(2, 2, 2), rho=1.2, u=(0.3, -0.1, 0.2), p=2.0, rk=0.0):
    v = np.zeros((7,) + shape)
    v[0] = rho
    for q in range(3):
        v[1 + q] = rho * u[q]
    v[5] = rk
    return v, np.full(shape, p), np.full(shape, p / rho)


def test_restatement_rigid_rotation():
    """Velocity gradient of a rigid rotation Omega about z (du/dy = -Omega, dv/dx = Omega):
    no strain, vorticity (0, 0, Omega) in the reference's half-curl convention, Om = 2 Omega,
    Q = Omega^2, Yita4 = 0, no shear stresses."""
    om = 0.7
    v, pp, t = _state()
    dq = _gradients([[0, -om, 0], [om, 0, 0], [0, 0, 0]])
    r = tf.field_vars(v, pp, t, np.full(pp.shape, 0.01), dq, np.zeros(pp.shape), np.ones(pp.shape),
                      np.ones(pp.shape), 0.0)
    np.testing.assert_allclose(r["Oz"], om, rtol=1e-15)
    np.testing.assert_allclose(r["Om"], 2 * om, rtol=1e-15)
    np.testing.assert_allclose(r["Q"], om * om, rtol=1e-15)
    np.testing.assert_allclose(r["Yita4"], 0.0, atol=1e-16)
    np.testing.assert_allclose(r["uv"], 0.0, atol=1e-16)
    assert np.all(r["Ox"] == 0) and np.all(r["Oy"] == 0)


def test_restatement_pure_strain_and_stresses():
    """Pure shear du/dy = dv/dx = s: S12 = S21 = s and no rotation, so Om = 0 and Q = -(S12^2 +
    S21^2) / 2 = -s^2; the Boussinesq shear stress uv = 2 s mu_t, and the normal stresses of a
    divergence-free field are -2/3 rho k."""
    s, mut, rk = 0.4, 0.02, 0.05
    v, pp, t = _state(rk=rk)
    dq = _gradients([[0, s, 0], [s, 0, 0], [0, 0, 0]])
    r = tf.field_vars(v, pp, t, np.full(pp.shape, mut), dq, np.zeros(pp.shape), np.ones(pp.shape),
                      np.ones(pp.shape), 0.0)
    np.testing.assert_allclose(r["Om"], 0.0, atol=1e-16)
    np.testing.assert_allclose(r["Q"], -0.5 * (2 * s * s), rtol=1e-15)
    np.testing.assert_allclose(r["uv"], 2 * s * mut, rtol=1e-15)
    np.testing.assert_allclose(r["uu"], -2 * rk / 3, rtol=1e-15)
    np.testing.assert_allclose(r["k"], rk / 1.2, rtol=1e-15)


def test_restatement_absolute_frame():
    """A fluid at rest in the absolute frame seen from a frame rotating at omega about x has
    the relative velocity (0, omega z, -omega y): the absolute velocity vanishes, so Mabs = 0,
    P0 = 3.5 Ps and T0 = Ts (the reference's product form), while the relative totals see the
    blade speed."""
    om, yc, zc = 1.3, np.full((2, 2, 2), 0.4), np.full((2, 2, 2), -0.25)
    v, pp, t = _state(u=(0.0, om * -0.25, -om * 0.4))
    r = tf.field_vars(v, pp, t, np.zeros(pp.shape), np.zeros((18, 2, 2, 2)), np.zeros(pp.shape), yc, zc, om,
                      scales=(2.0, 3.0, 5.0, 7.0))
    np.testing.assert_allclose(r["Mabs"], 0.0, atol=1e-15)
    np.testing.assert_allclose(r["P0"], 3.5 * 2.0 * 5.0, rtol=1e-15)
    np.testing.assert_allclose(r["T0"], t * 7.0, rtol=1e-15)
    ma = np.hypot(om * 0.25, om * 0.4) / np.sqrt(1.4 * 2.0 / 1.2)
    np.testing.assert_allclose(r["Ma"], ma, rtol=1e-14)
    np.testing.assert_allclose(r["P0rel"], 2.0 * (1 + 0.2 * ma * ma) ** 3.5 * 5.0, rtol=1e-14)
    np.testing.assert_allclose(r["density"], 1.2 * 2.0, rtol=1e-15)
    np.testing.assert_allclose(r["Vabs"], 0.0, atol=1e-15)
