"""The CPU checker for Doppler and speckle tracking (oracle/usx_oracle.cpp,
restating reference core/src/doppler.cpp), pinned three ways:

* bit for bit against the reference's own doppler.cpp compiled from its
  sources (oracle/_ref), over wall filters, orders, both estimators, window
  sizes, ensemble lengths and grid edge cases;
* against closed forms no restatement can fake: a slow-time phase ramp gives
  v = -w c PRF / (4 pi f) with coherence 1, a constant ensemble has no power
  after mean subtraction, an integer-sample shift is found as that lag, a
  linear displacement has that strain;
* the product's host-built wall-filter projector against the reference's
  (recovered through apply_wall_filter on impulse ensembles) and against the
  algebra of a projector (symmetric, idempotent, annihilates the polynomials
  it removes, trace n - order - 1).

No GPU is used here.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as Omod

HAVE_REF = Omod.available("reference")
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (no reference sources)")

C = 1540.0
FDEM = 5e6
PRF = 6000.0


def _grid(O, n0=7, n1=33, z0=0.020, z1=0.024):
    return O.GridSpec(0, -0.003, 0.003, n0, z0, z1, n1)


def _ensemble(n, n0, n1, seed=0, ramp=0.3, clutter=3.0):
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((n, n0, n1)) + 1j * rng.standard_normal((n, n0, n1))
    slow = np.exp(1j * ramp * np.arange(n))[:, None, None]
    fast = np.exp(1j * np.linspace(0.0, 3.0, n1))[None, None, :]
    return (0.2 * noise + slow * fast + clutter).astype(np.complex64)


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8)


@needs_ref
@pytest.mark.parametrize("n", [2, 3, 10, 17, 40])
@pytest.mark.parametrize("wall,order", [(0, 1), (1, 0), (1, 1), (1, 3), (2, 1), (1, 50)])
def test_wall_filter_and_power_doppler_equal_the_reference(O, port, n, wall, order):
    ref = O.load("reference")
    g = _grid(O)
    ens = _ensemble(n, g.n0, g.n1, seed=n + 7 * wall + order)
    a, b = port.wall_filter(g, ens, wall, order), ref.wall_filter(g, ens, wall, order)
    assert np.array_equal(_bits(a), _bits(b))
    a, b = port.power_doppler(g, ens, wall, order), ref.power_doppler(g, ens, wall, order)
    assert np.array_equal(_bits(a), _bits(b))


@needs_ref
@pytest.mark.parametrize("n", [2, 5, 16, 33])
@pytest.mark.parametrize("wall", [0, 1, 2])
@pytest.mark.parametrize("loupas,ka,kl", [(1, 4, 1), (0, 0, 0), (1, 6, 2), (1, -3, -1)])
def test_color_flow_equals_the_reference(O, port, n, wall, loupas, ka, kl):
    ref = O.load("reference")
    g = _grid(O)
    ens = _ensemble(n, g.n0, g.n1, seed=3 * n + wall, ramp=0.4, clutter=1.0 + wall)
    d = O.Doppler(wall_filter=wall, polynomial_order=1, loupas=loupas, kernel_axial=ka,
                  kernel_lateral=kl, power_threshold_db=-20.0, coherence_threshold=0.3)
    a = port.color_flow(g, ens, FDEM, PRF, C, d)
    b = ref.color_flow(g, ens, FDEM, PRF, C, d)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(_bits(x), _bits(y))
    assert a[4] == b[4]


@needs_ref
@pytest.mark.parametrize("shape", [(7, 33), (1, 40), (5, 1), (3, 2)])
def test_color_flow_grid_edges_equal_the_reference(O, port, shape):
    ref = O.load("reference")
    g = _grid(O, *shape)
    ens = _ensemble(6, *shape, seed=11)
    d = O.Doppler()
    a = port.color_flow(g, ens, FDEM, PRF, C, d)
    b = ref.color_flow(g, ens, FDEM, PRF, C, d)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(_bits(x), _bits(y))


@needs_ref
@pytest.mark.parametrize("shift,ka,kl,lag", [(2, 4, 1, 3), (0, 8, 2, 8), (-3, 5, 0, 0), (5, 1, 1, 8)])
def test_displacement_and_strain_equal_the_reference(O, port, shift, ka, kl, lag):
    ref = O.load("reference")
    g = _grid(O, 6, 48)
    before = _ensemble(1, g.n0, g.n1, seed=5 + shift, clutter=0.0)[0]
    after = (np.roll(before, shift, axis=1) * np.exp(-0.2j)).astype(np.complex64)
    a = port.axial_displacement(g, before, after, FDEM, C, ka, kl, lag)
    b = ref.axial_displacement(g, before, after, FDEM, C, ka, kl, lag)
    assert np.array_equal(_bits(a[0]), _bits(b[0]))
    assert np.array_equal(_bits(a[1]), _bits(b[1]))
    for fit in (1, 2, 15):
        assert np.array_equal(_bits(port.axial_strain(g, a[0], fit)),
                              _bits(ref.axial_strain(g, a[0], fit)))


def test_phase_ramp_gives_the_closed_form_velocity(O, port):
    """z_e = exp(j w e): R01 = (n - 1) * window * exp(j w) and R00 = n * window
    at every pixel, so v = -w c PRF / (4 pi f0) (Kasai) and the spectral width
    2 (1 - |R01| / R00) is 2 / n."""
    g = _grid(O, 5, 20)
    n, w = 8, 0.5
    ens = np.broadcast_to(np.exp(1j * w * np.arange(n))[:, None, None], (n, 5, 20)).astype(np.complex64)
    d = O.Doppler(wall_filter=0, loupas=0, kernel_axial=2, kernel_lateral=1)
    v, p, var, ok, nyq = port.color_flow(g, ens, FDEM, PRF, C, d)
    expect = -w * C * PRF / (4 * np.pi * FDEM)
    assert np.all(ok == 1)
    np.testing.assert_allclose(v, expect, rtol=1e-5)
    np.testing.assert_allclose(var, 2.0 / n, atol=1e-5)
    # interior pixel: 3 lines x 5 depths x 8 frames of unit power
    assert p[2, 10] == pytest.approx(3 * 5 * n, rel=1e-6)
    assert nyq == pytest.approx(C * PRF / (4 * FDEM), rel=1e-7)


def test_loupas_uses_the_axial_phase_gradient(O, port):
    """An axial phase advance of a rad per pixel means f = a c / (4 pi dz):
    the Loupas estimate rescales the velocity by f0 / f (a < pi here)."""
    g = _grid(O, 3, 81, 0.020, 0.022)
    dz = (g.a1_max - g.a1_min) / (g.n1 - 1)
    n, w = 6, 0.4
    f_true = 4.0e6
    a = 4 * np.pi * f_true * dz / C
    ens = (np.exp(1j * w * np.arange(n))[:, None, None] *
           np.exp(1j * a * np.arange(g.n1))[None, None, :] * np.ones((1, g.n0, 1))).astype(np.complex64)
    d = O.Doppler(wall_filter=0, loupas=1)
    v, *_ = port.color_flow(g, ens, FDEM, PRF, C, d)
    np.testing.assert_allclose(v[1, 5:-5], -w * C * PRF / (4 * np.pi * f_true), rtol=1e-4)


def test_mean_subtraction_removes_a_constant_ensemble(O, port):
    g = _grid(O, 4, 9)
    ens = np.full((7, 4, 9), 2.0 - 1.5j, np.complex64)
    assert np.max(np.abs(port.wall_filter(g, ens, 2, 0))) < 1e-6
    assert np.max(port.power_doppler(g, ens, 2, 0)) < 1e-10
    # and the polynomial filter of order 1 removes a linear trend exactly (to rounding)
    trend = (np.arange(7)[:, None, None] * (0.3 + 0.1j) + 1.0).astype(np.complex64) * np.ones((1, 4, 9))
    assert np.max(np.abs(port.wall_filter(g, trend.astype(np.complex64), 1, 1))) < 1e-5


def test_power_doppler_without_filter_is_the_energy(O, port):
    g = _grid(O, 3, 5)
    ens = _ensemble(4, 3, 5, seed=2)
    np.testing.assert_allclose(port.power_doppler(g, ens, 0, 1),
                               np.sum(np.abs(ens.astype(np.complex128)) ** 2, axis=0), rtol=1e-6)


def test_integer_shift_is_found_and_linear_displacement_has_that_strain(O, port):
    g = _grid(O, 4, 64)
    dz = (g.a1_max - g.a1_min) / (g.n1 - 1)
    before = _ensemble(1, 4, 64, seed=9, clutter=0.0)[0]
    after = np.roll(before, 3, axis=1)
    ax, co = port.axial_displacement(g, before, after, FDEM, C, 6, 1, 5)
    inner = ax[:, 12:-12]
    np.testing.assert_allclose(inner, 3 * dz, atol=1e-9)
    assert np.all(co[:, 12:-12] > 0.999)
    disp = np.broadcast_to((2e-3 * np.asarray(np.linspace(0, 1, 64)) * 0.004)[None, :], (4, 64))
    s = port.axial_strain(g, disp.astype(np.float32), 9)
    np.testing.assert_allclose(s[:, 5:-5], 2e-3 * 0.004 / (g.a1_max - g.a1_min), rtol=2e-3)


@needs_ref
@pytest.mark.parametrize("n,order", [(1, 1), (2, 1), (5, 0), (8, 1), (16, 2), (12, 20), (33, 3)])
def test_product_projector_is_the_reference_projector(usx_mod, O, n, order):
    """The reference builds its projector privately; applying its
    apply_wall_filter to unit impulses reads it out column by column."""
    ref = O.load("reference")
    g = O.GridSpec(0, -0.001, 0.001, 1, 0.02, 0.021, n)
    # impulse ensemble: frame j is 1 at depth j, so output frame i at depth j is M[i][j]
    ens = np.zeros((n, 1, n), np.complex64)
    for j in range(n):
        ens[j, 0, j] = 1.0
    out = ref.wall_filter(g, ens, 1, order)
    m_ref = np.real(out[:, 0, :]).astype(np.float32)
    m = np.asarray(usx_mod.polynomial_reject_matrix(n, order), np.float32)
    assert np.array_equal(m.view(np.uint32), m_ref.view(np.uint32))


@pytest.mark.parametrize("n,order", [(2, 1), (8, 0), (8, 1), (16, 2), (24, 3)])
def test_product_projector_algebra(usx_mod, n, order):
    m = np.asarray(usx_mod.polynomial_reject_matrix(n, order), np.float64)
    np.testing.assert_allclose(m, m.T, atol=1e-6)
    np.testing.assert_allclose(m @ m, m, atol=1e-5)
    x = np.linspace(-1, 1, n)
    for d in range(min(order, n - 1) + 1):
        assert np.max(np.abs(m @ x ** d)) < 1e-5
    assert np.trace(m) == pytest.approx(n - min(order, n - 1) - 1, abs=1e-4)
