"""Pins of the CPU oracle (oracle/sps_oracle.c) against mathematics.

The reference has no golden vectors or assertions for this path (its tests
print, tests/testregionalgrid.cpp and friends) and its build needs
CMake-generated headers, so the oracle is pinned here by:
  * polynomial exactness of every stencil (moment conditions of a consistent
    second-order stencil), at h != 1 so the divh scaling is pinned too;
  * Fourier symbols of lap5, lap9, lap7 and bilap = lap∘lap for <13>/<25>;
  * an independent numpy np.roll implementation of the Model B and
    Allen-Cahn right-hand sides (temp-grid semantics included);
  * exact recursions: a single Fourier mode of dpsi/dt = lap psi, a constant
    Allen-Cahn field, a constant Model B field;
  * invariants: Model B mass conservation and free-energy decrease;
  * the free-energy functional: a constant field's closed form (pins the
    1/2 and 1/4 and h^D), a single Fourier mode's closed form for the
    forward and the central gradient (pins kappa/2 and the difference
    scheme), discrete summation by parts against the pinned lap5/lap7, and
    linearity in (c1, c2, kappa);
  * the 2D ghost strips against modular indexing;
  * the reference's own quirks: lap19's "vz + vz", the 2D 21-point
    bilaplacian's 1/12 normalisation (a quarter of the bilaplacian) and the
    negative-stride sign of the Z-axis gradlaplacian.
"""
import itertools
import math

import numpy as np
import pytest

from symphas_amd import abi, models
from symphas_amd.models import Stencil2d2h, Stencil3d2h, TermSpec

LAP, BILAP, GRADLAP = abi.OP_LAP, abi.OP_BILAP, abi.OP_GRADLAP

ALL_2D = [(L, G, B) for L in (5, 9) for G in (6, 8, 12, 16) for B in (13, 17, 21)]
LAPS_2D, GRADS_2D, BILAPS_2D = (5, 9), (6, 8, 12, 16), (13, 17, 21)
LAPS_3D, GRADS_3D, BILAPS_3D = (7, 15, 19, 21, 27), (10, 12, 28, 36, 40), (21, 25, 41, 52, 57)


# ----------------------------------------------------------- polynomials
def _mono_deriv(exp, coef, axis, times=1):
    terms = {exp: coef}
    for _ in range(times):
        nxt = {}
        for e, c in terms.items():
            if e[axis] == 0:
                continue
            e2 = list(e)
            e2[axis] -= 1
            nxt[tuple(e2)] = nxt.get(tuple(e2), 0.0) + c * e[axis]
        terms = nxt
    return terms


def _poly_add(a, b, s=1.0):
    out = dict(a)
    for e, c in b.items():
        out[e] = out.get(e, 0.0) + s * c
    return out


def _lap(poly, dim):
    out = {}
    for e, c in poly.items():
        for ax in range(dim):
            out = _poly_add(out, _mono_deriv(e, c, ax, 2))
    return out


def _d(poly, ax, sign=1.0):
    out = {}
    for e, c in poly.items():
        out = _poly_add(out, _mono_deriv(e, c, ax, 1), sign)
    return out


def _eval(poly, x):
    return sum(c * np.prod([x[i] ** e[i] for i in range(len(e))]) for e, c in poly.items())


def _monomials(dim, deg):
    for e in itertools.product(range(deg + 1), repeat=dim):
        if sum(e) <= deg:
            yield e


def _desc(dim, L, G, B, n=8, h=0.5):
    st = Stencil2d2h(L, G, B) if dim == 2 else Stencil3d2h(L, G, B)
    return models.make_desc([n] * dim, st, h, 0.01, [TermSpec(LAP, 1.0, models.PSI)])


def _fill_poly(desc, mono, origin, h):
    """padded array holding the monomial at coordinates (p - origin)*h + x0"""
    shape = tuple(int(desc.n[a]) + 6 for a in reversed(range(desc.dim)))
    x0 = (0.3, -0.2, 0.1)[: desc.dim]
    grids = np.meshgrid(*[np.arange(s) for s in reversed(shape)], indexing="ij")
    coords = [(grids[a] - origin[a]) * h + x0[a] for a in range(desc.dim)]
    val = np.ones_like(coords[0], dtype=np.float64)
    for a in range(desc.dim):
        val = val * coords[a] ** mono[a]
    # meshgrid axes are (x, y[, z]); the padded array is (z, y, x)
    return np.ascontiguousarray(np.transpose(val)), np.array(x0)


def _direction(dim, axis, component):
    """unit vector of stride s0 (or s1) of get_stride<axis> (grid.h:537-578)"""
    if dim == 2:
        s = {0: [(1, 0), (0, 1)], 1: [(0, 1), (-1, 0)]}[axis]
    else:
        s = {0: [(1, 0, 0)], 1: [(0, 1, 0)], 2: [(0, 0, -1)]}[axis]
    return np.array(s[component], dtype=float)


# The reference's 2D 21-point bilaplacian (stencilh2.h:143-157) divides its
# weight sum by 12; a consistent bilaplacian with those weights needs 1/3, so
# the stencil returns a quarter of the bilaplacian (x^4 -> 6, x^2 y^2 -> 2).
# The oracle reproduces the reference.
QUIRK_SCALE = {(2, "bilap", 21): 0.25}


def _expected(op, dim, mono, x0, axis=0, component=0, L=None, h=None, B=None):
    poly = {tuple(mono): 1.0}
    if op == LAP:
        val = _eval(_lap(poly, dim), x0)
        if dim == 3 and L == 19:
            # stencilh2.h:273 sums "vz + vz": the -z neighbour is replaced by
            # +z, adding (2/6)*divh2*(f(x + h e_z) - f(x - h e_z))
            up, dn = x0.copy(), x0.copy()
            up[2] += h
            dn[2] -= h
            val += (2.0 / 6.0) / h ** 2 * (_eval(poly, up) - _eval(poly, dn))
        return val
    if op == BILAP:
        val = _eval(_lap(_lap(poly, dim), dim), x0)
        return val * QUIRK_SCALE.get((dim, "bilap", B), 1.0)
    u = _direction(dim, axis, component)
    lp = _lap(poly, dim)
    g = {}
    for a in range(dim):
        if u[a] != 0:
            g = _poly_add(g, _d(lp, a), u[a])
    return _eval(g, x0)


def _site_index(desc, site):
    X, Y = desc.n[0] + 6, desc.n[1] + 6
    if desc.dim == 2:
        return site[0] + X * site[1]
    return site[0] + X * (site[1] + Y * site[2])


def _check_exact(orc, dim, L, G, B, op, deg, axis=0, component=0):
    h = 0.5
    desc = _desc(dim, L, G, B, n=6, h=h)
    site = (5, 6, 7)[:dim]
    idx = _site_index(desc, site)
    for mono in _monomials(dim, deg):
        g, x0 = _fill_poly(desc, mono, site, h)
        got = orc.stencil_at(desc, op, axis, component, g, idx)
        want = _expected(op, dim, mono, x0, axis, component, L=L, h=h, B=B)
        scale = max(1.0, abs(want)) * (64.0 / h ** 4)
        assert abs(got - want) <= 1e-11 * scale, (dim, L, G, B, op, mono, got, want)


@pytest.mark.parametrize("L", LAPS_2D)
def test_lap2d_exact_to_cubics(orc, L):
    _check_exact(orc, 2, L, 6, 13, LAP, 3)


@pytest.mark.parametrize("B", BILAPS_2D)
def test_bilap2d_exact_to_quintics(orc, B):
    _check_exact(orc, 2, 5, 6, B, BILAP, 5)


@pytest.mark.parametrize("G", GRADS_2D)
@pytest.mark.parametrize("axis", (0, 1))
def test_gradlap2d_exact_to_quartics(orc, G, axis):
    comps = (0, 1) if G == 6 else (0,)
    for comp in comps:
        _check_exact(orc, 2, 5, G, 13, GRADLAP, 4, axis, comp)


@pytest.mark.parametrize("L", LAPS_3D)
def test_lap3d_exact_to_cubics(orc, L):
    _check_exact(orc, 3, L, 10, 25, LAP, 3)


@pytest.mark.parametrize("B", BILAPS_3D)
def test_bilap3d_exact_to_quintics(orc, B):
    _check_exact(orc, 3, 7, 10, B, BILAP, 5)


@pytest.mark.parametrize("G", GRADS_3D)
@pytest.mark.parametrize("axis", (0, 1, 2))
def test_gradlap3d_exact_to_quartics(orc, G, axis):
    _check_exact(orc, 3, 7, G, 25, GRADLAP, 4, axis, 0)


def test_stencils_not_exact_beyond_order(orc):
    """The exactness checks are sharp: lap5 of x^4 carries the O(h^2) error."""
    h = 0.5
    desc = _desc(2, 5, 6, 13, n=6, h=h)
    site = (5, 6)
    g, x0 = _fill_poly(desc, (4, 0), site, h)
    got = orc.stencil_at(desc, LAP, 0, 0, g, _site_index(desc, site))
    want = 12 * x0[0] ** 2
    assert abs(got - want - 2 * h * h) < 1e-12  # (f(x+h)+f(x-h)-2f)/h^2 = 12x^2 + 2h^2


def test_lap19_z_typo_value(orc):
    """lap19 applied to f = z gives 2/(3h) instead of 0 (stencilh2.h:273)."""
    h = 0.25
    desc = _desc(3, 19, 10, 25, n=6, h=h)
    site = (5, 5, 5)
    g, _ = _fill_poly(desc, (0, 0, 1), site, h)
    got = orc.stencil_at(desc, LAP, 0, 0, g, _site_index(desc, site))
    assert abs(got - 2.0 / (3.0 * h)) < 1e-12


def test_gradlap_z_sign(orc):
    """get_stride<Z> starts with -X*Y, so the Z gradlaplacian is -d/dz lap."""
    h = 0.5
    desc = _desc(3, 7, 12, 25, n=6, h=h)
    site = (5, 5, 5)
    g, _ = _fill_poly(desc, (0, 0, 3), site, h)
    got = orc.stencil_at(desc, GRADLAP, 2, 0, g, _site_index(desc, site))
    assert abs(got - (-6.0)) < 1e-10


# ----------------------------------------------------------- Fourier symbols
def _plane_wave_2d(nx, ny, mx, my, h):
    x = np.arange(nx) * h
    y = np.arange(ny) * h
    kx, ky = 2 * np.pi * mx / (nx * h), 2 * np.pi * my / (ny * h)
    return np.cos(kx * x[None, :] + ky * y[:, None] + 0.4), kx * h, ky * h


@pytest.mark.parametrize("h", (1.0, 0.5))
def test_fourier_symbols_2d(orc, h):
    nx, ny = 16, 12
    f, a, b = _plane_wave_2d(nx, ny, 3, 2, h)
    lam5 = (2 * math.cos(a) + 2 * math.cos(b) - 4) / h ** 2
    lam9 = (4 * math.cos(a) * math.cos(b) + 8 * (math.cos(a) + math.cos(b)) - 20) / (6 * h * h)
    for L, lam in ((5, lam5), (9, lam9)):
        desc = models.make_desc((nx, ny), Stencil2d2h(L, 6, 13), h, 0.01, [])
        g = orc.refresh(desc, orc.pad(desc, f))
        out = orc.unpad(desc, orc.apply(desc, LAP, 0, 0, g))
        np.testing.assert_allclose(out, lam * f, atol=1e-12 / h ** 2)
    desc = models.make_desc((nx, ny), Stencil2d2h(5, 6, 13), h, 0.01, [])
    g = orc.refresh(desc, orc.pad(desc, f))
    out = orc.unpad(desc, orc.apply(desc, BILAP, 0, 0, g))
    np.testing.assert_allclose(out, lam5 ** 2 * f, atol=1e-11 / h ** 4)


def test_fourier_symbols_3d(orc):
    n, h = (8, 6, 10), 0.5
    x = np.arange(n[0]) * h
    y = np.arange(n[1]) * h
    z = np.arange(n[2]) * h
    k = [2 * np.pi * m / (nn * h) for m, nn in zip((1, 2, 3), n)]
    f = np.cos(k[0] * x[None, None, :] + k[1] * y[None, :, None] + k[2] * z[:, None, None])
    lam7 = sum(2 * math.cos(kk * h) for kk in k) / h ** 2 - 6 / h ** 2
    desc = models.make_desc(n, Stencil3d2h(7, 10, 25), h, 0.01, [])
    g = orc.refresh(desc, orc.pad(desc, f))
    lap = orc.unpad(desc, orc.apply(desc, LAP, 0, 0, g))
    bil = orc.unpad(desc, orc.apply(desc, BILAP, 0, 0, g))
    np.testing.assert_allclose(lap, lam7 * f, atol=1e-11)
    np.testing.assert_allclose(bil, lam7 ** 2 * f, atol=1e-9)


# ----------------------------------------------------------- numpy cross-check
def _lap5_roll(f, h):
    return (np.roll(f, 1, 1) + np.roll(f, -1, 1) + np.roll(f, 1, 0) + np.roll(f, -1, 0) - 4 * f) / h ** 2


def _lap7_roll(f, h):
    s = sum(np.roll(f, d, a) for a in range(3) for d in (1, -1))
    return (s - 6 * f) / h ** 2


def test_model_b_rhs_matches_numpy(orc):
    from symphas_amd import inputs
    n, h = (20, 14), 0.7
    psi = inputs.uniform((n[1], n[0]), seed=11, lo=-0.8, hi=0.8)
    desc = models.model_b(n, h=h)
    g = orc.refresh(desc, orc.pad(desc, psi))
    rhs = orc.unpad(desc, orc.equation(desc, g))
    want = -_lap5_roll(_lap5_roll(psi, h), h) - _lap5_roll(psi, h) + _lap5_roll(psi ** 3, h)
    np.testing.assert_allclose(rhs, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_model_b_3d_rhs_matches_numpy(orc):
    from symphas_amd import inputs
    n, h = (6, 8, 7), 1.0
    psi = inputs.uniform((n[2], n[1], n[0]), seed=3, lo=-0.9, hi=0.9)
    desc = models.model_b(n, h=h, stencil=Stencil3d2h(7, 10, 25))
    g = orc.refresh(desc, orc.pad(desc, psi))
    rhs = orc.unpad(desc, orc.equation(desc, g))
    want = -_lap7_roll(_lap7_roll(psi, h), h) - _lap7_roll(psi, h) + _lap7_roll(psi ** 3, h)
    np.testing.assert_allclose(rhs, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_allen_cahn_rhs_matches_numpy(orc):
    from symphas_amd import inputs
    n, h = (16, 16), 0.5
    psi = inputs.uniform((16, 16), seed=5)
    desc = models.allen_cahn(n, h=h)
    g = orc.refresh(desc, orc.pad(desc, psi))
    rhs = orc.unpad(desc, orc.equation(desc, g))
    want = _lap5_roll(psi, h) + psi - psi ** 3
    np.testing.assert_allclose(rhs, want, rtol=0, atol=1e-13)


# ----------------------------------------------------------- time stepping
def test_single_mode_recursion(orc):
    nx, ny, h, dt = 16, 8, 0.5, 0.01
    f, a, b = _plane_wave_2d(nx, ny, 2, 1, h)
    lam = (2 * math.cos(a) + 2 * math.cos(b) - 4) / h ** 2
    desc = models.make_desc((nx, ny), Stencil2d2h(5, 6, 13), h, dt,
                            [TermSpec(LAP, 1.0, models.PSI)])
    out = orc.unpad(desc, orc.run(desc, f, 50))
    np.testing.assert_allclose(out, (1 + dt * lam) ** 50 * f, atol=1e-12)


def test_constant_allen_cahn_scalar_recursion(orc):
    c, dt = 0.3, 0.01
    desc = models.allen_cahn((8, 8), h=0.5, dt=dt)
    out = orc.unpad(desc, orc.run(desc, np.full((8, 8), c), 100))
    for _ in range(100):
        lap = 0.0  # lap5 of a constant is exactly 0
        rhs = lap + (1.0 * c + (-1.0) * (c * c * c))
        c = c + dt * rhs
    assert np.all(out == c)


def test_constant_model_b_is_fixed(orc):
    desc = models.model_b((8, 8))
    out = orc.unpad(desc, orc.run(desc, np.full((8, 8), -0.37), 50))
    assert np.all(out == -0.37)


def test_model_b_mass_conserved(orc):
    from symphas_amd import inputs
    desc = models.model_b((32, 32))
    psi = inputs.uniform((32, 32), seed=9)
    g0 = orc.refresh(desc, orc.pad(desc, psi))
    m0 = orc.mass(desc, g0)
    g = orc.run(desc, psi, 200)
    assert abs(orc.mass(desc, g) - m0) <= 1e-12 * np.abs(psi).sum()


def test_model_b_energy_decreases(orc):
    from symphas_amd import inputs
    desc = models.model_b((32, 32))
    psi = inputs.uniform((32, 32), seed=2, lo=-0.5, hi=0.5)
    g = orc.refresh(desc, orc.pad(desc, psi))
    e_prev = orc.energy(desc, g)
    for _ in range(10):
        g = orc.run(desc, orc.unpad(desc, g), 20)
        e = orc.energy(desc, g)
        assert e < e_prev
        e_prev = e


# ----------------------------------------------------------- free energy
def _energy_desc(dim, h):
    if dim == 2:
        return models.model_b((24, 20), h=h)
    return models.model_b((12, 10, 8), h=h, stencil=Stencil3d2h(7, 10, 25))


@pytest.mark.parametrize("dim,h", [(2, 0.5), (2, 1.3), (3, 0.5), (3, 1.3)])
@pytest.mark.parametrize("gradient", [0, 1])
def test_energy_constant_field_closed_form(orc, dim, h, gradient):
    """f(c) = -c1/2 c^2 + c2/4 c^4 (no gradient) times N h^D, or its mean."""
    desc = _energy_desc(dim, h)
    c, c1, c2 = -0.83, 0.7, 1.9
    shape = tuple(int(desc.n[a]) for a in reversed(range(dim)))
    g = orc.refresh(desc, orc.pad(desc, np.full(shape, c)))
    n = int(np.prod(shape))
    dens = -c1 * c * c / 2 + c2 * c ** 4 / 4
    e = orc.energy(desc, g, c1=c1, c2=c2, kappa=3.0, gradient=gradient)
    assert e == pytest.approx(n * h ** dim * dens, rel=1e-14)
    em = orc.energy(desc, g, c1=c1, c2=c2, kappa=3.0, gradient=gradient, mean=True)
    assert em == pytest.approx(dens, rel=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("gradient", [0, 1])
def test_energy_fourier_mode_closed_form(orc, dim, gradient):
    """psi = A cos(k.x): sum over sites of |grad psi|^2 is N A^2/2 times
    sum_a (2 - 2cos(k_a h))/h^2 (forward) or sin^2(k_a h)/h^2 (central)."""
    h = 0.7
    desc = _energy_desc(dim, h)
    n = [int(desc.n[a]) for a in range(dim)]
    modes = [3, 2, 1][:dim]
    k = [2 * math.pi * m / (n[a] * h) for a, m in enumerate(modes)]
    axes = np.meshgrid(*[np.arange(n[a]) * h for a in reversed(range(dim))], indexing="ij")
    phase = sum(k[a] * axes[dim - 1 - a] for a in range(dim))
    amp = 0.37
    psi = amp * np.cos(phase + 0.4)
    g = orc.refresh(desc, orc.pad(desc, psi))
    if gradient == 0:
        sym = sum((2 - 2 * math.cos(k[a] * h)) / h ** 2 for a in range(dim))
    else:
        sym = sum(math.sin(k[a] * h) ** 2 / h ** 2 for a in range(dim))
    N = int(np.prod(n))
    kappa = 0.6
    want = h ** dim * kappa / 2 * N * amp ** 2 / 2 * sym
    got = orc.energy(desc, g, c1=0.0, c2=0.0, kappa=kappa, gradient=gradient)
    assert got == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_energy_summation_by_parts(orc, dim):
    """Forward differences: sum |grad+ psi|^2 = -sum psi * lap psi on the
    periodic grid, with lap the (pinned) 5-point / 7-point stencil."""
    from symphas_amd import inputs
    h = 0.8
    desc = _energy_desc(dim, h)
    shape = tuple(int(desc.n[a]) for a in reversed(range(dim)))
    psi = inputs.uniform(shape, seed=41, lo=-1.0, hi=1.0)
    g = orc.refresh(desc, orc.pad(desc, psi))
    lap = orc.unpad(desc, orc.apply(desc, LAP, 0, 0, g))
    rhs = -np.sum(psi * lap)
    got = orc.energy(desc, g, c1=0.0, c2=0.0, kappa=2.0) / h ** dim
    assert got == pytest.approx(rhs, rel=1e-12)
    # the central gradient is not the one summation by parts pairs with lap5
    cen = orc.energy(desc, g, c1=0.0, c2=0.0, kappa=2.0, gradient=1) / h ** dim
    assert abs(cen - rhs) > 1e-3 * abs(rhs)


@pytest.mark.parametrize("gradient", [0, 1])
def test_energy_linear_in_coefficients(orc, gradient):
    from symphas_amd import inputs
    desc = _energy_desc(2, 1.1)
    psi = inputs.uniform((20, 24), seed=42, lo=-1.2, hi=1.2)
    g = orc.refresh(desc, orc.pad(desc, psi))
    e = lambda c1, c2, k: orc.energy(desc, g, c1=c1, c2=c2, kappa=k, gradient=gradient)
    parts = e(1, 0, 0), e(0, 1, 0), e(0, 0, 1)
    assert e(0.3, 1.7, 2.5) == pytest.approx(0.3 * parts[0] + 1.7 * parts[1] + 2.5 * parts[2],
                                             rel=1e-13)
    # the default functional is c1 = c2 = kappa = 1, forward, integral
    if gradient == 0:
        assert orc.energy(desc, g) == e(1, 1, 1)


def test_mass_is_sequential_sum(orc):
    desc = models.model_b((8, 6))
    psi = np.arange(48, dtype=float).reshape(6, 8) * 0.1
    assert orc.mass(desc, orc.pad(desc, psi)) == sum(psi.ravel().tolist())


# ----------------------------------------------------------- ghosts
@pytest.mark.parametrize("n", [(7, 9), (3, 3), (16, 5)])
def test_refresh_2d_is_modular(orc, n):
    desc = models.model_b(n)
    rng = np.random.default_rng(1)
    g = rng.standard_normal((n[1] + 6, n[0] + 6))
    a = orc.refresh(desc, g.copy())
    b = orc.refresh(desc, g.copy(), modular=True)
    assert np.array_equal(a, b)
    assert np.array_equal(orc.refresh(desc, a.copy()), a)  # idempotent
    assert np.array_equal(orc.unpad(desc, a), orc.unpad(desc, g))  # interior untouched


def test_refresh_3d_matches_wrap(orc):
    n = (4, 5, 6)
    desc = models.model_b(n, stencil=Stencil3d2h(7, 10, 25))
    rng = np.random.default_rng(2)
    inner = rng.standard_normal((6, 5, 4))
    g = orc.refresh(desc, orc.pad(desc, inner))
    want = np.pad(inner, 3, mode="wrap")
    assert np.array_equal(g, want)


def test_check_rejects_bad_descriptors(orc):
    d = models.model_b((8, 8))
    assert orc.check(d) == 0
    d.lap_points = 7
    assert orc.check(d) != 0
    d = models.model_b((2, 8))
    assert orc.check(d) != 0


def test_generator_known_answer(orc):
    from symphas_amd import inputs
    # SplitMix64, seed 0: published first outputs
    z = inputs.splitmix64(0, 2)
    assert int(z[0]) == 0xE220A8397B1DCDAF
    assert int(z[1]) == 0x6E789E6AA1B965F4
    a = inputs.uniform((1000,), seed=123, lo=-0.1, hi=0.1)
    b = orc.fill_uniform(1000, 123, -0.1, 0.1)
    assert np.array_equal(a, b)
    assert a.min() >= -0.1 and a.max() < 0.1


# ----------------------------------------------------------- provisional variables
@pytest.mark.parametrize("stencil,n", [(Stencil2d2h(5, 6, 13), (18, 12)),
                                       (Stencil2d2h(9, 6, 17), (11, 9)),
                                       (Stencil3d2h(7, 10, 25), (7, 6, 9))])
def test_mbb_provisional_equals_state_speed_bitwise(orc, stencil, n):
    """MBB evaluates var(1) = -(c1 - 4 c2 psi^2) psi into its own grid and
    refreshes its ghosts; STATE_SPEED differentiates the same polynomial through
    a temporary grid over the padded array.  The values the laplacian reads
    are the same numbers, so every step agrees bit for bit."""
    from symphas_amd import inputs
    psi = inputs.uniform(tuple(reversed(n)), seed=31, lo=-0.7, hi=0.7)
    prov = models.model_mbb(n, h=0.8, dt=0.002, stencil=stencil, c1=1.3, c2=0.6)
    flat = models.state_speed(n, h=0.8, dt=0.002, stencil=stencil, a=1.3, b=0.6)
    assert np.array_equal(orc.run(prov, psi, 7), orc.run(flat, psi, 7))


def test_provisional_grid_has_periodic_ghosts(orc):
    from symphas_amd import inputs
    n = (9, 7)
    psi = inputs.uniform((7, 9), seed=32, lo=-0.9, hi=0.9)
    desc = models.model_bb(n, h=0.5, a=0.8, b=0.3)
    g = orc.refresh(desc, orc.pad(desc, psi))
    var = orc.provisional(desc, g, 1)
    inner = orc.unpad(desc, var)
    want = _lap5_roll(psi, 0.5) + (0.8 - 4 * 0.3 * psi * psi) * psi
    np.testing.assert_allclose(inner, want, rtol=0, atol=1e-12 * np.abs(want).max())
    np.testing.assert_array_equal(var, np.pad(inner, 3, mode="wrap"))


@pytest.mark.parametrize("L,h", [(5, 1.0), (9, 0.5)])
def test_bb_single_mode_recursion(orc, L, h):
    """BB with B = 0 is linear: var = lap psi + a psi, dpsi = -lap var, so a
    plane wave with laplacian symbol lam follows (1 - dt (lam^2 + a lam))^n."""
    nx, ny, dt, a = 16, 12, 0.002, 0.7
    f, kx, ky = _plane_wave_2d(nx, ny, 2, 1, h)
    if L == 5:
        lam = (2 * math.cos(kx) + 2 * math.cos(ky) - 4) / h ** 2
    else:
        lam = (4 * math.cos(kx) * math.cos(ky) + 8 * (math.cos(kx) + math.cos(ky)) - 20) / (6 * h * h)
    desc = models.model_bb((nx, ny), h=h, dt=dt, stencil=Stencil2d2h(L, 6, 13), a=a, b=0.0)
    out = orc.unpad(desc, orc.run(desc, f, 40))
    np.testing.assert_allclose(out, (1 - dt * (lam * lam + a * lam)) ** 40 * f, atol=1e-11)


def test_bb_matches_numpy_and_model_b(orc):
    """BB against an np.roll restatement of its two passes, and against Model B
    written with the 13-point bilaplacian (= lap5 of lap5 exactly, so the two
    differ by rounding only)."""
    from symphas_amd import inputs
    n, h, dt = (14, 10), 0.9, 0.004
    psi = inputs.uniform((10, 14), seed=33, lo=-0.8, hi=0.8)
    desc = models.model_bb(n, h=h, dt=dt, a=1.0, b=0.25)
    g = orc.refresh(desc, orc.pad(desc, psi))
    rhs = orc.unpad(desc, orc.equation(desc, g))
    var = _lap5_roll(psi, h) + (1.0 - psi * psi) * psi
    want = -_lap5_roll(var, h)
    np.testing.assert_allclose(rhs, want, rtol=0, atol=1e-12 * np.abs(want).max())
    mb = models.model_b(n, h=h, dt=dt)
    a = orc.unpad(desc, orc.run(desc, psi, 20))
    b = orc.unpad(mb, orc.run(mb, psi, 20))
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)


def test_bb_3d_matches_numpy(orc):
    from symphas_amd import inputs
    n, h = (6, 7, 5), 1.0
    psi = inputs.uniform((5, 7, 6), seed=34, lo=-0.8, hi=0.8)
    desc = models.model_bb(n, h=h, stencil=Stencil3d2h(7, 10, 25), a=0.9, b=0.2)
    g = orc.refresh(desc, orc.pad(desc, psi))
    rhs = orc.unpad(desc, orc.equation(desc, g))
    want = -_lap7_roll(_lap7_roll(psi, h) + (0.9 - 0.8 * psi * psi) * psi, h)
    np.testing.assert_allclose(rhs, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_check_rejects_bad_provisionals(orc):
    d = models.model_bb((8, 8))
    assert orc.check(d) == 0
    d.terms[0].operand = 2  # var(2) is not defined
    assert orc.check(d) != 0
    d = models.model_bb((8, 8))
    d.prov[0][0].operand = 1  # var(1) may not read itself
    assert orc.check(d) != 0
    d = models.model_bb((8, 8))
    d.prov_nterms[0] = 0
    assert orc.check(d) != 0


# ----------------------------------------------------------- coupled fields
MT = models.MTermSpec


def test_model_c_without_rho_is_model_b_bitwise(orc):
    from symphas_amd import inputs
    n = (20, 14)
    psi = inputs.uniform((14, 20), seed=61, lo=-0.6, hi=0.6)
    mc = models.model_c(n, h=0.9, dt=0.003, c=(1.2, 0.7, 0.4, 0.9, 0.3))
    mb = models.model_b(n, h=0.9, dt=0.003, c1=1.2, c2=0.7)
    got = orc.mrun(mc, [psi, np.zeros_like(psi)], 9)
    want = orc.unpad(mb, orc.run(mb, psi, 9))
    assert np.array_equal(got[0], want)
    assert np.array_equal(got[1], np.zeros_like(psi))


def test_coupled_linear_mode_recursion(orc):
    """dpsi = a lap psi + b rho, drho = c psi + e lap rho: a plane wave with
    laplacian symbol lam follows (I + dt [[a lam, b], [c, e lam]])^n."""
    nx, ny, h, dt = 16, 12, 0.5, 0.004
    f, kx, ky = _plane_wave_2d(nx, ny, 2, 1, h)
    lam = (2 * math.cos(kx) + 2 * math.cos(ky) - 4) / h ** 2
    a, b, c, e = 0.8, -0.3, 0.5, 1.4
    eqs = [[MT(LAP, a, (1, 0)), MT(abi.OP_ID, b, (0, 1))],
           [MT(abi.OP_ID, c, (1, 0)), MT(LAP, e, (0, 1))]]
    d = models.make_mdesc((nx, ny), Stencil2d2h(5, 6, 13), h, dt, eqs)
    got = orc.mrun(d, [f, 0.5 * f], 30)
    A = np.linalg.matrix_power(np.eye(2) + dt * np.array([[a * lam, b], [c, e * lam]]), 30)
    want = A @ np.array([1.0, 0.5])
    np.testing.assert_allclose(got[0], want[0] * f, atol=1e-12)
    np.testing.assert_allclose(got[1], want[1] * f, atol=1e-12)


def _rhs_roll(name, psi, rho, h, c, lapf):
    if name == "model_c":
        c1, c2, c3, c4, c5 = c
        return (-lapf(lapf(psi, h), h) - lapf((c1 - c2 * psi * psi) * psi + c5 * rho * rho, h),
                lapf(rho, h) + (c3 - c4 * rho * rho) * rho + 2 * c5 * psi * rho)
    if name == "gray_scott":
        c1, c2, c3, c4 = c
        prr = psi * rho * rho
        return (c1 * lapf(psi, h) - prr + c3 * (1 - psi),
                c2 * lapf(rho, h) + prr - (c3 + c4) * rho)
    if name == "mixed":
        k0, k1 = c
        return (lapf(psi ** 3 - psi + 0.3 * psi * rho + 0.2 * psi * psi * rho
                     + 0.15 * psi * rho * rho, h) - k0 * lapf(lapf(psi, h), h)
                + 0.1 * psi - 0.05 * psi * rho ** 3,
                lapf(rho ** 3 - rho + 0.2 * psi * psi + 0.1 * psi * rho + 0.1 * rho * rho, h)
                - k1 * lapf(lapf(rho, h), h) - 0.1 * rho + 0.05 * psi * psi * rho * rho)
    c1, c2, c3, c4, c5, c6 = c
    return (c1 * lapf(psi, h) + c2 * (psi + c3 * rho - psi * rho * rho - c4 * psi * rho),
            lapf(rho, h) + c2 * (c5 * rho + c6 * psi + psi * rho * rho + c4 * psi * rho))


@pytest.mark.parametrize("name,c", [("model_c", (1.1, 0.8, 0.6, 1.3, 0.4)),
                                    ("gray_scott", (0.16, 0.08, 0.035, 0.065)),
                                    ("turing", (1.0, 0.7, 1.2, 0.5, -0.4, 0.9)),
                                    ("mixed", (0.6, 0.8))])
def test_coupled_rhs_matches_numpy(orc, name, c):
    from symphas_amd import inputs
    n, h = (18, 13), 0.8
    psi = inputs.uniform((13, 18), seed=62, lo=-0.9, hi=0.9)
    rho = inputs.uniform((13, 18), seed=63, lo=-0.9, hi=0.9)
    d = getattr(models, name)(n, h=h, dt=0.01, c=c)
    got = orc.mequation(d, [psi, rho])
    want = _rhs_roll(name, psi, rho, h, c, _lap5_roll)
    for g, w in zip(got, want):
        np.testing.assert_allclose(g, w, rtol=0, atol=1e-12 * np.abs(w).max())


def test_model_c_3d_rhs_matches_numpy(orc):
    from symphas_amd import inputs
    n, h, c = (6, 7, 5), 1.0, (1.0, 1.0, 0.5, 0.8, 0.3)
    psi = inputs.uniform((5, 7, 6), seed=64, lo=-0.8, hi=0.8)
    rho = inputs.uniform((5, 7, 6), seed=65, lo=-0.8, hi=0.8)
    d = models.model_c(n, h=h, stencil=Stencil3d2h(7, 10, 25), c=c)
    got = orc.mequation(d, [psi, rho])
    want = _rhs_roll("model_c", psi, rho, h, c, _lap7_roll)
    for g, w in zip(got, want):
        np.testing.assert_allclose(g, w, rtol=0, atol=1e-12 * np.abs(w).max())


def test_mixed_3d_rhs_matches_numpy(orc):
    from symphas_amd import inputs
    n, h, c = (6, 7, 5), 0.9, (0.6, 0.8)
    psi = inputs.uniform((5, 7, 6), seed=66, lo=-0.8, hi=0.8)
    rho = inputs.uniform((5, 7, 6), seed=67, lo=-0.8, hi=0.8)
    d = models.mixed(n, h=h, stencil=Stencil3d2h(7, 10, 25), c=c)
    got = orc.mequation(d, [psi, rho])
    want = _rhs_roll("mixed", psi, rho, h, c, _lap7_roll)
    for g, w in zip(got, want):
        np.testing.assert_allclose(g, w, rtol=0, atol=1e-12 * np.abs(w).max())


def test_gray_scott_homogeneous_state_is_fixed(orc):
    d = models.gray_scott((8, 8))
    got = orc.mrun(d, [np.ones((8, 8)), np.zeros((8, 8))], 25)
    assert np.array_equal(got[0], np.ones((8, 8))) and np.array_equal(got[1], np.zeros((8, 8)))


def test_mcheck_rejects_bad_descriptors(orc):
    d = models.gray_scott((8, 8))
    assert orc.mcheck(d) == 0
    d.terms[0][1].pow[2] = 1  # a third field the model does not have
    assert orc.mcheck(d) != 0
    d = models.gray_scott((8, 8))
    d.nfields = 5
    assert orc.mcheck(d) != 0
