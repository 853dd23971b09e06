"""Pins of the oracle's regional-grid restatement (SolverSystemFDwSD).

No reference test or fixture covers the regional grid, and the reference
does not build here, so these pins come from what the rules fix by
themselves:
  - the covering interval equals a brute-force search over every periodic
    interval on small circles;
  - a box that spans the whole domain reproduces the plain periodic run bit
    for bit (the regional layer is then the periodic ghost layer);
  - a blob on an exactly stationary background (0 for Model B) matches the
    plain run bit for bit for as many steps as its dependency cone stays
    inside the box interior, and the exterior stays exactly `empty` for any
    number of steps;
  - shifting the initial field periodically shifts the result and the box
    origin by the same amount;
  - a resize leaves every active site inside the box, centred to within a
    site, with the dims the padding factor asks for.
Plausible mistakes each break one: an off-by-one in the wrap, a layer that
is not refreshed on a spanning axis, a move that shifts values, a resize
that keeps the old origin.
"""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as po
from symphas_amd import abi, inputs, models


def brute_cover(occ):
    n = len(occ)
    idx = [i for i in range(n) if occ[i]]
    if not idx:
        return (0, 0)
    best = None
    for L in range(1, n + 1):
        for s in range(n):
            cov = {(s + k) % n for k in range(L)}
            if all(i in cov for i in idx):
                gap_start = (s + L) % n if L < n else 0
                key = (L, gap_start if L < n else 0, s)
                if best is None or key < best[0]:
                    best = (key, (s if L < n else 0, L))
        if best is not None:
            return best[1]
    return best[1]


def test_cover_interval_matches_brute_force():
    rng = np.random.default_rng(5)
    for n in range(1, 13):
        for _ in range(60):
            occ = (rng.random(n) < rng.random()).astype(np.uint8)
            assert po.cover_interval(occ) == brute_cover(occ), (n, occ.tolist())
    # every pattern on a circle of 7
    for bits in itertools.product([0, 1], repeat=7):
        assert po.cover_interval(bits) == brute_cover(bits), bits


def _desc2(shape=(48, 40), dt=0.01):
    return models.model_b(shape, h=1.0, dt=dt, dtype=abi.F64)


def _desc3(shape=(20, 18, 16)):
    return models.model_b(shape, h=1.0, dt=0.005, dtype=abi.F64,
                          stencil=models.Stencil3d2h(7, 10, 25))


@pytest.mark.parametrize("kind", ["mb2", "ac2", "mb3"])
def test_spanning_box_equals_plain_run(kind):
    if kind == "mb2":
        d = _desc2()
    elif kind == "ac2":
        d = models.allen_cahn((40, 36), h=0.5, dt=0.01, dtype=abi.F64)
    else:
        d = _desc3()
    psi = inputs.uniform(po.interior_shape(d), seed=11)
    rp = po.regional_defaults()
    rp.resize_factor = 50.0  # every axis spans the domain
    rp.resize_time = 0.03     # resizes along the way keep it spanning
    got, st = po.regional_run(d, rp, psi, 40)
    want = po.unpad(d, po.run(d, psi, 40))
    assert np.array_equal(got, want)
    for a in range(d.dim):
        assert st.dims[a] == d.n[a] + 6


def _blob(d, value=0.5, lo=(18, 14, 6), size=(10, 7, 5)):
    psi = np.zeros(po.interior_shape(d))
    if d.dim == 2:
        psi[lo[1]:lo[1] + size[1], lo[0]:lo[0] + size[0]] = value
    else:
        psi[lo[2]:lo[2] + size[2], lo[1]:lo[1] + size[1], lo[0]:lo[0] + size[0]] = value
    return psi


@pytest.mark.parametrize("dim", [2, 3])
def test_blob_matches_plain_run_inside_its_cone(dim):
    if dim == 2:
        d, lo, size = _desc2(), (18, 14), (10, 7)
    else:
        d, lo, size = _desc3((32, 28, 24)), (12, 10, 8), (6, 5, 4)
    psi = _blob(d, lo=lo + (0,), size=size + (1,))
    _, st0 = po.regional_run(d, po.regional_defaults(), psi, 1)
    # steps whose radius-2 cone stays inside the box interior: the plain run
    # is then exactly 0 on the layer and outside the box, as the regional
    # run is
    margin = min(min((lo[a] - st0.origin[a]) % d.n[a],
                     (st0.origin[a] + st0.dims[a] - 6 - lo[a] - size[a]) % d.n[a])
                 for a in range(dim))
    k = margin // 2
    assert k >= 2
    got, _ = po.regional_run(d, po.regional_defaults(), psi, k)
    want = po.unpad(d, po.run(d, psi, k))
    assert np.array_equal(got, want)
    got2, _ = po.regional_run(d, po.regional_defaults(), psi, k + 2)
    want2 = po.unpad(d, po.run(d, psi, k + 2))
    assert not np.array_equal(got2, want2)  # the frozen layer now matters


def _outside_mask(d, st):
    shape = po.interior_shape(d)
    inside = np.ones(shape, dtype=bool)
    for a in range(d.dim):
        ax = d.dim - 1 - a
        c = np.arange(d.n[a])
        k = (c - st.origin[a]) % d.n[a]
        ins = k < st.dims[a] - 6
        sh = [1] * d.dim
        sh[ax] = d.n[a]
        inside &= ins.reshape(sh)
    return ~inside


@pytest.mark.parametrize("dim", [2, 3])
def test_exterior_stays_empty(dim):
    d = _desc2() if dim == 2 else _desc3()
    # empty = min = -0.25; the exterior is stored as `empty`, whatever the
    # plain dynamics would do there
    psi = _blob(d, lo=(8, 8, 6), size=(6, 5, 4)) - 0.25
    got, st = po.regional_run(d, po.regional_defaults(), psi, 30)
    assert st.empty == -0.25
    out = _outside_mask(d, st)
    assert out.any()
    assert np.all(got[out] == -0.25)


def test_shift_covariance():
    d = _desc2((40, 36))
    psi = _blob(d, lo=(9, 11), size=(8, 5))
    psi += 0.01 * inputs.uniform(psi.shape, seed=3, lo=0.0, hi=1.0) * (psi > 0)
    rp = po.regional_defaults()
    rp.resize_time = 0.05
    rp.fixed_resize = 0
    base, st = po.regional_run(d, rp, psi, 25)
    for sx, sy in [(5, 0), (0, 7), (33, 29), (-4, 12)]:
        moved, st2 = po.regional_run(d, rp, np.roll(psi, (sy, sx), axis=(0, 1)), 25)
        assert np.array_equal(moved, np.roll(base, (sy, sx), axis=(0, 1)))
        assert st2.origin[0] == (st.origin[0] + sx) % d.n[0]
        assert st2.origin[1] == (st.origin[1] + sy) % d.n[1]
        assert list(st2.dims) == list(st.dims)


def test_resize_places_a_centred_box():
    d = _desc2((64, 56))
    psi = _blob(d, lo=(20, 22), size=(11, 6))
    rp = po.regional_defaults()
    for steps in (1, 7):
        got, st = po.regional_run(d, rp, psi, steps)
        # the first update's box: factor 2 around the active box (11 x 6)
        for a, L in enumerate((11, 6)):
            dims = L + 6
            assert st.dims[a] == min(d.n[a] + 6, int(dims * rp.resize_factor))
            inner = st.dims[a] - 6
            lo = (20, 22)[a]
            left = (lo - st.origin[a]) % d.n[a]
            right = inner - left - L
            assert left >= 0 and right >= 0 and abs(left - right) <= 1
    # a non-fixed resize later tracks the grown active set
    rp.fixed_resize = 0
    rp.resize_time = 0.5
    got, st = po.regional_run(d, rp, psi, 120)
    active = got > st.cutoff
    assert not np.any(active & _outside_mask(d, st))


def test_no_active_site_gives_empty_box():
    d = _desc2((24, 20))
    psi = np.full(po.interior_shape(d), 0.3)
    got, st = po.regional_run(d, po.regional_defaults(), psi, 3)
    assert list(st.dims[:2]) == [0, 0]
    assert np.all(got == 0.3)
