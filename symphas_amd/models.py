"""Problem descriptors for the SymPhas models on the solver path.

Each builder returns the canonical term list that SolverFT::form_expr_one
produces for the reference's EVOLUTION expression (examples/solvers/
solverft.h:112-119): derivatives distributed over sums
(sym/inc/expressionapply.h:1476-1488), one term per derivative, a derivative
of a non-linear operand taking that operand as a polynomial in psi.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from typing import Sequence

from . import abi


@dataclass(frozen=True)
class Stencil2d2h:
    """Stencil2d2h<L, G, B> (sol/inc/stencilincludes.h:315-356).

    The reference's default Stencil2d2h<> is <9, 6, 13> under the default
    build option AVAILABLE_STENCILS_BASIC_ONLY (sol/inc/stencilincludes.h:
    894-900, sol/CMakeLists.txt:73,100-106)."""
    L: int = 9
    G: int = 6
    B: int = 13
    dim: int = 2


@dataclass(frozen=True)
class Stencil3d2h:
    """Stencil3d2h<L, G, B> (sol/inc/stencilincludes.h:368-409); default
    <15, 10, 25> (:784-789, :913-917)."""
    L: int = 15
    G: int = 10
    B: int = 25
    dim: int = 3


@dataclass(frozen=True)
class TermSpec:
    op: int
    coeff: float
    poly: Sequence[float]  # poly[k] multiplies u^k
    axis: int = 0
    component: int = 0
    operand: int = 0       # u: 0 the field psi, k >= 1 provisional var(k)


PSI = (0.0, 1.0)
PSI3 = (0.0, 0.0, 0.0, 1.0)


def make_desc(shape: Sequence[int], stencil, h: float, dt: float,
              terms: Sequence[TermSpec], dtype: int = abi.F64,
              path: int = abi.PATH_AUTO,
              provisionals: Sequence[Sequence[TermSpec]] = ()) -> abi.Desc:
    """shape is the interior extent (nx, ny[, nz]), x fastest.  provisionals[k-1]
    lists the terms defining var(k) (PROVISIONAL_DEF)."""
    if len(shape) != stencil.dim:
        raise ValueError(f"shape {tuple(shape)} does not match a {stencil.dim}D stencil")
    if len(terms) > abi.MAX_TERMS:
        raise ValueError(f"at most {abi.MAX_TERMS} terms")
    if len(provisionals) > abi.MAX_PROV:
        raise ValueError(f"at most {abi.MAX_PROV} provisional variables")
    d = abi.Desc()
    d.dim = stencil.dim
    d.dtype = dtype
    for a in range(3):
        d.n[a] = int(shape[a]) if a < len(shape) else 1
    d.lap_points, d.gradlap_points, d.bilap_points = stencil.L, stencil.G, stencil.B
    d.nterms = len(terms)
    d.path = path
    d.h = float(h)
    d.dt = float(dt)

    def fill(dst, t):
        if len(t.poly) > abi.MAX_DEGREE + 1:
            raise ValueError("operand degree above 6")
        dst.op, dst.axis, dst.component, dst.operand = t.op, t.axis, t.component, t.operand
        dst.coeff = float(t.coeff)
        for k, c in enumerate(t.poly):
            dst.poly[k] = float(c)

    for i, t in enumerate(terms):
        fill(d.terms[i], t)
    d.nprov = len(provisionals)
    for k, defn in enumerate(provisionals):
        if not 1 <= len(defn) <= abi.MAX_TERMS:
            raise ValueError(f"var({k + 1}) needs 1..{abi.MAX_TERMS} terms")
        d.prov_nterms[k] = len(defn)
        for i, t in enumerate(defn):
            fill(d.prov[k][i], t)
    return d


def allen_cahn_terms(mobility: float = 1.0, kappa: float = 1.0):
    """dpsi = M*(kappa*lap(psi) + psi - psi^3).

    With M = kappa = 1 this is the README model
    dop(1) = lap(op(1)) - (power(op(1), 2) - 1_n) * op(1) (README.md:83-85):
    a laplacian term and a pointwise term."""
    return [TermSpec(abi.OP_LAP, mobility * kappa, PSI),
            TermSpec(abi.OP_ID, 1.0, (0.0, mobility, 0.0, -mobility))]


def model_a_terms(c1: float = 1.0, c2: float = 1.0):
    """MA: dpsi = lap(psi) + (c1 - c2*psi*psi)*psi
    (examples/models/modeldefinitions.h:48)."""
    return [TermSpec(abi.OP_LAP, 1.0, PSI),
            TermSpec(abi.OP_ID, 1.0, (0.0, c1, 0.0, -c2))]


def model_b_terms(c1: float = 1.0, c2: float = 1.0, mobility: float = 1.0,
                  kappa: float = 1.0):
    """MB: dpsi = -bilap(psi) - lap((c1 - c2*psi*psi)*psi)
    (examples/models/modeldefinitions.h:53-55), i.e.
    -bilap(psi) - c1*lap(psi) + c2*lap(psi^3) after distribution; with a
    mobility M and gradient coefficient kappa,
    dpsi = M*lap(c2*psi^3 - c1*psi - kappa*lap(psi))."""
    return [TermSpec(abi.OP_BILAP, -mobility * kappa, PSI),
            TermSpec(abi.OP_LAP, -mobility * c1, PSI),
            TermSpec(abi.OP_LAP, mobility * c2, PSI3)]


def allen_cahn(shape, h=0.5, dt=0.01, stencil=Stencil2d2h(5, 6, 13), dtype=abi.F64,
               path=abi.PATH_AUTO, mobility=1.0, kappa=1.0) -> abi.Desc:
    d = make_desc(shape, stencil, h, dt, allen_cahn_terms(mobility, kappa), dtype, path)
    d.landau = (1.0, 1.0, float(kappa))
    return d


def model_b(shape, h=1.0, dt=0.01, stencil=Stencil2d2h(5, 6, 13), dtype=abi.F64,
            path=abi.PATH_AUTO, mobility=1.0, kappa=1.0, c1=1.0, c2=1.0) -> abi.Desc:
    d = make_desc(shape, stencil, h, dt, model_b_terms(c1, c2, mobility, kappa),
                  dtype, path)
    d.landau = (float(c1), float(c2), float(kappa))
    return d


def landau_energy(desc=None, gradient: int = abi.GRAD_FORWARD, mean: bool = False,
                  c1=None, c2=None, kappa=None) -> "abi.Energy":
    """The free-energy functional -c1/2 psi^2 + c2/4 psi^4 + kappa/2 |grad psi|^2
    whose gradient flow a Model B / Allen-Cahn descriptor made here follows
    (mu = -c1 psi + c2 psi^3 - kappa lap psi); explicit values override the
    descriptor's, and a descriptor from elsewhere defaults to 1, 1, 1."""
    base = getattr(desc, "landau", (1.0, 1.0, 1.0)) if desc is not None else (1.0, 1.0, 1.0)
    return abi.Energy(base[0] if c1 is None else c1, base[1] if c2 is None else c2,
                      base[2] if kappa is None else kappa, int(gradient), int(bool(mean)))


# ------------------------------------------------- provisional variables

def mbb_terms(c1: float = 1.0, c2: float = 1.0):
    """MBB (examples/models/modeldefinitions.h:150-155):
        var(1) <= -(c(1) - 4*c(2)*psi*psi)*psi
        dpsi = -bilap(psi) + lap(var(1))
    Returns (evolution terms, [definition of var(1)])."""
    var1 = [TermSpec(abi.OP_ID, 1.0, (0.0, -c1, 0.0, 4.0 * c2))]
    evo = [TermSpec(abi.OP_BILAP, -1.0, PSI), TermSpec(abi.OP_LAP, 1.0, PSI, operand=1)]
    return evo, [var1]


def state_speed_terms(a: float = 1.0, b: float = 1.0):
    """STATE_SPEED (tests/testdefs.h:23-26), MBB's right-hand side written
    without a provisional variable:
        dpsi = -bilap(psi) + lap(-(A - 4*B*psi*psi)*psi)"""
    return [TermSpec(abi.OP_BILAP, -1.0, PSI),
            TermSpec(abi.OP_LAP, 1.0, (0.0, -a, 0.0, 4.0 * b))]


def bb_terms(a: float = 1.0, b: float = 1.0):
    """BB (tests/testdefs.h:29-37), Model B with the laplacian moved into the
    provisional variable so that no bilaplacian stencil is used:
        var(1) = lap(psi) + (A - 4*B*psi*psi)*psi
        dpsi = -lap(var(1))
    Returns (evolution terms, [definition of var(1)])."""
    var1 = [TermSpec(abi.OP_LAP, 1.0, PSI), TermSpec(abi.OP_ID, 1.0, (0.0, a, 0.0, -4.0 * b))]
    evo = [TermSpec(abi.OP_LAP, -1.0, PSI, operand=1)]
    return evo, [var1]


def model_mbb(shape, h=1.0, dt=0.01, stencil=Stencil2d2h(5, 6, 13), dtype=abi.F64,
              path=abi.PATH_AUTO, c1=1.0, c2=1.0) -> abi.Desc:
    evo, prov = mbb_terms(c1, c2)
    return make_desc(shape, stencil, h, dt, evo, dtype, path, prov)


def model_bb(shape, h=1.0, dt=0.01, stencil=Stencil2d2h(5, 6, 13), dtype=abi.F64,
             path=abi.PATH_AUTO, a=1.0, b=1.0) -> abi.Desc:
    evo, prov = bb_terms(a, b)
    return make_desc(shape, stencil, h, dt, evo, dtype, path, prov)


def state_speed(shape, h=1.0, dt=0.01, stencil=Stencil2d2h(5, 6, 13), dtype=abi.F64,
                path=abi.PATH_AUTO, a=1.0, b=1.0) -> abi.Desc:
    return make_desc(shape, stencil, h, dt, state_speed_terms(a, b), dtype, path)


# ------------------------------------------------------- coupled fields

@dataclass
class MTermSpec:
    """coeff * D[prod_f psi_f^pow[f]] in the equation of one field."""
    op: int
    coeff: float
    pow: Sequence[int]
    axis: int = 0
    component: int = 0


def make_mdesc(shape: Sequence[int], stencil, h: float, dt: float,
               equations: Sequence[Sequence[MTermSpec]], dtype: int = abi.F64,
               path: int = abi.PATH_AUTO) -> abi.MDesc:
    """equations[f] lists the terms of d(field f)/dt."""
    if len(shape) != stencil.dim:
        raise ValueError(f"shape {tuple(shape)} does not match a {stencil.dim}D stencil")
    if not 1 <= len(equations) <= abi.MAX_FIELDS:
        raise ValueError(f"1..{abi.MAX_FIELDS} fields")
    d = abi.MDesc()
    d.dim = stencil.dim
    d.dtype = dtype
    for a in range(3):
        d.n[a] = int(shape[a]) if a < len(shape) else 1
    d.lap_points, d.gradlap_points, d.bilap_points = stencil.L, stencil.G, stencil.B
    d.nfields = len(equations)
    d.path = path
    d.h, d.dt = float(h), float(dt)
    for f, eq in enumerate(equations):
        if len(eq) > abi.MAX_MTERMS:
            raise ValueError(f"at most {abi.MAX_MTERMS} terms per field")
        d.nterms[f] = len(eq)
        for i, t in enumerate(eq):
            tt = d.terms[f][i]
            tt.op, tt.axis, tt.component, tt.coeff = t.op, t.axis, t.component, float(t.coeff)
            for g, p in enumerate(t.pow):
                tt.pow[g] = int(p)
    return d


def model_c_terms(c=(1.0, 1.0, 1.0, 1.0, 1.0)):
    """MC (examples/models/modeldefinitions.h:74-80), fields (psi, rho):
        dpsi = -bilap(psi) - lap((c1 - c2*psi*psi)*psi + c5*rho*rho)
        drho = lap(rho) + (c3 - c4*rho*rho)*rho + 2*c5*psi*rho"""
    c1, c2, c3, c4, c5 = c
    psi = [
        MTermSpec(abi.OP_BILAP, -1.0, (1, 0)),
        MTermSpec(abi.OP_LAP, -c1, (1, 0)),
        MTermSpec(abi.OP_LAP, c2, (3, 0)),
        MTermSpec(abi.OP_LAP, -c5, (0, 2)),
    ]
    rho = [
        MTermSpec(abi.OP_LAP, 1.0, (0, 1)),
        MTermSpec(abi.OP_ID, c3, (0, 1)),
        MTermSpec(abi.OP_ID, -c4, (0, 3)),
        MTermSpec(abi.OP_ID, 2.0 * c5, (1, 1)),
    ]
    return [psi, rho]


def gray_scott_terms(c=(0.16, 0.08, 0.035, 0.065)):
    """GRAYSCOTT (examples/models/modeldefinitions.h:123-128), fields (psi, rho),
    prr = psi*rho*rho:
        dpsi = c1*lap(psi) - prr + c3*(1 - psi)
        drho = c2*lap(rho) + prr - (c3 + c4)*rho"""
    c1, c2, c3, c4 = c
    psi = [
        MTermSpec(abi.OP_LAP, c1, (1, 0)),
        MTermSpec(abi.OP_ID, -1.0, (1, 2)),
        MTermSpec(abi.OP_ID, c3, (0, 0)),
        MTermSpec(abi.OP_ID, -c3, (1, 0)),
    ]
    rho = [
        MTermSpec(abi.OP_LAP, c2, (0, 1)),
        MTermSpec(abi.OP_ID, 1.0, (1, 2)),
        MTermSpec(abi.OP_ID, -(c3 + c4), (0, 1)),
    ]
    return [psi, rho]


def turing_terms(c=(1.0, 1.0, 1.0, 1.0, 1.0, 1.0)):
    """Turing (examples/models/modeldefinitions.h:130-136), fields (psi, rho):
        dpsi = c1*lap(psi) + c2*(psi + c3*rho - psi*rho*rho - c4*psi*rho)
        drho = lap(rho) + c2*(c5*rho + c6*psi + psi*rho*rho + c4*psi*rho)"""
    c1, c2, c3, c4, c5, c6 = c
    psi = [
        MTermSpec(abi.OP_LAP, c1, (1, 0)),
        MTermSpec(abi.OP_ID, c2, (1, 0)),
        MTermSpec(abi.OP_ID, c2 * c3, (0, 1)),
        MTermSpec(abi.OP_ID, -c2, (1, 2)),
        MTermSpec(abi.OP_ID, -c2 * c4, (1, 1)),
    ]
    rho = [
        MTermSpec(abi.OP_LAP, 1.0, (0, 1)),
        MTermSpec(abi.OP_ID, c2 * c5, (0, 1)),
        MTermSpec(abi.OP_ID, c2 * c6, (1, 0)),
        MTermSpec(abi.OP_ID, c2, (1, 2)),
        MTermSpec(abi.OP_ID, c2 * c4, (1, 1)),
    ]
    return [psi, rho]


def mixed_terms(k=(0.6, 0.8)):
    """A synthetic two-field model whose pointwise polynomials carry mixed
    monomials of both fields, and where both fields have a bilaplacian:
        dpsi = lap(psi^3 - psi + 0.3 psi rho + 0.2 psi^2 rho + 0.15 psi rho^2)
               - k0 bilap psi + 0.1 psi - 0.05 psi rho^3
        drho = lap(rho^3 - rho + 0.2 psi^2 + 0.1 psi rho + 0.1 rho^2)
               - k1 bilap rho - 0.1 rho + 0.05 psi^2 rho^2
    B_psi and B_rho have total degree 3 and are linear in neither field nor
    separable; A_psi and A_rho have total degree 4 (it exercises every
    polynomial class of the fused coupled kernel's general radius-two form)."""
    k0, k1 = k
    psi = [MTermSpec(abi.OP_LAP, 1.0, (3, 0)), MTermSpec(abi.OP_LAP, -1.0, (1, 0)),
           MTermSpec(abi.OP_LAP, 0.3, (1, 1)), MTermSpec(abi.OP_LAP, 0.2, (2, 1)),
           MTermSpec(abi.OP_LAP, 0.15, (1, 2)), MTermSpec(abi.OP_BILAP, -k0, (1, 0)),
           MTermSpec(abi.OP_ID, 0.1, (1, 0)), MTermSpec(abi.OP_ID, -0.05, (1, 3))]
    rho = [MTermSpec(abi.OP_LAP, 1.0, (0, 3)), MTermSpec(abi.OP_LAP, -1.0, (0, 1)),
           MTermSpec(abi.OP_LAP, 0.2, (2, 0)), MTermSpec(abi.OP_LAP, 0.1, (1, 1)),
           MTermSpec(abi.OP_LAP, 0.1, (0, 2)), MTermSpec(abi.OP_BILAP, -k1, (0, 1)),
           MTermSpec(abi.OP_ID, -0.1, (0, 1)), MTermSpec(abi.OP_ID, 0.05, (2, 2))]
    return [psi, rho]


def mixed(shape, h=1.0, dt=0.002, stencil=Stencil2d2h(5, 6, 13), dtype=abi.F64,
          path=abi.PATH_AUTO, c=(0.6, 0.8)) -> abi.MDesc:
    return make_mdesc(shape, stencil, h, dt, mixed_terms(c), dtype, path)


def model_c(shape, h=1.0, dt=0.01, stencil=Stencil2d2h(5, 6, 13), dtype=abi.F64,
            path=abi.PATH_AUTO, c=(1.0, 1.0, 1.0, 1.0, 1.0)) -> abi.MDesc:
    return make_mdesc(shape, stencil, h, dt, model_c_terms(c), dtype, path)


def gray_scott(shape, h=1.0, dt=1.0, stencil=Stencil2d2h(5, 6, 13), dtype=abi.F64,
               path=abi.PATH_AUTO, c=(0.16, 0.08, 0.035, 0.065)) -> abi.MDesc:
    return make_mdesc(shape, stencil, h, dt, gray_scott_terms(c), dtype, path)


def turing(shape, h=1.0, dt=0.01, stencil=Stencil2d2h(5, 6, 13), dtype=abi.F64,
           path=abi.PATH_AUTO, c=(1.0, 1.0, 1.0, 1.0, 1.0, 1.0)) -> abi.MDesc:
    return make_mdesc(shape, stencil, h, dt, turing_terms(c), dtype, path)


def batched_instances(count: int = 256, seed: int = 7,
                      sides=(128, 192, 256, 384, 512), h: float = 1.0,
                      dt: float = 0.01, coeff_range=(0.5, 2.0), dtype=abi.F32,
                      stability_margin: float = 0.95):
    """The batched workload: count independent periodic 2D instances, the
    first half Allen-Cahn dpsi = M*(kappa*lap psi + psi - psi^3), the second
    half Model B dpsi = M*lap(psi^3 - psi - kappa*lap psi), square sides drawn
    uniformly from `sides`, M and kappa uniform in coeff_range.

    Explicit Euler on -M*kappa*bilap(psi) with the 13-point bilaplacian is
    stable only while dt*M*kappa*64/h^4 <= 2 (the largest eigenvalue of
    lap5∘lap5 is 64/h^4), so Model B draws are redrawn until M*kappa is
    within stability_margin of that bound; with dt = 0.01, h = 1 and
    coefficients in [0.5, 2] about 5% of raw Model B draws exceed it.

    Returns a list of (descriptor, field_seed) pairs; draws come from the
    SplitMix64 stream of `seed`, so the workload is reproducible."""
    from . import inputs
    stream = iter(((inputs.splitmix64(seed, 64 * count) >> 11).astype(np.float64)
                   * (1.0 / 9007199254740992.0)).tolist())
    lo, hi = coeff_range
    bound = stability_margin * 2.0 * h ** 4 / (64.0 * dt)
    out = []
    for i in range(count):
        side = sides[min(int(next(stream) * len(sides)), len(sides) - 1)]
        model_b_inst = i >= count // 2
        while True:
            m = lo + (hi - lo) * next(stream)
            k = lo + (hi - lo) * next(stream)
            if not model_b_inst or m * k <= bound:
                break
        if model_b_inst:
            d = model_b((side, side), h=h, dt=dt, dtype=dtype, path=abi.PATH_FUSED,
                        mobility=m, kappa=k)
        else:
            d = allen_cahn((side, side), h=h, dt=dt, dtype=dtype, path=abi.PATH_FUSED,
                           mobility=m, kappa=k, stencil=Stencil2d2h(5, 6, 13))
        out.append((d, seed * 1000003 + i))
    return out
