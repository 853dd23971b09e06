"""Batched persistent kernel: every instance must equal the single-field
fused kernel bit for bit (same per-site arithmetic, different schedule), and
the FP64 oracle within the FP32 tolerance."""
import numpy as np
import pytest

from symphas_amd import abi, inputs, models
from symphas_amd.solver import FDBatch, FDPlan

pytestmark = pytest.mark.gpu


def _mixed(count=10, seed=3):
    inst = models.batched_instances(count=count, seed=seed)
    fields = [inputs.uniform((d.n[1], d.n[0]), seed=s, dtype=np.float32) for d, s in inst]
    return [d for d, _ in inst], fields


def _single(desc, psi, steps):
    with FDPlan(desc) as p:
        p.upload(psi)
        p.step(steps)
        return p.download()


@pytest.mark.parametrize("K,rows", [(1, 32), (2, 32), (3, 16), (4, 64), (2, 7)])
def test_batch_matches_single_field_bitwise(gpu_lib, K, rows):
    descs, fields = _mixed()
    print("[case] batch", K, rows, [d.n[0] for d in descs], flush=True)
    with FDBatch(descs) as b:
        b.tune(K, rows)
        for i, f in enumerate(fields):
            b.upload(i, f)
        b.step(40)
        got = [b.download(i) for i in range(len(descs))]
    for d, f, g in zip(descs, fields, got):
        want = _single(d, f, 40)
        assert np.array_equal(g, want), (d.n[0], d.terms[0].op)


def test_batch_graph_replay_and_reduce(gpu_lib):
    descs, fields = _mixed(count=16, seed=5)
    print("[case] batch graph replay", flush=True)
    with FDBatch(descs) as b:
        for i, f in enumerate(fields):
            b.upload(i, f)
        m0, e0 = b.reduce()
        b.step(200)                 # graph path
        got = [b.download(i) for i in range(len(descs))]
        m1, e1 = b.reduce()
    with FDBatch(descs) as b:
        for i, f in enumerate(fields):
            b.upload(i, f)
        for _ in range(5):
            b.step(40)              # eager launches
        eager = [b.download(i) for i in range(len(descs))]
    for g, e in zip(got, eager):
        assert np.array_equal(g, e)
    for i, d in enumerate(descs):
        with FDPlan(d) as p:
            p.upload(got[i])
            m, e = p.reduce()
        assert abs(m - m1[i]) <= 1e-12 * np.abs(got[i].astype(np.float64)).sum() + 1e-12
        assert abs(e - e1[i]) <= 1e-12 * abs(e) + 1e-12
        if d.terms[0].op == abi.OP_BILAP:   # Model B conserves mass
            assert abs(m1[i] - m0[i]) <= 1e-6 * np.abs(fields[i].astype(np.float64)).sum()


def test_batch_energy_per_instance(gpu_lib, orc):
    """Each instance reports the functional its own dynamics descend,
    -1/2 psi^2 + 1/4 psi^4 + kappa_i/2 |grad+ psi|^2, equal to the oracle's,
    and that energy decreases under the instance's steps."""
    descs, fields = _mixed(count=8, seed=11)
    energies = [models.landau_energy(d) for d in descs]
    assert len({round(e.kappa, 12) for e in energies}) == len(descs)
    print("[case] batch energy per instance", flush=True)
    with FDBatch(descs) as b:
        b.set_energy(energies)
        for i, f in enumerate(fields):
            b.upload(i, f)
        _, e0 = b.reduce()
        traj = [e0]
        for _ in range(3):
            b.step(40)
            traj.append(b.reduce()[1])
        last = [b.download(i) for i in range(len(descs))]
        b.set_energy([models.landau_energy(d, gradient=abi.GRAD_CENTRAL, mean=True) for d in descs])
        _, ec = b.reduce()
    for i, d in enumerate(descs):
        g = orc.refresh(d, orc.pad(d, last[i].astype(np.float64)))
        want = orc.energy(d, g, kappa=energies[i].kappa)
        assert abs(traj[-1][i] - want) <= 1e-12 * abs(want)
        wantc = orc.energy(d, g, kappa=energies[i].kappa, gradient=1, mean=True)
        assert abs(ec[i] - wantc) <= 1e-12 * abs(wantc)
        seq = [t[i] for t in traj]
        assert all(b_ < a_ for a_, b_ in zip(seq, seq[1:])), (i, seq)


def test_batch_against_oracle(gpu_lib, orc):
    descs, fields = _mixed(count=4, seed=9)
    print("[case] batch vs oracle", flush=True)
    with FDBatch(descs) as b:
        for i, f in enumerate(fields):
            b.upload(i, f)
        b.step(100)
        got = [b.download(i).astype(np.float64) for i in range(len(descs))]
    for d, f, g in zip(descs, fields, got):
        d64 = models.make_desc((d.n[0], d.n[1]), models.Stencil2d2h(5, 6, 13), d.h, d.dt,
                               [models.TermSpec(d.terms[t].op, d.terms[t].coeff,
                                                tuple(d.terms[t].poly)) for t in range(d.nterms)],
                               dtype=abi.F64, path=abi.PATH_REFERENCE)
        want = orc.unpad(d64, orc.run(d64, f.astype(np.float64), 100, 8))
        err = np.linalg.norm(g - want) / np.linalg.norm(want)
        assert err <= 1e-5, err


def test_batch_uneven_extents_bitwise(gpu_lib):
    """Widths that are not multiples of the strip width, rectangular
    instances and a width narrower than one strip, against single-field
    runs bit for bit."""
    shapes = [(70, 50), (130, 90), (8, 12), (250, 34), (514, 40)]
    descs, fields = [], []
    for i, (x, y) in enumerate(shapes):
        mk = models.model_b if i % 2 == 0 else models.allen_cahn
        descs.append(mk((x, y), h=1.0, dt=0.01, dtype=abi.F32, path=abi.PATH_FUSED))
        fields.append(inputs.uniform((y, x), seed=30 + i, dtype=np.float32))
    print("[case] batch uneven extents", shapes, flush=True)
    with FDBatch(descs) as b:
        for i, f in enumerate(fields):
            b.upload(i, f)
        b.step(37)
        got = [b.download(i) for i in range(len(descs))]
    for d, f, g in zip(descs, fields, got):
        assert np.array_equal(g, _single(d, f, 37)), (d.n[0], d.n[1])


def test_batch_rejects_unsupported(gpu_lib):
    d = models.model_b((101, 128), dtype=abi.F32)
    with pytest.raises(abi.SpsError) as ei:
        FDBatch([d])
    assert ei.value.status == abi.E_UNSUPPORTED
    d = models.model_b((128, 128), dtype=abi.F64)
    with pytest.raises(abi.SpsError):
        FDBatch([d])


def test_batch_autotune_keeps_fields(gpu_lib):
    descs, fields = _mixed(count=6, seed=13)
    print("[case] batch autotune", flush=True)
    with FDBatch(descs) as b:
        for i, f in enumerate(fields):
            b.upload(i, f)
        k, rows = b.autotune()
        assert k in (1, 2, 3, 4) and 4 <= rows <= 512
        for i, f in enumerate(fields):
            assert np.array_equal(b.download(i), f)
        b.step(24)
        got = [b.download(i) for i in range(len(descs))]
    for d, f, g in zip(descs, fields, got):
        assert np.array_equal(g, _single(d, f, 24))
