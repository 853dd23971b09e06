"""Surrogate: training data come from the GPU solver, the trained network
beats the method without it (persistence psi(t+k) = psi(t), and the untrained
network), and the plain solver takes over when the surrogate fails its
checks, returning exactly the plain solver's trajectory."""
import numpy as np
import pytest
import torch

from symphas_amd import abi, models
from symphas_amd.nn import UNet
from symphas_amd import surrogate as sg
from symphas_amd.solver import FDPlan

pytestmark = pytest.mark.gpu


def test_trajectory_pairs_are_solver_advances(gpu_lib):
    print("[case] trajectory pairs", flush=True)
    ts = sg.TrajectoryStream(batch=4, n=128, k=50, horizon=400, seed=3)
    for _ in range(3):
        x, y = ts.next_pair()
    ts.close()
    for i in range(4):
        with FDPlan(sg.model_b_desc(128)) as p:
            p.upload(x[i].cpu().numpy())
            p.step(50)
            want = p.download()
        assert np.array_equal(y[i].cpu().numpy(), want)


def _eval_loss(net, n, k, batch, seed, steps_before=600):
    """relative increment error on held-out seeds after steps_before steps"""
    solver = sg.PlainSolver(batch, n)
    psi0 = torch.stack([torch.from_numpy(sg.initial_field(n, 10**7 + seed + i)) for i in range(batch)]).cuda()
    x = solver.advance(psi0, steps_before)
    y = solver.advance(x, k)
    solver.close()
    with torch.no_grad():
        pred = net(x)
    return float(sg.relative_increment_loss(pred, x, y)), float(sg.relative_increment_loss(x, x, y))


def test_training_beats_persistence(gpu_lib):
    print("[case] training 128^2", flush=True)
    torch.manual_seed(0)
    n, k = 128, 50
    net = UNet(seed=3).cuda()
    untrained, persistence = _eval_loss(net, n, k, 16, seed=1)
    log = sg.train(net, iters=400, batch=16, n=n, k=k, horizon=3000, seed=5,
                   log_every=100)
    trained, _ = _eval_loss(net, n, k, 16, seed=1)
    print(f"held-out relative increment error: persistence {persistence:.3f}, "
          f"untrained {untrained:.3f}, trained {trained:.4f} ({log.seconds:.1f} s)")
    assert persistence == pytest.approx(1.0)
    assert trained < 0.1 * persistence
    assert trained < 0.1 * untrained


def _unfit_net():
    net = UNet(seed=4).cuda()
    with torch.no_grad():  # an unfit network: large random increments
        net.w_out.copy_(torch.randn_like(net.w_out))
    return net


def test_fallback_returns_plain_trajectory(gpu_lib):
    print("[case] fallback", flush=True)
    n, k, batch = 128, 50, 8
    net = _unfit_net()
    psi0 = torch.stack([torch.from_numpy(sg.initial_field(n, 500 + i)) for i in range(batch)]).cuda()
    ctl = sg.SurrogateSolver(net, batch, n, k=k, tol=1e-3, check_every=4)
    out, rep = ctl.advance(psi0, 6)
    ctl.close()
    assert rep.fallbacks == batch
    # all samples fell back at the first check: one network call, one referee
    # advance, then the remaining five advances as one plain call
    assert rep.surrogate_calls == 1 and rep.checks == 1 and rep.advances == 6
    plain, _ = sg.time_plain(psi0, 6 * k, n)
    assert torch.equal(out, plain)


def test_all_fallback_costs_about_the_plain_solver(gpu_lib):
    """256^2 in batches of 64: once every sample has fallen back, the network
    is not called again, so the advance costs the plain solver plus the one
    network call that found the failures."""
    print("[case] all-fallback cost 256^2 x 64", flush=True)
    n, k, batch, n_adv = 256, 50, 64, 40
    net = _unfit_net()
    psi0 = torch.stack([torch.from_numpy(sg.initial_field(n, 700 + i)) for i in range(batch)]).cuda()
    ctl = sg.SurrogateSolver(net, batch, n, k=k, tol=1e-3, check_every=4)
    ctl.advance(psi0, 2)  # warm: kernels, graphs
    out, rep = ctl.advance(psi0, n_adv)
    ctl.close()
    assert rep.fallbacks == batch and rep.surrogate_calls == 1 and rep.checks == 1
    plain, t_plain = sg.time_plain(psi0, n_adv * k, n)
    t_net = sg.time_network(net, psi0)
    print(f"all-fallback advance {rep.seconds * 1e3:.2f} ms; plain {t_plain * 1e3:.2f} ms, "
          f"one network call {t_net * 1e3:.2f} ms")
    assert torch.equal(out, plain)
    # host copies and checks around the single network call: a few ms of slack
    assert rep.seconds <= 1.05 * (t_plain + t_net) + 5e-3


class _HalfBad(torch.nn.Module):
    """The plain solver itself for fields of negative mean; a mass-breaking
    prediction for fields of positive mean.  Records the batch sizes it sees."""

    def __init__(self, n, k):
        super().__init__()
        self.n, self.k, self.sizes, self.solvers = n, k, [], {}

    def forward(self, psi):
        b = psi.shape[0]
        self.sizes.append(b)
        if b not in self.solvers:
            self.solvers[b] = sg.PlainSolver(b, self.n)
        exact = self.solvers[b].advance(psi, self.k)
        bad = (psi.mean((1, 2)) > 0)[:, None, None]
        return torch.where(bad, exact + 0.05, exact)

    def close(self):
        for s in self.solvers.values():
            s.close()


def test_partial_fallback_drops_samples_from_the_network(gpu_lib):
    print("[case] partial fallback", flush=True)
    n, k, batch = 128, 50, 8
    psi0 = torch.stack([torch.from_numpy(sg.initial_field(n, 800 + i)) for i in range(batch)]).cuda()
    positive = int((psi0.mean((1, 2)) > 0).sum())
    assert 0 < positive < batch
    net = _HalfBad(n, k)
    ctl = sg.SurrogateSolver(net, batch, n, k=k, tol=1e-3, check_every=4)
    out, rep = ctl.advance(psi0, 5)
    ctl.close()
    net.close()
    assert rep.fallbacks == positive
    assert net.sizes == [batch] + [batch - positive] * 4
    plain, _ = sg.time_plain(psi0, 5 * k, n)
    assert torch.equal(out, plain)


class _Exact(torch.nn.Module):
    """A 'surrogate' that is the plain solver itself."""

    def __init__(self, batch, n, k):
        super().__init__()
        self.s, self.k = sg.PlainSolver(batch, n), k

    def forward(self, psi):
        return self.s.advance(psi, self.k)


def test_controller_keeps_an_accurate_surrogate(gpu_lib):
    print("[case] accurate surrogate", flush=True)
    n, k, batch = 128, 50, 8
    psi0 = torch.stack([torch.from_numpy(sg.initial_field(n, 900 + i)) for i in range(batch)]).cuda()
    ex = _Exact(batch, n, k)
    ctl = sg.SurrogateSolver(ex, batch, n, k=k, tol=1e-3, check_every=2)
    out, rep = ctl.advance(psi0, 6)
    ctl.close()
    assert rep.fallbacks == 0 and rep.checks == 3 and rep.max_check_err == 0.0
    plain, _ = sg.time_plain(psi0, 6 * k, n)
    assert torch.equal(out, plain)
