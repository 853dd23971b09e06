"""Learned surrogate for k-step advances of the Model B solver, with the
plain finite-difference solver as its data source, fallback and referee.

* Data: trajectories of the batched GPU solver (Model B 2D, periodic, h = 1,
  dt = 0.01), psi0 uniform in [-0.1, 0.1) plus a mean offset uniform in
  [-0.3, 0.3), sampled every k = 50 steps up to a horizon (20000 steps by
  default).  Each instance has its own random lifetime, so a batch mixes early
  (spinodal) and late (coarsening) states.  Pairs (psi_t, psi_{t+k}) are
  copied device to device into torch tensors.
* Model: nn.UNet (BF16 MFMA kernels), trained with Adam on the per-sample
  relative increment loss mean_i ||pred_i - y_i||^2 / ||y_i - x_i||^2.
* Inference with fallback (SurrogateSolver.advance): the network advances the
  samples still on the surrogate, each checked for mass drift and free-energy
  increase; every `check_every`-th advance (and any advance failing a cheap
  check) is compared with the plain solver.  A sample whose check exceeds
  `tol` relative takes the plain solver's state and the plain solver advances
  it from then on; the network no longer sees it.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import abi, inputs, models
from . import nn
from .nn import UNet
from .solver import FDBatch


def model_b_desc(n, h=1.0, dt=0.01):
    return models.model_b((n, n), h=h, dt=dt, dtype=abi.F32, path=abi.PATH_FUSED)


def initial_field(n, seed, offset_range=(-0.3, 0.3)):
    noise = inputs.uniform((n, n), seed=seed, lo=-0.1, hi=0.1)
    u = float(((inputs.splitmix64(seed ^ 0x5DEECE66D, 1) >> 11)[0]) * (1.0 / 9007199254740992.0))
    return (noise + offset_range[0] + (offset_range[1] - offset_range[0]) * u).astype(np.float32)


class PlainSolver:
    """A batch of Model B fields advanced by the batched HIP kernel, with
    device-to-device copies to and from torch tensors [B, n, n] FP32."""

    def __init__(self, batch, n, h=1.0, dt=0.01, stream=None):
        self.batch, self.n = batch, n
        self.fd = FDBatch([model_b_desc(n, h, dt)] * batch)
        self.stream = stream or torch.cuda.current_stream()
        self.fd.set_stream(self.stream.cuda_stream)

    def load(self, psi):
        psi = psi.contiguous()
        for i in range(self.batch):
            self.fd.copy_from_device(i, psi[i].data_ptr())

    def store(self, out):
        for i in range(self.batch):
            self.fd.copy_to_device(i, out[i].data_ptr())
        return out

    def advance(self, psi, steps):
        self.load(psi)
        self.fd.step(steps)
        return self.store(torch.empty_like(psi))

    def close(self):
        self.fd.close()


class TrajectoryStream:
    """(psi_t, psi_{t+k}) pairs from `batch` running solver trajectories."""

    def __init__(self, batch=64, n=256, k=50, horizon=20000, seed=1, dt=0.01):
        self.batch, self.n, self.k, self.horizon = batch, n, k, horizon
        self.rng = np.random.default_rng(seed)
        self.seed_base = seed * 100003
        self.count = 0
        self.solver = PlainSolver(batch, n, dt=dt)
        self.age = np.zeros(batch, dtype=np.int64)
        self.life = np.zeros(batch, dtype=np.int64)
        self.cur = torch.empty((batch, n, n), device="cuda", dtype=torch.float32)
        for i in range(batch):
            self._reset(i)
        self.solver.load(self.cur)

    def _reset(self, i):
        self.count += 1
        f = initial_field(self.n, self.seed_base + self.count)
        self.cur[i].copy_(torch.from_numpy(f))
        self.age[i] = 0
        self.life[i] = self.k * int(self.rng.integers(1, self.horizon // self.k + 1))

    def next_pair(self):
        x = self.solver.store(torch.empty_like(self.cur))
        self.solver.fd.step(self.k)
        y = self.solver.store(torch.empty_like(self.cur))
        self.age += self.k
        dead = np.nonzero(self.age >= self.life)[0]
        if len(dead):
            self.cur.copy_(y)
            for i in dead:
                self._reset(int(i))
            self.solver.load(self.cur)
        return x, y

    def close(self):
        self.solver.close()


def relative_increment_loss(pred, x, y):
    """mean over samples of ||pred - y||^2 / ||y - x||^2: every sample's
    increment error counts relative to its own increment, so the small late-time
    increments weigh as much as the large early-time ones."""
    num = ((pred - y) ** 2).flatten(1).sum(1)
    den = ((y - x) ** 2).flatten(1).sum(1).clamp_min(1e-30)
    return (num / den).mean()


@dataclass
class TrainLog:
    iters: int = 0
    seconds: float = 0.0
    losses: list = field(default_factory=list)
    evaluations: list = field(default_factory=list)  # (iterations, evaluate(net))


def train(net, iters=400, batch=32, n=256, k=50, horizon=20000, lr=3e-4, seed=1,
          log_every=50, verbose=True, evaluate=None, eval_at=()):
    """Train on pairs generated on the fly by the GPU solver (Adam, constant
    learning rate: a probe of lr in {3e-4, 1e-3} and output scale in {0.01,
    0.1} on 128^2 fields picked 3e-4 with scale 0.1)."""
    stream = TrajectoryStream(batch=batch, n=n, k=k, horizon=horizon, seed=seed)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    log = TrainLog()
    t0 = time.perf_counter()
    for it in range(iters):
        x, y = stream.next_pair()
        pred = net(x)
        loss = relative_increment_loss(pred, x, y)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        if it % log_every == 0 or it == iters - 1:
            lv = float(loss.detach())
            log.losses.append((it, lv))
            if verbose:
                print(f"[train] iter {it} rel-increment loss {lv:.4e}", flush=True)
        if evaluate is not None and it + 1 in eval_at:
            log.evaluations.append((it + 1, float(evaluate(net))))
    torch.cuda.synchronize()
    log.iters = iters
    log.seconds = time.perf_counter() - t0
    stream.close()
    return log


def free_energy(psi, h=1.0):
    """h^2 * sum[-1/2 psi^2 + 1/4 psi^4 + 1/2 |grad+ psi|^2] per sample (FP64)."""
    p = psi.double()
    gx = (torch.roll(p, -1, dims=2) - p) / h
    gy = (torch.roll(p, -1, dims=1) - p) / h
    f = -0.5 * p * p + 0.25 * p ** 4 + 0.5 * (gx * gx + gy * gy)
    return f.sum(dim=(1, 2)) * h * h


def rel_l2(a, b):
    return ((a.double() - b.double()).flatten(1).norm(dim=1)
            / b.double().flatten(1).norm(dim=1).clamp_min(1e-300))


@dataclass
class AdvanceReport:
    advances: int = 0
    surrogate_calls: int = 0
    checks: int = 0
    fallbacks: int = 0          # samples taken over by the plain solver
    max_check_err: float = 0.0
    seconds: float = 0.0


class SurrogateSolver:
    """Surrogate k-step advances checked by the plain solver.

    Per advance: the network predicts only the samples still on the surrogate
    ("live"); a cheap check (mass drift, free-energy increase) runs on those.
    The plain solver (one batched launch over every sample) runs when some
    sample is already on the plain path, when a cheap check failed, or every
    `check_every`-th advance; that one call is both the referee of the live
    samples and the advance of the fallen-back ones.  A live sample whose
    relative L2 error against the referee (or whose cheap check) exceeds `tol`
    takes the referee's state and stays on the plain path.  Once every sample
    is on the plain path the network is not called again and the remaining
    advances run as one plain call, so an all-fallback batch costs the plain
    solver plus the network calls made before the fallbacks."""

    def __init__(self, net, batch, n, k=50, tol=1e-3, check_every=4):
        self.net, self.k, self.tol, self.check_every = net, k, tol, check_every
        # the U-Net's inference replays a captured HIP graph per batch shape
        self.forward = nn.GraphedForward(net) if isinstance(net, nn.UNet) else net
        self.referee = PlainSolver(batch, n)

    def advance(self, psi, n_adv):
        rep = AdvanceReport()
        psi = psi.clone()
        batch = psi.shape[0]
        on_plain = np.zeros(batch, dtype=bool)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a in range(n_adv):
            live = np.nonzero(~on_plain)[0]
            if len(live) == 0:
                # every sample is on the plain path: the rest in one call
                psi = self.referee.advance(psi, (n_adv - a) * self.k)
                rep.advances += n_adv - a
                break
            all_live = len(live) == batch
            idx = None if all_live else torch.from_numpy(live).to(psi.device)
            x = psi if all_live else psi.index_select(0, idx)
            with torch.no_grad():
                pred = self.forward(x)
            rep.surrogate_calls += 1
            m0, m1 = x.double().sum((1, 2)), pred.double().sum((1, 2))
            mass_bad = (m1 - m0).abs() > self.tol * x.double().abs().sum((1, 2))
            f0, f1 = free_energy(x), free_energy(pred)
            energy_bad = (f1 - f0) > self.tol * f0.abs()
            cheap_bad = (mass_bad | energy_bad).cpu().numpy()
            need_ref = (not all_live) or bool(cheap_bad.any()) or a % self.check_every == 0
            if need_ref:
                # one batched plain advance: referee of the live samples and
                # the advance of those already on the plain path
                ref = self.referee.advance(psi, self.k)
                rep.checks += 1
                ref_live = ref if all_live else ref.index_select(0, idx)
                err = rel_l2(pred, ref_live)
                rep.max_check_err = max(rep.max_check_err, float(err.max()))
                bad = (err > self.tol).cpu().numpy() | cheap_bad
                rep.fallbacks += int(bad.sum())
                on_plain[live[bad]] = True
                keep = torch.from_numpy(~bad).to(psi.device)[:, None, None]
                merged = torch.where(keep, pred, ref_live)
                if all_live:
                    psi = merged
                else:
                    psi = ref
                    psi.index_copy_(0, idx, merged)
            else:
                psi = pred  # every sample is live here
            rep.advances += 1
        torch.cuda.synchronize()
        rep.seconds = time.perf_counter() - t0
        return psi, rep

    def close(self):
        self.referee.close()


def time_plain(psi, steps, n):
    solver = PlainSolver(psi.shape[0], n)
    solver.load(psi)
    solver.fd.step(min(steps, 50))
    solver.load(psi)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.fd.step(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = solver.store(torch.empty_like(psi))
    solver.close()
    return out, dt


def time_network(net, psi, reps=5):
    """Seconds per inference of net on psi, as the controller runs it (the
    U-Net through its captured graph)."""
    fwd = nn.GraphedForward(net) if isinstance(net, nn.UNet) else net
    with torch.no_grad():
        fwd(psi)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fwd(psi)
        torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps
