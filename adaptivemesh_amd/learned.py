"""Learned per-octant prolongation: training on the GPU through the native
MLP kernels (am_mlp_* in include/adaptivemesh.h).

Device memory comes from torch (plumbing only); every computation --
pairs, forward, backward, loss, Adam -- is a hand-written HIP kernel.

    pairs = PairSet.from_sim(sim, tiles)          # restrict/prolong pairs of a run
    net = Prolongation(seed=0)
    net.fit(pairs, steps=2000, batch=4096, lr=1e-3)
    sim.set_prolongation(net)                     # learned coarse-face ghosts

The network predicts the 8 fine children of a coarse cell from its 4^3
coarse stencil (24 fields), as a normalised correction to the tricubic
baseline; see am_mlp_pairs for the normalisation.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import MlpWork, check, lib, stream_handle

N_IN, N_HID, N_OUT = 1536, 128, 192


def _torch():
    import torch
    return torch


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform fan-in initialisation, zero biases; the last layer is scaled
    by 0.1 so the untrained network stays close to the tricubic baseline."""
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (N_IN, N_HID)) * math.sqrt(6.0 / N_IN)
    w2 = rng.uniform(-1, 1, (N_HID, N_HID)) * math.sqrt(6.0 / N_HID)
    w3 = rng.uniform(-1, 1, (N_HID, N_OUT)) * math.sqrt(6.0 / N_HID) * 0.1
    parts = [w1, np.zeros(N_HID), w2, np.zeros(N_HID), w3, np.zeros(N_OUT)]
    p = np.concatenate([np.asarray(a, np.float32).ravel() for a in parts])
    assert p.size == lib().am_mlp_param_count()
    return p


class PairSet:
    """Device-resident training pairs: x (n, 1536), t (n, 192), mu/sigma (n, 24)."""

    def __init__(self, x, t, mu, sigma):
        self.x, self.t, self.mu, self.sigma = x, t, mu, sigma

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @staticmethod
    def from_fields(field_ptrs, tiles, geom, device="cuda"):
        """Pairs of the listed tiles of 24 padded device fields."""
        torch = _torch()
        per = int(lib().am_mlp_pairs_per_tile(C.byref(geom)))
        n = per * len(tiles)
        x = torch.empty((n, N_IN), dtype=torch.float32, device=device)
        t = torch.empty((n, N_OUT), dtype=torch.float32, device=device)
        mu = torch.empty((n, 24), dtype=torch.float32, device=device)
        sg = torch.empty((n, 24), dtype=torch.float32, device=device)
        tl = torch.tensor(np.asarray(tiles, np.int32), device=device)
        ptrs = (C.c_void_p * 24)(*field_ptrs)
        check(lib().am_mlp_pairs(ptrs, tl.data_ptr(), len(tiles), C.byref(geom), x.data_ptr(), t.data_ptr(),
                                 mu.data_ptr(), sg.data_ptr(), stream_handle()), "am_mlp_pairs")
        return PairSet(x, t, mu, sg)

    @staticmethod
    def from_sim(sim, tiles, which: int = 1):
        """Pairs of a driver's fields (`which` 0 old, 1 new), converted to
        padded tiles first (the driver keeps them interior-first)."""
        padded = sim.padded_fields(which)
        out = PairSet.from_fields([padded[v].data_ptr() for v in range(24)], tiles, sim.geom)
        _torch().cuda.synchronize()  # the pair kernel has read `padded` before it is freed
        return out

    @staticmethod
    def cat(sets):
        torch = _torch()
        return PairSet(*(torch.cat([getattr(s, k) for s in sets]) for k in ("x", "t", "mu", "sigma")))

    def split(self, n_test: int, seed: int = 0):
        torch = _torch()
        g = torch.Generator(device="cpu").manual_seed(seed)
        perm = torch.randperm(self.n, generator=g).to(self.x.device)
        te, tr = perm[:n_test], perm[n_test:]
        pick = lambda idx: PairSet(self.x[idx], self.t[idx], self.mu[idx], self.sigma[idx])
        return pick(tr), pick(te)


class Prolongation:
    """The MLP's parameters and Adam state on the device."""

    def __init__(self, seed: int = 0, params: np.ndarray | None = None):
        torch = _torch()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.params = torch.tensor(p, device="cuda")
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grads = torch.zeros_like(self.params)
        self.step_count = 0
        self._work = {}

    def work(self, batch: int):
        if batch not in self._work:
            torch = _torch()
            w = {k: torch.empty((batch, N_HID), dtype=torch.float32, device="cuda")
                 for k in ("z1", "h1", "z2", "h2", "dz2", "dz1")}
            w["y"] = torch.empty((batch, N_OUT), dtype=torch.float32, device="cuda")
            w["dy"] = torch.empty((batch, N_OUT), dtype=torch.float32, device="cuda")
            w["partial"] = torch.empty(int(lib().am_mlp_mse_blocks(batch)), dtype=torch.float64, device="cuda")
            s = MlpWork()
            for k in ("z1", "h1", "z2", "h2", "y", "dy", "dz2", "dz1"):
                setattr(s, k, w[k].data_ptr())
            w["struct"] = s
            self._work[batch] = w
        return self._work[batch]

    def forward(self, x, batch: int | None = None):
        """Network output for the first `batch` rows of x (batch % 64 == 0)."""
        b = int(x.shape[0] if batch is None else batch)
        w = self.work(b)
        check(lib().am_mlp_forward(self.params.data_ptr(), x.data_ptr(), b, C.byref(w["struct"]),
                                   stream_handle()), "am_mlp_forward")
        return w["y"]

    def loss_and_grads(self, x, t):
        """MSE loss (FP64 from the device partial sums) and self.grads."""
        b = int(x.shape[0])
        w = self.work(b)
        self.forward(x)
        check(lib().am_mlp_mse(w["y"].data_ptr(), t.data_ptr(), b, w["dy"].data_ptr(), w["partial"].data_ptr(),
                               stream_handle()), "am_mlp_mse")
        check(lib().am_mlp_backward(self.params.data_ptr(), x.data_ptr(), b, C.byref(w["struct"]),
                                    self.grads.data_ptr(), stream_handle()), "am_mlp_backward")
        return w["partial"]

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        check(lib().am_mlp_adam(self.params.data_ptr(), self.grads.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                                self.params.numel(), self.step_count, lr, beta1, beta2, eps, stream_handle()),
              "am_mlp_adam")

    def fit(self, pairs: PairSet, steps: int = 2000, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
            report_every: int = 0, lr_final: float | None = None):
        """Adam on shuffled mini-batches (cosine decay to lr_final when
        given); returns the loss history (FP64) at the reporting steps."""
        torch = _torch()
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = pairs.n - pairs.n % batch
        assert n >= batch and batch % 64 == 0
        perm = torch.randperm(pairs.n, generator=g).to(pairs.x.device)
        xb = torch.empty((batch, N_IN), dtype=torch.float32, device="cuda")
        tb = torch.empty((batch, N_OUT), dtype=torch.float32, device="cuda")
        hist = []
        pos = 0
        for s in range(steps):
            if pos + batch > n:
                perm = torch.randperm(pairs.n, generator=g).to(pairs.x.device)
                pos = 0
            idx = perm[pos:pos + batch]
            pos += batch
            torch.index_select(pairs.x, 0, idx, out=xb)
            torch.index_select(pairs.t, 0, idx, out=tb)
            part = self.loss_and_grads(xb, tb)
            rate = lr if lr_final is None else lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * s / steps))
            self.adam(lr=rate)
            if report_every and (s % report_every == 0 or s == steps - 1):
                hist.append((s, float(part.sum().item()) / (batch * N_OUT)))
        return hist

    def evaluate(self, pairs: PairSet, batch: int = 4096):
        """Mean squared normalised error of the network and of the tricubic
        baseline (whose normalised error is the target itself)."""
        torch = _torch()
        n = pairs.n - pairs.n % batch
        se_net, se_base = 0.0, 0.0
        for i in range(0, n, batch):
            y = self.forward(pairs.x[i:i + batch].contiguous())
            t = pairs.t[i:i + batch]
            se_net += float(((y.double() - t.double()) ** 2).sum().item())
            se_base += float((t.double() ** 2).sum().item())
        return se_net / (n * N_OUT), se_base / (n * N_OUT)

    def numpy_params(self) -> np.ndarray:
        return self.params.detach().cpu().numpy()


# ---------------------------------------------------------------- workflow
def _sim_pairs(am, sim, min_depth: int = 1, max_tiles: int = 0):
    """Pairs of every tile of depth >= min_depth, or of max_tiles of them
    evenly spaced through the slot order when there are more."""
    e = sim.export()
    tiles = [s for s, p in enumerate(e["positions"]) if int(p) != 0xFFFFFFFF and (int(p) >> 27) >= min_depth]
    if max_tiles and len(tiles) > max_tiles:
        tiles = [tiles[i] for i in np.linspace(0, len(tiles) - 1, max_tiles).round().astype(int)]
    return PairSet.from_sim(sim, tiles)


def harvest(am, outer: float = 32.0, masses=(0.5, 1.0, 2.0), snapshots: int = 3,
            gauge=((0.02, 0.0, 0), (0.05, 8.0, 1), (0.1, 16.0, 2), (0.05, 24.0, 0)), S: int = 16, P: int = 3):
    """Training pairs from the plain solver at twice the resolution of the
    puncture run on [-32, 32]^3 (outer = 32 by default): for each puncture
    mass, the tricubic octree run's fields after 0..snapshots-1 coarse steps,
    264 of the tiles of depth >= 2; plus analytic gauge waves (amplitude, time,
    axis) on a uniform depth-2 mesh of one wavelength, every tile of depth
    >= 1."""
    from . import workloads as WL
    sets = []
    g = am.geometry(S, P, 4, outer, 0.25)
    for m in masses:
        ops, n = WL.puncture_ops(outer)  # the run's mesh at half the size: twice the resolution
        sim = am.BssnSim(g, n, am.bssn_options(periodic=False, interp=am.AM_INTERP_TRICUBIC))
        for p_, c in ops:
            sim.add_node(p_, c)
        f = WL.puncture_fields(sim.export()["positions"], S, P, outer, mass=m)
        for v in range(24):
            sim.set_field(0, v, f[v])
            sim.set_field(1, v, f[v])
        sim.fill_ghosts()
        for k in range(snapshots):
            if k:
                sim.timestep()
            sets.append(_sim_pairs(am, sim, 2, max_tiles=264))
        del sim
    for (a, t, axis) in gauge:
        ops, n = WL.uniform_ops(2)
        gw = am.geometry(S, P, 4, outer, 0.25)
        sim = am.BssnSim(gw, n, am.bssn_options(periodic=True))
        for p_, c in ops:
            sim.add_node(p_, c)
        f = WL.gauge_wave_fields(sim.export()["positions"], S, P, outer, amplitude=a, t=t, axis=axis)
        for v in range(24):
            sim.set_field(1, v, f[v])
        sets.append(_sim_pairs(am, sim, 1))
        del sim
    return PairSet.cat(sets)


def puncture_run(am, net=None, threshold: float = 1e-5, coarse_steps: int = 2, mass: float = 1.0,
                 outer: float = 64.0, S: int = 16, P: int = 3, ko_sigma: float = 0.3):
    """The octree puncture run (upwinded advection + Kreiss-Oliger ko_sigma)
    with tricubic ghosts (net None) or the learned prolongation: constraint
    norms after `coarse_steps` coarse steps, the throughput of the timed
    steps and the learned fill's statistics."""
    import time

    import torch

    from . import workloads as WL
    ops, n = WL.puncture_ops(outer)
    g = am.geometry(S, P, 4, outer, 0.25)
    interp = am.AM_INTERP_TRICUBIC if net is None else am.AM_INTERP_LEARNED
    sim = am.BssnSim(g, n, am.bssn_options(periodic=False, interp=interp))
    for p_, c in ops:
        sim.add_node(p_, c)
    if net is not None:
        sim.set_prolongation(net, threshold)
    sim.set_dissipation(ko_sigma, True)
    f = WL.puncture_fields(sim.export()["positions"], S, P, outer, mass=mass)
    for v in range(24):
        sim.set_field(0, v, f[v])
        sim.set_field(1, v, f[v])
    sim.fill_ghosts()
    c0 = sim.constraints()
    sim.sync()
    torch.cuda.synchronize()
    z0, t0 = sim.zone_cycles, time.perf_counter()
    for _ in range(coarse_steps):
        sim.timestep()
    sim.sync()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    c1 = sim.constraints()
    out = dict(H_l2=c1["H_l2"], M_l2=c1["M_l2"], H_l2_initial=c0["H_l2"],
               zone_cycles_per_s=(sim.zone_cycles - z0) / dt, ms_per_coarse_step=1e3 * dt / coarse_steps)
    if net is not None:
        out["learned"] = sim.prolongation_stats()
    return out


def train_and_compare(am, train_steps: int = 3000, batch: int = 4096, seed: int = 0, threshold: float = 1e-5,
                      coarse_steps: int = 2):
    """The learned-versus-plain comparison: harvest pairs at twice the
    resolution, train on the GPU, evaluate on held-out pairs of an unseen
    mass at the run's resolution, then run the puncture with both fills."""
    import time

    import torch
    t0 = time.perf_counter()
    pairs = harvest(am)
    train, val = pairs.split(16384, seed=seed)
    del pairs
    net = Prolongation(seed=seed)
    t1 = time.perf_counter()
    hist = net.fit(train, steps=train_steps, batch=batch, lr=1e-3, lr_final=3e-5, seed=seed,
                   report_every=max(1, train_steps // 4))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    n_train = train.n
    del train
    e_net, e_tri = net.evaluate(val, batch=4096)
    # unseen data at the run's own resolution: mass 1.5 on [-32, 32]^3
    test = _unseen(am)
    u_net, u_tri = net.evaluate(test, batch=4096)
    # each run twice, the second (warm: graphs, allocator, caches) reported
    plain = puncture_run(am, None, coarse_steps=coarse_steps)
    plain = puncture_run(am, None, coarse_steps=coarse_steps)
    learned = puncture_run(am, net, threshold=threshold, coarse_steps=coarse_steps)
    learned = puncture_run(am, net, threshold=threshold, coarse_steps=coarse_steps)
    # the same run with every block that can use the network using it
    learned_all = puncture_run(am, net, threshold=float("inf"), coarse_steps=coarse_steps)
    learned_all = puncture_run(am, net, threshold=float("inf"), coarse_steps=coarse_steps)
    return net, dict(pairs=n_train, train_steps=train_steps, batch=batch, loss_start=hist[0][1],
                     loss_end=hist[-1][1], harvest_s=round(t1 - t0, 2), train_s=round(t2 - t1, 2),
                     heldout_mse_network=e_net, heldout_mse_tricubic=e_tri,
                     unseen_mse_network=u_net, unseen_mse_tricubic=u_tri, plain=plain, learned=learned,
                     learned_all=learned_all, threshold=threshold)


def _unseen(am, S: int = 16, P: int = 3):
    from . import workloads as WL
    outer = 64.0
    ops, n = WL.puncture_ops(outer)
    g = am.geometry(S, P, 4, outer, 0.25)
    sim = am.BssnSim(g, n, am.bssn_options(periodic=False, interp=am.AM_INTERP_TRICUBIC))
    for p_, c in ops:
        sim.add_node(p_, c)
    f = WL.puncture_fields(sim.export()["positions"], S, P, outer, mass=1.5)
    for v in range(24):
        sim.set_field(0, v, f[v])
        sim.set_field(1, v, f[v])
    sim.fill_ghosts()
    sim.timestep()
    return _sim_pairs(am, sim, 2, max_tiles=264)
