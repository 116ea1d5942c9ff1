"""Writes the FP64 reference runs of the precision study: the FP64 scheme
(oracle/oracle.py Bssn64Sim: the RK4 octree schedule with exact FP64
derivatives) on the 512-tile uniform periodic mesh (depth 3, S = 16, P = 3,
fd_order 4, [-32, 32]^3, CFL 0.25), 50 coarse steps, for two data sets:

* noise: flat space + uniform noise of +-1e-6 on every field (the
  throughput workload's data), default parameters (1+log slicing,
  Gamma-driver shift, eta 2, advected shift);
* gauge: the harmonic gauge wave of amplitude 0.01 along x, one wavelength
  across the domain, harmonic slicing with the shift frozen at zero
  (shift_coeff 0, no advection), an exact solution of the continuum
  equations.

The whole state is far too large for a fixture, so each file keeps the
values of 24 fields at 1024 sampled interior cells (4 random cells drawn in
each depth-3 tile with numpy PCG64 seed 4321, the first 2 of each kept) at
coarse steps 0, 1, 10 and 50, the FP64 constraint norms over the leaves
after each of those steps, and for the gauge wave the exact solution at the
sampled cells after the last step.

Run from the repository root:  python tests/golden/make_bssn512_fp64.py
(about 25 minutes per data set on 8 cores).

With --float-coefficients it writes nothing: it reruns the gauge wave for 10
coarse steps with the FD weights rounded to float (the reference's tables,
AMR.h:1144-1150) and prints, per moving variable, the RMS deviation from the
committed exact-weight run relative to the variable's RMS deviation from
flat space -- the share of the fp32 runs' error that those tables explain.
"""
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from adaptivemesh_amd import workloads as WL  # noqa: E402
from oracle import oracle as orc  # noqa: E402

S, P, ORDER, OUTER, CFL = 16, 3, 4, 64.0, 0.25
T = S + 2 * P
DEPTH = 3
CHECKPOINTS = (0, 1, 10, 50)
PER_TILE = 4  # drawn per tile; KEEP of them are stored
KEEP = 2
GAUGE_A = 0.01
DATASETS = {
    "noise": dict(params=orc.BSSN_DEFAULT),
    "gauge": dict(params=(2.0, 0.0, 1e-4, 0, 1)),
}


def mesh():
    ops, n = WL.uniform_ops(DEPTH)
    t = orc.OracleTree(n, True)
    for a, c in ops:
        t.add_node(a, c)
    return t.export(), n


def samples(e):
    """Flat indices (padded layout) of the sampled interior cells."""
    rng = np.random.default_rng(4321)
    leaves = [s for s, p in enumerate(e["positions"]) if int(p) != 0xFFFFFFFF and int(p) >> 27 == DEPTH]
    idx = []
    for s in leaves:
        c = rng.integers(0, S, size=(PER_TILE, 3))
        for z, y, x in c[:KEEP]:
            idx.append(s * T ** 3 + ((z + P) * T + (y + P)) * T + (x + P))
    return np.asarray(idx, np.int64)


def initial(name, e, n):
    if name == "noise":
        return WL.flat_noise_fields(n, S, P)
    return WL.gauge_wave_fields(e["positions"], S, P, OUTER, amplitude=GAUGE_A)


def run(name, steps=CHECKPOINTS[-1]):
    e, n = mesh()
    idx = samples(e)
    sim = orc.Bssn64Sim(e, OUTER / S, CFL, mode=0, params=DATASETS[name]["params"])
    f0 = initial(name, e, n)
    sim.set_state([a.astype(np.float64) for a in f0])
    sim.fill_ghosts()
    odt = float(np.float32(np.float32(OUTER / S) * np.float32(CFL)))
    vals, cons, exact = [], [], []
    t0 = time.time()
    for k in range(steps + 1):
        if k in CHECKPOINTS:
            vals.append(np.stack([sim.new[v][idx] for v in range(24)]))
            cons.append(np.asarray(sim.constraints()[:3], np.float64))
            if name == "gauge" and k == CHECKPOINTS[-1]:
                ex = WL.gauge_wave_fields(e["positions"], S, P, OUTER, amplitude=GAUGE_A, t=k * odt,
                                          dtype=np.float64)
                exact.append(np.stack([ex[v][idx] for v in range(24)]))
            print(f"{name}: coarse step {k}, rms H {cons[-1][0]:.4e}, {time.time() - t0:.0f} s", flush=True)
        if k == steps:
            break
        for _ in range(1 << DEPTH):
            sim.substep()
    out = dict(idx=idx, checkpoints=np.asarray(CHECKPOINTS[:len(vals)]), values=np.stack(vals),
               constraints=np.stack(cons), initial=np.stack([f0[v][idx] for v in range(24)]).astype(np.float64),
               params=np.asarray(DATASETS[name]["params"], np.float64))
    if exact:
        out["exact"] = exact[-1]
    return out


def float_coefficient_check(steps=10):
    global CHECKPOINTS
    orc.lib().orc_fd64_float_coefficients(1)
    CHECKPOINTS = (0, 1, steps)
    out = run("gauge", steps=steps)
    with np.load(Path(__file__).with_name("bssn512_fp64_gauge.npz"), allow_pickle=False) as z:
        ref_cp, ref_vals = list(z["checkpoints"]), z["values"]
    flat = WL.FLAT.astype(np.float64)[:, None]
    names = "chi gxx gxy gxz gyy gyz gzz K Axx Axy Axz Ayy Ayz Azz Gx Gy Gz alpha bx by bz Bx By Bz".split()
    for ci, cp in enumerate(CHECKPOINTS):
        ref = ref_vals[ref_cp.index(cp)]
        dev = np.sqrt(np.mean((ref - flat) ** 2, axis=1))
        err = np.sqrt(np.mean((out["values"][ci] - ref) ** 2, axis=1))
        print(f"coarse step {cp}: " + " ".join(f"{names[v]}={err[v] / dev[v]:.2e}" for v in range(24)
                                              if dev[v] > 1e-9), flush=True)


if __name__ == "__main__":
    if "--float-coefficients" in sys.argv:
        float_coefficient_check()
        sys.exit(0)
    names = sys.argv[1:] or list(DATASETS)
    for nm in names:
        np.savez_compressed(Path(__file__).with_name(f"bssn512_fp64_{nm}.npz"), **run(nm))
