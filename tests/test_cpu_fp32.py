"""The FP32 host implementation of the scheme (oracle/am_cpu_fp32.c, the
CPU baseline of bench.py) against the FP64 oracle on a periodic mesh of
eight depth-1 tiles with a smooth perturbed state:

* the equal-depth ghost refresh equals the oracle's copy rule
  (copy_boundaries with equal_copy) on every ghost cell it writes;
* one RK4 step, centred and upwinded + Kreiss-Oliger, agrees with the FP64
  scheme run the same way (orc_bssn64_stage + orc_ghost64 between stages)
  to 1e-3 of the largest change of each field (fp32 state and arithmetic
  against FP64: measured 2.3e-4 at most, in K and At_yy, whose right-hand
  sides difference second derivatives of the O(1) metric).
"""
import numpy as np
import pytest

from oracle import cpu_fp32 as CPU
from oracle import oracle as orc
from tests import fields as F
from tests import meshes

S, P, Q, OUTER, CFL = 16, 3, 2, 64.0, 0.25
T = S + 2 * P


def setup():
    ops, n = meshes.uniform(1)
    ot = orc.OracleTree(n, True)
    meshes.apply(ot, ops)
    e = ot.export()
    tiles = np.array([s for s in F.used_slots(e) if F.depth_of(e, s) == 1], np.int32)
    flat = np.zeros(24)
    flat[[0, 1, 4, 6, 17]] = 1.0
    k = 2 * np.pi / OUTER  # periodic on the domain
    st = [F.sample(e["positions"], n, S, P, OUTER,
                   lambda x, y, z, v=v: flat[v] + 1e-3 * (1 + v % 5) * np.sin(k * (x + 2 * y + 3 * z) + v),
                   slots=list(tiles), dtype=np.float64) for v in range(24)]
    return e, n, tiles, st


def test_equal_ghosts_match_copy_rule():
    e, n, tiles, st = setup()
    rng = np.random.default_rng(0)
    f = [rng.standard_normal(n * T ** 3).astype(np.float32) for _ in range(3)]
    want = [x.copy() for x in f]
    for x in want:
        orc.copy_boundaries(x.copy(), x, None, e["boundaries"], list(tiles), 0, e["max_depth"], True, 0, S, P,
                            equal_copy=True)
    got = [x.copy() for x in f]
    CPU.lib().cpu32_equal_ghosts(CPU.ptrs(got), 3, e["boundaries"].ctypes.data, tiles.ctypes.data, len(tiles), S, P)
    changed = np.zeros(n * T ** 3, bool)
    for v in range(3):
        changed |= got[v] != f[v]
        assert np.array_equal(got[v].view(np.uint32), want[v].view(np.uint32)), v
    assert changed.sum() == len(tiles) * 6 * P * S * S


@pytest.mark.parametrize("upwind", [False, True])
def test_rk4_step_tracks_fp64(upwind):
    e, n, tiles, st = setup()
    outer_dx = OUTER / S
    inv_dx = 2.0 / outer_dx
    dt = outer_dx * CFL / 2
    sigma = 0.3 if upwind else 0.0
    params = np.array(orc.BSSN_DEFAULT, np.float64)
    # FP64: state with filled ghosts, four stages, ghosts of each stage output refreshed
    y32 = [x.astype(np.float32) for x in st]
    for x in y32:
        orc.copy_boundaries(x.copy(), x, None, e["boundaries"], list(tiles), 0, e["max_depth"], True, 0, S, P,
                            equal_copy=True)
    L = orc.lib()
    L.orc_bssn64_set_dissipation(sigma, 1 if upwind else 0)
    try:
        acc = [np.zeros(n * T ** 3) for _ in range(24)]
        cur = [x.astype(np.float64) for x in y32]
        y0 = [x.copy() for x in cur]
        inv = np.full(9, inv_dx)
        dtd = np.full(9, dt)
        for stage in range(4):
            new = [x.copy() for x in cur]
            L.orc_bssn64_stage(orc.ptrs(cur), orc.ptrs(y0), orc.ptrs(acc), orc.ptrs(new), e["positions"].ctypes.data,
                               tiles.ctypes.data, len(tiles), 0, e["max_depth"], stage, inv.ctypes.data,
                               dtd.ctypes.data, params.ctypes.data, S, P, Q)
            L.orc_ghost64(orc.ptrs(new), orc.ptrs(new), None, 24, e["boundaries"].ctypes.data, tiles.ctypes.data,
                          len(tiles), 0, e["max_depth"], 1.0, 0, S, P, 1)
            cur = new
    finally:
        L.orc_bssn64_set_dissipation(0.0, 0)
    a = [x.copy() for x in y32]
    b = [x.copy() for x in y32]
    acc32 = [np.zeros(n * T ** 3, np.float32) for _ in range(24)]
    CPU.rk4_step(y32, a, b, acc32, e["boundaries"], tiles, S, P, Q, inv_dx, dt, params, sigma, upwind)
    mask = np.zeros((n, T, T, T), bool)
    mask[tiles, P:P + S, P:P + S, P:P + S] = True
    mask = mask.reshape(-1)
    worst = 0.0
    for v in range(24):
        d64 = cur[v][mask] - y0[v][mask]
        d32 = b[v][mask].astype(np.float64) - y0[v][mask]
        scale = max(np.max(np.abs(d64)), 1e-12)
        err = np.max(np.abs(d32 - d64)) / scale
        worst = max(worst, err)
        assert err < 1e-3, (v, err)
    print(f"upwind={upwind}: worst error / largest change {worst:.2e}")
