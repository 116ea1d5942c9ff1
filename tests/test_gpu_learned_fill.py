"""The learned ghost fill on the GPU against the FP64 oracle (oracle/mlp.py
learned_fill), on the puncture octree one sub-step into the run: depth-5
tiles at counter 1 fill their faces from depth-4 targets at stage time
(1 + 1/2) / 2 with quadratic time interpolation.

* With a threshold no residual reaches, the learned cells match the oracle's
  FP64 network to 2e-5 (1 + |value|) (fp32 MFMA over 1536 inputs); every
  other ghost cell -- tricubic cells whose stencil would read the target's
  edge ghosts, equal-depth copies -- is bit-identical; the statistics (64-
  sample blocks, two per face) agree.
* With a negative threshold every block keeps the tricubic values, and the
  whole fill equals the tricubic fill of the float restatement bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from adaptivemesh_amd import workloads as WL
from oracle import mlp as M

pytestmark = pytest.mark.gpu

S, P, ORDER, OUTER, CFL = 16, 3, 4, 64.0, 0.25
T = S + 2 * P


@pytest.fixture(scope="module")
def setup(am, orc, torch_cuda):
    torch = torch_cuda
    ops, n = WL.puncture_ops(OUTER)
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    sim = am.BssnSim(g, n, am.bssn_options(periodic=False, interp=am.AM_INTERP_TRICUBIC))
    tree = am.Tree(n, periodic=False)
    ot = orc.OracleTree(n, False)
    for p_, c in ops:
        sim.add_node(p_, c)
        tree.add_node(p_, c)
        ot.add_node(p_, c)
    e = ot.export()
    st = WL.puncture_fields(e["positions"], S, P, OUTER)
    for v in range(24):
        sim.set_field(0, v, st[v])
        sim.set_field(1, v, st[v])
    sim.fill_ghosts()
    sim.substep()
    sim.sync()
    bufs = {w: [sim.get_field(w, v) for v in range(24)] for w in (0, 1, 3)}
    rng = np.random.default_rng(7)
    params = M.init_params(3)
    params[-M.N_OUT:] = rng.normal(0, 0.05, M.N_OUT)  # visible corrections
    params32 = params.astype(np.float32)
    tiles = [s for s in range(n) if int(e["positions"][s]) != 0xFFFFFFFF and int(e["positions"][s]) >> 27 == 5]
    # faces of those tiles with a coarser neighbour: two 64-sample blocks each
    coarse = 0
    for t in tiles:
        for f in range(6):
            nd = int(e["boundaries"][t * 6 + f]) & 0xFFFFFFFF
            if not (nd & 0x800000) and nd >> 28 not in (0, 15) and (nd >> 24) & 15:
                coarse += 1
    return dict(sim=sim, tree=tree, e=e, n=n, g=g, bufs=bufs, params=params32, tiles=tiles, blocks=2 * coarse)


def run_gpu(am, torch, s, threshold):
    sim, g = s["sim"], s["g"]
    for v in range(24):
        sim.set_field(1, v, s["bufs"][1][v])  # the same input every call
    fb = am.FieldBatch()
    fb.n_vars = 24
    fb.tile_layout = am.AM_LAYOUT_INTERIOR_FIRST  # the driver's state layout
    for v in range(24):
        fb.old_[v] = sim.field_ptr(0, v)
        fb.new_[v] = sim.field_ptr(1, v)
        fb.older[v] = sim.field_ptr(3, v)
    m = s["tree"].mesh(5)
    m.substep = 1
    prm = torch.tensor(s["params"], device="cuda")
    stats = torch.zeros(4, dtype=torch.int64, device="cuda")
    am.check(am.lib().am_ghost_fill_learned(C.byref(fb), C.byref(m), C.byref(g), 0.5, prm.data_ptr(), threshold,
                                            stats.data_ptr(), am.AM_EQUAL_COPY, am.stream_handle()), "learned fill")
    torch.cuda.synchronize()
    return [sim.get_field(1, v) for v in range(24)], tuple(int(x) for x in stats.cpu().numpy())


def run_oracle(orc, s, threshold):
    new = [a.copy() for a in s["bufs"][1]]
    stats, tric, idx = M.learned_fill(s["bufs"][0], new, s["bufs"][3], s["e"]["boundaries"], s["e"]["positions"],
                                      s["tiles"], 1, 5, 0.5, s["params"].astype(np.float64), threshold, S, P, orc)
    return new, tric, stats, idx


def test_learned_fill_matches_oracle(am, orc, torch_cuda, setup):
    print("learned ghost fill, depth-5 puncture tiles, counter 1", flush=True)
    got, gstats = run_gpu(am, torch_cuda, setup, 1e30)
    want, tric, ostats, idx = run_oracle(orc, setup, 1e30)
    print(f"  stats GPU {gstats}, oracle {ostats}", flush=True)
    assert gstats == ostats and ostats[0] == setup["blocks"] > 0 and ostats[1] == 0
    worst = 0.0
    n_learned = 0
    learned = np.zeros(want[0].size, bool)
    learned[idx] = True
    for v in range(24):
        n_learned += int(learned.sum())
        err = np.abs(got[v][learned].astype(np.float64) - want[v][learned].astype(np.float64))
        tol = 2e-5 * (1 + np.abs(want[v][learned].astype(np.float64)))
        if err.size:
            worst = max(worst, float(np.max(err / tol)))
        assert np.all(err <= tol), v
        bad = np.nonzero(~learned & (got[v].view(np.uint32) != want[v].view(np.uint32)))[0]
        if bad.size:
            slot, rem = np.divmod(bad[:6], T ** 3)
            zz, rem = np.divmod(rem, T * T)
            yy, xx = np.divmod(rem, T)
            print(f"  field {v}: {bad.size} mismatches, e.g. slot/z/y/x {list(zip(slot, zz, yy, xx))}, "
                  f"got {got[v][bad[:6]]}, want {want[v][bad[:6]]}, tricubic {tric[v][bad[:6]]}", flush=True)
        assert bad.size == 0, v
    print(f"  {n_learned} learned cell values, worst error / tolerance {worst:.3f}", flush=True)
    assert n_learned > 0


def test_learned_fill_falls_back_to_tricubic(am, orc, torch_cuda, setup):
    got, gstats = run_gpu(am, torch_cuda, setup, -1.0)
    want, tric, ostats, idx = run_oracle(orc, setup, -1.0)
    assert idx.size == 0
    assert gstats == ostats and gstats[0] == 0 and gstats[1] == setup["blocks"] > 0
    for v in range(24):
        assert np.array_equal(got[v].view(np.uint32), tric[v].view(np.uint32)), v


def test_driver_learned_fallback_equals_tricubic(am, orc, torch_cuda, setup):
    """The RK4 driver with the learned prolongation and a threshold every
    tile exceeds runs bit for bit like the tricubic driver over a coarse step
    (the per-tile gate refills every learned face; its statistics count
    them all as fallbacks)."""
    torch = torch_cuda
    ops, n = WL.puncture_ops(OUTER)
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    sims = [am.BssnSim(g, n, am.bssn_options(periodic=False, interp=i))
            for i in (am.AM_INTERP_TRICUBIC, am.AM_INTERP_LEARNED)]
    st = WL.puncture_fields(setup["e"]["positions"], S, P, OUTER)
    prm = torch.tensor(setup["params"], device="cuda")
    sims[1].set_prolongation(prm, -1.0)
    for s in sims:
        for p_, c in ops:
            s.add_node(p_, c)
        for v in range(24):
            s.set_field(0, v, st[v])
            s.set_field(1, v, st[v])
        s.fill_ghosts()
        s.timestep()
        s.sync()
    for w in (0, 1, 2, 3):
        for v in range(24):
            a, b = sims[0].get_field(w, v), sims[1].get_field(w, v)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (w, v)
    stats = sims[1].prolongation_stats()
    print(f"  driver fallback statistics {stats}", flush=True)
    assert stats["learned_faces"] == 0 and stats["fallback_faces"] > 0


def test_learned_faces_gate_per_tile(am, orc, torch_cuda, setup):
    """am_ghost_fill_learned_faces (the driver's call: a face list and the
    gate per tile, 32-sample workgroups) on the depth-5 faces whose target is
    coarser: with an unreachable threshold it writes what
    am_ghost_fill_learned (64-sample workgroups) writes with its per-block
    gate passing every block, with a negative one what it
    writes with none passing (the tricubic fill), bit for bit; with the
    median of the tiles' residuals, each tile's slot equals the first
    result when its residual is within the threshold and the second
    otherwise, and the statistics count its learned faces accordingly."""
    torch = torch_cuda
    s = setup
    sim, g, e = s["sim"], s["g"], s["e"]
    faces, learned_faces = [], {}
    for t in s["tiles"]:
        for f in range(6):
            nd = int(e["boundaries"][t * 6 + f]) & 0xFFFFFFFF
            tgt = nd & 0xFFFFFF
            if tgt & 0x800000 or nd >> 28 in (0, 15) or not (nd >> 24) & 15:
                continue
            faces.append(t * 6 + f)
            if (nd >> 24) & 15 == 1 and tgt > 0:
                learned_faces[t] = learned_faces.get(t, 0) + 1
    dfaces = torch.tensor(np.array(faces, np.int32), device="cuda")
    prm = torch.tensor(s["params"], device="cuda")
    resid = torch.full((s["n"],), float("nan"), dtype=torch.float32, device="cuda")

    def run(threshold, per_tile):
        for v in range(24):
            sim.set_field(1, v, s["bufs"][1][v])
        fb = am.FieldBatch()
        fb.n_vars = 24
        fb.tile_layout = am.AM_LAYOUT_INTERIOR_FIRST
        for v in range(24):
            fb.old_[v], fb.new_[v], fb.older[v] = sim.field_ptr(0, v), sim.field_ptr(1, v), sim.field_ptr(3, v)
        m = s["tree"].mesh(5)
        m.substep = 1
        stats = torch.zeros(4, dtype=torch.int64, device="cuda")
        if per_tile:
            am.check(am.lib().am_ghost_fill_learned_faces(C.byref(fb), C.byref(m), dfaces.data_ptr(), len(faces),
                                                          C.byref(g), 0.5, prm.data_ptr(), threshold,
                                                          resid.data_ptr(), stats.data_ptr(), am.AM_EQUAL_SKIP,
                                                          am.stream_handle()), "learned faces")
        else:
            am.check(am.lib().am_ghost_fill_learned(C.byref(fb), C.byref(m), C.byref(g), 0.5, prm.data_ptr(),
                                                    threshold, stats.data_ptr(), am.AM_EQUAL_SKIP,
                                                    am.stream_handle()), "learned fill")
        torch.cuda.synchronize()
        return [sim.get_field(1, v) for v in range(24)], [int(x) for x in stats.cpu().numpy()]

    every, _ = run(1e30, False)
    none, _ = run(-1.0, False)
    got, st = run(1e30, True)
    r = resid.cpu().numpy()
    n_learned = sum(learned_faces.values())
    assert st[0] == n_learned > 0 and st[1] == 0
    for v in range(24):
        assert np.array_equal(got[v].view(np.uint32), every[v].view(np.uint32)), v
    got, st = run(-1.0, True)
    assert st[0] == 0 and st[1] == n_learned
    for v in range(24):
        assert np.array_equal(got[v].view(np.uint32), none[v].view(np.uint32)), v
    tr = np.array([r[t] for t in learned_faces])
    assert np.all(np.isfinite(tr)) and np.all(tr >= 0)
    thr = float(np.median(tr))
    got, st = run(thr, True)
    keep = [t for t in learned_faces if r[t] <= thr]
    assert 0 < len(keep) < len(learned_faces)
    assert st[0] == sum(learned_faces[t] for t in keep) and st[0] + st[1] == n_learned
    T3 = T ** 3
    for v in range(24):
        a, w_every, w_none = (x.reshape(s["n"], T3) for x in (got[v], every[v], none[v]))
        for t in learned_faces:
            want = w_every[t] if t in keep else w_none[t]
            assert np.array_equal(a[t].view(np.uint32), want.view(np.uint32)), (v, t)
    print(f"  per-tile gate at {thr:.3e}: {len(keep)} of {len(learned_faces)} tiles keep the network", flush=True)
