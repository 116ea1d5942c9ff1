"""The BSSN RK4 driver on the GPU against the oracle, stage by stage.

For every stage the oracle is given the GPU's stage input, and then:
* the FD words must equal the oracle's bit for bit (same seeds, drawn in
  the driver's order);
* the fp32 RHS + RK4 update must match the FP64 oracle evaluated on the same
  words within 5e-5 of the largest |k| of each field (fp32 round-off of a
  ~1.5 kflop expression; the words themselves are identical; both raise the
  conformal Christoffel symbols from the metric gradient words);
* the ghost cells written at the stage time (equal-depth faces pushed by the
  RHS kernel as exact copies, coarser faces interpolated) must equal the
  oracle's fill of the GPU's own new interior bit for bit.
The constraint reduction (FP64) must match the oracle's sums over the same
words to 1e-4.
"""
import ctypes as C

import numpy as np
import pytest

from tests import fields as F
from tests import meshes

pytestmark = pytest.mark.gpu

S, P, ORDER, OUTER, CFL = 16, 3, 4, 16.0, 0.25
T = S + 2 * P
HESS = (0, 1, 2, 3, 4, 5, 6, 17, 18, 19, 20)
STAGE_FRAC = (0.5, 0.5, 1.0, 1.0)


def smooth_state(e, n, seed=0, eps=0.04, S=S, P=P):
    """Flat space plus smooth periodic perturbations on every field."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(24)
    flat[[0, 1, 4, 6, 17]] = 1.0
    out = []
    L = OUTER
    for v in range(24):
        k = rng.integers(1, 3, size=3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        a = eps * rng.uniform(0.5, 1.0)
        fn = (lambda x, y, z, v=v, k=k, ph=ph, a=a:
              flat[v] + a * np.sin(2 * np.pi * k[0] * x / L + ph[0]) * np.cos(2 * np.pi * k[1] * y / L + ph[1])
              + 0.5 * a * np.sin(2 * np.pi * k[2] * z / L + ph[2]))
        out.append(F.sample(e["positions"], n, S, P, OUTER, fn))
    return out


def download(am, dev_ptr, count, dtype, offset_elems=0):
    a = np.empty(count, dtype)
    itemsize = np.dtype(dtype).itemsize
    am.check(am.lib().am_memcpy_d2h(a.ctypes.data, dev_ptr + offset_elems * itemsize, count * itemsize, None), "d2h")
    return a


def sim_words(am, sim, n, S=S):
    W = n * S ** 3
    d1 = [download(am, sim.device_ptr(4), W, np.uint32, v * W) for v in range(24)]
    d2d = [download(am, sim.device_ptr(5), W, np.uint32, h * W) for h in range(11)]
    d2o = [download(am, sim.device_ptr(6), W, np.uint32, h * W) for h in range(11)]
    return d1, d2d, d2o


def tile_mask(n, tiles, inner=True, S=S, P=P):
    T = S + 2 * P
    m = np.zeros((n, T, T, T) if not inner else (n, S, S, S), bool)
    m[tiles] = True
    return m.reshape(-1)


def interior_mask(n, tiles, S=S, P=P):
    T = S + 2 * P
    m = np.zeros((n, T, T, T), bool)
    for t in tiles:
        m[t, P:P + S, P:P + S, P:P + S] = True
    return m.reshape(-1)


def fd_input(orc, e, n, tiles, state_in, old, stage, direct, S=S, P=P):
    return F.fd_input(orc, e, n, tiles, state_in, old, stage, direct, S, P, STAGE_FRAC[stage])


@pytest.mark.parametrize("recipe,periodic,depth,layers,direct,S,P", [
    ("uniform1", True, 1, 0, 4, S, P), ("graded", False, 2, 0, 4, S, P), ("graded", False, 3, 0, 4, S, P),
    ("uniform1", True, 1, P, 4, S, P), ("graded", False, 2, 0, 0, S, P), ("uniform1", True, 1, P, 0, S, P),
    ("graded", False, 3, P, 0, S, P), ("graded", False, 3, P, 4, S, P), ("uniform1", True, 1, 0, 6, S, P),
    ("graded", False, 3, 0, 6, S, P), ("graded", False, 2, P, 2, S, P),
    ("uniform1", True, 1, 0, 4, 28, 2), ("graded", False, 2, 0, 0, 28, 2), ("graded", False, 3, 0, 4, 28, 2),
    ("graded", False, 3, 0, 6, 28, 2)])
def test_bssn_stages_match_oracle(am, orc, torch_cuda, recipe, periodic, depth, layers, direct, S, P):
    """Also at the reference's default tile, S = 28, P = 2 (1/S inexact in
    float; every ghost layer pushed, the x faces by the general push)."""
    print(f"bssn stages {recipe} periodic={periodic} depth={depth} push layers={layers} direct={direct} S={S} "
          f"P={P}", flush=True)
    T = S + 2 * P
    ops, n = {"uniform1": meshes.uniform(1), "graded": meshes.graded()}[recipe]
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    sim = am.BssnSim(g, n, am.bssn_options(periodic=periodic))
    sim.set_push_layers(layers)
    assert sim.direct_faces == 4
    sim.set_direct_faces(direct)
    assert sim.direct_faces == direct
    ot = orc.OracleTree(n, periodic)
    for p_, c in ops:
        assert sim.add_node(p_, c) == ot.add_node(p_, c)
    e = ot.export()
    state0 = smooth_state(e, n, seed=depth, S=S, P=P)
    for v in range(24):
        sim.set_field(0, v, state0[v])
        sim.set_field(1, v, state0[v])
    tiles = [s for s in F.used_slots(e) if F.depth_of(e, s) == depth]
    assert tiles
    outer_dx = float(np.float32(OUTER) / np.float32(S))
    dtd = np.array([float(np.float32(np.float32(outer_dx) * np.float32(CFL)) / np.float32(2 ** d)) for d in range(9)])
    params = np.array(orc.BSSN_DEFAULT, np.float64)
    imask = interior_mask(n, tiles, S, P)
    wmask = tile_mask(n, tiles, True, S, P)
    old = [x.copy() for x in state0]
    for stage in range(4):
        state_in = old if stage == 0 else [sim.get_field(1, v) for v in range(24)]
        state_fd = fd_input(orc, e, n, tiles, state_in, old, stage, direct, S, P)
        acc_in = [sim.get_field(2, v).astype(np.float64) for v in range(24)]
        seed = sim.rng_state
        sim.stage_depth(depth, stage)
        sim.sync()
        d1, d2d, d2o = sim_words(am, sim, n, S)
        # words: bit-exact with the oracle on the same input and seeds
        s_ = seed
        h = 0
        for v in range(24):
            sd = orc.prng_ints(s_, 1)[0]
            s_ = sd
            hess = v in HESS
            w1, w2 = orc.differentiate(state_fd[v], e["positions"], tiles, 0, e["max_depth"], sd, S, P, ORDER // 2,
                                       outer_dx, n, with_d2=hess)
            assert np.array_equal(d1[v][wmask], w1[wmask]), (stage, v)
            if hess:
                sm = orc.prng_ints(s_, 1)[0]
                s_ = sm
                wo = orc.mixed_partials(state_fd[v], e["positions"], tiles, 0, e["max_depth"], sm, S, P, outer_dx, n)
                assert np.array_equal(d2d[h][wmask], w2[wmask]), (stage, v)
                assert np.array_equal(d2o[h][wmask], wo[wmask]), (stage, v)
                h += 1
        assert sim.rng_state == s_
        # RHS + update in FP64 on the GPU's words
        acc_o = [np.zeros(n * T ** 3) for _ in range(24)]
        new_o = [np.zeros(n * T ** 3) for _ in range(24)]
        sin32 = [np.ascontiguousarray(x, np.float32) for x in state_in]
        tl = np.array(tiles, np.int32)  # alive through the call
        orc.lib().orc_bssn_stage_from_words(orc.ptrs(sin32), orc.ptrs(old), orc.ptrs(acc_in), orc.ptrs(acc_o),
                                            orc.ptrs(new_o), orc.ptrs(d1), orc.ptrs(d2d), orc.ptrs(d2o),
                                            None, e["positions"].ctypes.data,
                                            tl.ctypes.data, len(tiles), 0, e["max_depth"],
                                            stage, dtd.ctypes.data, params.ctypes.data, S, P)
        new_g = [sim.get_field(1, v) for v in range(24)]
        acc_g = [sim.get_field(2, v) for v in range(24)]
        worst = 0.0
        for v in range(24):
            if stage < 3:
                k_o = (acc_o[v] - acc_in[v])[imask] / (1 if stage == 0 else 2)
                k_g = (acc_g[v].astype(np.float64) - acc_in[v])[imask] / (1 if stage == 0 else 2)
                scale = max(np.max(np.abs(k_o)), 1e-3)
                err = np.max(np.abs(k_g - k_o)) / scale
                worst = max(worst, err)
                assert err < 5e-5, (stage, v, err)
            dn = np.max(np.abs(new_g[v][imask] - new_o[v][imask]))
            assert dn < 1e-6 * max(1.0, np.max(np.abs(new_o[v][imask]))), (stage, v, dn)
        print(f"  stage {stage}: worst relative k error {worst:.2e}", flush=True)
        # ghost cells of the stage output: the pushed layers (the FD
        # half-width) are exact copies, the outer equal-depth layers keep what
        # they held; with direct faces stages 0-2 push nothing
        if depth > 0:
            ref = [x.copy() for x in new_g]
            keep = F.unpushed_layers(e, n, tiles, S, P, sim.push_layers, direct, stage)
            assert sim.push_layers == (layers or ORDER // 2)
            for v in range(24):
                orc.copy_boundaries(old[v].copy(), ref[v], None, e["boundaries"], tiles, 0, e["max_depth"], True, 0,
                                    S, P, STAGE_FRAC[stage], equal_copy=True)
                ref[v][keep] = state_in[v][keep]
                assert np.array_equal(ref[v].view(np.uint32), new_g[v].view(np.uint32)), (stage, v)


def test_bssn_constraints_match_oracle(am, orc, torch_cuda):
    ops, n = meshes.uniform(1)
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    sim = am.BssnSim(g, n, am.bssn_options(periodic=True))
    ot = orc.OracleTree(n, True)
    for p_, c in ops:
        assert sim.add_node(p_, c) == ot.add_node(p_, c)
    e = ot.export()
    st = smooth_state(e, n, seed=9)
    for v in range(24):
        sim.set_field(0, v, st[v])
        sim.set_field(1, v, st[v])
    cons = sim.constraints()
    d1, d2d, d2o = sim_words(am, sim, n)
    leaves = np.array(list(range(1, 9)), np.int32)
    out = np.zeros(4 * len(leaves))
    params = np.array(orc.BSSN_DEFAULT, np.float64)
    orc.lib().orc_bssn_constraints_from_words(orc.ptrs(st), orc.ptrs(d1), orc.ptrs(d2d), orc.ptrs(d2o), None,
                                              leaves.ctypes.data, len(leaves), params.ctypes.data, S, P,
                                              out.ctypes.data)
    q = out.reshape(-1, 4)
    h_l2, m_l2 = np.sqrt(q[:, 0].sum() / q[:, 2].sum()), np.sqrt(q[:, 1].sum() / q[:, 2].sum())
    assert cons["cells"] == q[:, 2].sum()
    assert abs(cons["H_l2"] - h_l2) <= 1e-4 * h_l2, (cons, h_l2)
    assert abs(cons["M_l2"] - m_l2) <= 1e-4 * m_l2, (cons, m_l2)
    assert abs(cons["H_max"] - q[:, 3].max()) <= 1e-4 * q[:, 3].max()


def test_bssn_flat_space_stays_flat(am, torch_cuda):
    """Flat data stays flat to round-off.  Not bit-identically: the
    reference's float second-derivative coefficients (AMR.h:1148-1150) do
    not sum to exactly zero, so d^2 of a constant metric component is a few
    1e-7 inv_dx^2 rather than 0, and that feeds the Ricci tensor."""
    ops, n = meshes.uniform(1)
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    sim = am.BssnSim(g, n, am.bssn_options(periodic=True))
    meshes.apply(sim, ops)
    flat = np.zeros(24, np.float32)
    flat[[0, 1, 4, 6, 17]] = 1.0
    for v in range(24):
        a = np.full(n * T ** 3, flat[v], np.float32)
        sim.set_field(0, v, a)
        sim.set_field(1, v, a)
    for _ in range(2):
        sim.timestep()
    sim.sync()
    for v in range(24):
        got = sim.get_field(1, v).reshape(n, T, T, T)[:, P:P + S, P:P + S, P:P + S]
        assert np.max(np.abs(got - flat[v])) < 1e-5, v
    c = sim.constraints()
    assert c["H_l2"] < 1e-4 and c["M_l2"] < 1e-4, c


@pytest.mark.parametrize("recipe,periodic", [("graded", False), ("uniform1", True)])
def test_direct_faces_leave_every_step_unchanged(am, torch_cuda, recipe, periodic):
    """Whole coarse steps (graph replays) with the direct-face masks 0, 2, 4
    and 6: old, new and the accumulator, ghost cells included, and the seed
    state are the same bit for bit after every step.  Masks with x faces are
    refused (the driver's interior-first tiles give those ghost cells only as
    strided columns)."""
    print(f"direct faces over whole steps: {recipe} periodic={periodic}", flush=True)
    ops, n = {"uniform1": meshes.uniform(1), "graded": meshes.graded()}[recipe]
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    results = {}
    for mask in (0, 2, 4, 6):
        sim = am.BssnSim(g, n, am.bssn_options(periodic=periodic))
        meshes.apply(sim, ops)
        for bad in (1, 3, 5, 7, 8, -2):
            with pytest.raises(am.AMError):
                sim.set_direct_faces(bad)
        sim.set_direct_faces(mask)
        e = sim.export()
        st = smooth_state(e, n, seed=11)
        for v in range(24):
            sim.set_field(0, v, st[v])
            sim.set_field(1, v, st[v])
        sim.fill_ghosts()
        sim.use_graph(True)
        snaps = []
        for _ in range(2):
            sim.timestep()
            sim.sync()
            snaps.append(([sim.get_field(w, v).view(np.uint32) for w in (0, 1, 2) for v in range(24)],
                          sim.rng_state))
        results[mask] = snaps
    for mask in (2, 4, 6):
        for step, ((fa, sa), (fb, sb)) in enumerate(zip(results[0], results[mask])):
            assert sa == sb, (mask, step)
            for i, (a, b) in enumerate(zip(fa, fb)):
                assert np.array_equal(a, b), (mask, step, i // 24, i % 24)


def sim_word_bytes(am, sim, n, fmt):
    """The driver's d1 / d2diag / d2off arrays as raw bytes per field."""
    wb = am.lib().am_word_bytes(fmt)
    W = n * S ** 3 * wb
    get = lambda which, k: [download(am, sim.device_ptr(which), W, np.uint8, i * W) for i in range(k)]
    return get(4, 24), get(5, 11), get(6, 11)


@pytest.mark.parametrize("fmt", [1, 2])
def test_bssn_word_formats_match_oracle(am, orc, torch_cuda, fmt):
    """fp32 / fp16 derivative words (the precision study's formats) through
    two RK4 stages of a depth-1 periodic mesh: the words equal the oracle's
    undithered fp32 stencils bit for bit (fp16: those values rounded to
    nearest even by numpy); the RHS + update match the FP64 oracle on the
    same words to the cf3 path's bar; the constraint sums match to 1e-4."""
    print(f"bssn word format {fmt}", flush=True)
    ops, n = meshes.uniform(1)
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    sim = am.BssnSim(g, n, am.bssn_options(periodic=True, word_format=fmt))
    ot = orc.OracleTree(n, True)
    for p_, c in ops:
        assert sim.add_node(p_, c) == ot.add_node(p_, c)
    e = ot.export()
    state0 = smooth_state(e, n, seed=5)
    for v in range(24):
        sim.set_field(0, v, state0[v])
        sim.set_field(1, v, state0[v])
    tiles = [s for s in F.used_slots(e) if F.depth_of(e, s) == 1]
    outer_dx = float(np.float32(OUTER) / np.float32(S))
    dtd = np.array([float(np.float32(np.float32(outer_dx) * np.float32(CFL)) / np.float32(2 ** d)) for d in range(9)])
    params = np.array(orc.BSSN_DEFAULT, np.float64)
    imask = interior_mask(n, tiles, S, P)
    wmask = tile_mask(n, tiles, True, S, P)
    old = [x.copy() for x in state0]
    as_words = (lambda a: a.view(np.float32).reshape(-1, 3)) if fmt == 1 else \
        (lambda a: a.view(np.float16).reshape(-1, 3))
    to_fmt = (lambda x: x) if fmt == 1 else (lambda x: x.astype(np.float16))
    for stage in range(2):
        state_in = old if stage == 0 else [sim.get_field(1, v) for v in range(24)]
        state_fd = fd_input(orc, e, n, tiles, state_in, old, stage, sim.direct_faces)
        acc_in = [sim.get_field(2, v).astype(np.float64) for v in range(24)]
        sim.stage_depth(1, stage)
        sim.sync()
        d1, d2d, d2o = sim_word_bytes(am, sim, n, fmt)
        h = 0
        for v in range(24):
            hess = v in HESS
            r1, r2 = orc.differentiate_f32(state_fd[v], e["positions"], tiles, 0, e["max_depth"], S, P, ORDER // 2,
                                           outer_dx, n, with_d2=hess)
            assert np.array_equal(as_words(d1[v])[wmask].view(np.uint8), to_fmt(r1[wmask]).view(np.uint8)), \
                (stage, v)
            if hess:
                ro = orc.mixed_partials_f32(state_fd[v], e["positions"], tiles, 0, e["max_depth"], S, P, outer_dx, n)
                assert np.array_equal(as_words(d2d[h])[wmask].view(np.uint8), to_fmt(r2[wmask]).view(np.uint8))
                assert np.array_equal(as_words(d2o[h])[wmask].view(np.uint8), to_fmt(ro[wmask]).view(np.uint8))
                h += 1
        acc_o = [np.zeros(n * T ** 3) for _ in range(24)]
        new_o = [np.zeros(n * T ** 3) for _ in range(24)]
        sin32 = [np.ascontiguousarray(x, np.float32) for x in state_in]
        tl = np.array(tiles, np.int32)  # alive through the call
        orc.lib().orc_bssn_stage_from_words_fmt(orc.ptrs(sin32), orc.ptrs(old), orc.ptrs(acc_in), orc.ptrs(acc_o),
                                                orc.ptrs(new_o), orc.ptrs(d1), orc.ptrs(d2d), orc.ptrs(d2o),
                                                e["positions"].ctypes.data, tl.ctypes.data,
                                                len(tiles), 0, e["max_depth"], stage, dtd.ctypes.data,
                                                params.ctypes.data, S, P, fmt)
        new_g = [sim.get_field(1, v) for v in range(24)]
        acc_g = [sim.get_field(2, v) for v in range(24)]
        worst = 0.0
        for v in range(24):
            k_o = (acc_o[v] - acc_in[v])[imask] / (1 if stage == 0 else 2)
            k_g = (acc_g[v].astype(np.float64) - acc_in[v])[imask] / (1 if stage == 0 else 2)
            scale = max(np.max(np.abs(k_o)), 1e-3)
            err = np.max(np.abs(k_g - k_o)) / scale
            worst = max(worst, err)
            assert err < 5e-5, (stage, v, err)
            dn = np.max(np.abs(new_g[v][imask] - new_o[v][imask]))
            assert dn < 1e-6 * max(1.0, np.max(np.abs(new_o[v][imask]))), (stage, v, dn)
        print(f"  stage {stage}: worst relative k error {worst:.2e}", flush=True)
    # constraints of the stage-1 output, on its own words
    cons = sim.constraints()
    st = [sim.get_field(1, v) for v in range(24)]
    d1, d2d, d2o = sim_word_bytes(am, sim, n, fmt)
    out = np.zeros(4 * len(tiles))
    tl = np.array(tiles, np.int32)  # alive through the call
    orc.lib().orc_bssn_constraints_from_words_fmt(orc.ptrs(st), orc.ptrs(d1), orc.ptrs(d2d), orc.ptrs(d2o),
                                                  tl.ctypes.data, len(tiles),
                                                  params.ctypes.data, S, P, out.ctypes.data, fmt)
    q = out.reshape(-1, 4)
    h_l2, m_l2 = np.sqrt(q[:, 0].sum() / q[:, 2].sum()), np.sqrt(q[:, 1].sum() / q[:, 2].sum())
    assert cons["cells"] == q[:, 2].sum()
    assert abs(cons["H_l2"] - h_l2) <= 1e-4 * h_l2, (cons, h_l2)
    assert abs(cons["M_l2"] - m_l2) <= 1e-4 * m_l2, (cons, m_l2)


def test_word_format_rejected_off_geometry(am, torch_cuda):
    """fp32 / fp16 words exist for S=16, P=3, order-4 tiles only: other
    geometries and unknown formats are refused at creation."""
    g = am.geometry(16, 2, 4, OUTER, CFL)
    with pytest.raises(am.AMError):
        am.BssnSim(g, 9, am.bssn_options(periodic=True, word_format=1))
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    with pytest.raises(am.AMError):
        am.BssnSim(g, 9, am.bssn_options(periodic=True, word_format=7))


def test_push_layers_lower_bound(am, torch_cuda):
    """The push must refresh every ghost layer the FD stencils read
    (fd_order / 2): fewer layers are refused and leave the setting as it
    was; 0 (the default, fd_order / 2) and fd_order / 2 .. P are taken."""
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    sim = am.BssnSim(g, 9, am.bssn_options(periodic=True))
    assert sim.push_layers == ORDER // 2
    for bad in range(1, ORDER // 2):
        with pytest.raises(am.AMError):
            sim.set_push_layers(bad)
        assert sim.push_layers == ORDER // 2
    for ok in range(ORDER // 2, P + 1):
        sim.set_push_layers(ok)
        assert sim.push_layers == ok
    with pytest.raises(am.AMError):
        sim.set_push_layers(P + 1)
    sim.set_push_layers(0)
    assert sim.push_layers == ORDER // 2
