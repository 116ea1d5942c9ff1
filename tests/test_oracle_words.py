"""Pins the oracle's handling of the precision study's word formats.

* differentiate_f32 / mixed_partials_f32 are the values the cf3 entry
  points pack: re-packing them with the per-thread dither seeds (the
  differentiate order, Hessian diagonal before gradient) reproduces the cf3
  words bit for bit, on a graded mesh at two depths.
* The format-aware stage and constraint evaluators decode fp32 and fp16
  triples: words whose cf3 values are exactly representable in fp32 and in
  fp16 (8-bit mantissas, exponents inside fp16's range, subnormals included)
  give bit-identical RK4 stage outputs and constraint sums in all three
  formats -- which also checks the IEEE binary16 decoder and the Hessian
  assembly order of every format.
"""
import numpy as np

from tests import fields, meshes

S, P, OUTER = 16, 3, 16.0
T = S + 2 * P
HESS = (0, 1, 2, 3, 4, 5, 6, 17, 18, 19, 20)


def test_fd_f32_values_repack_to_cf3_words(orc):
    ops, n = meshes.graded()
    t = orc.OracleTree(n, False)
    meshes.apply(t, ops)
    e = t.export()
    rng = np.random.default_rng(3)
    f = rng.standard_normal(n * T ** 3).astype(np.float32)
    tiles = fields.used_slots(e)
    gs, gm = 123457, -98765
    w1, w2 = orc.differentiate(f, e["positions"], tiles, 0, e["max_depth"], gs, S, P, 2, 1.0, n)
    wo = orc.mixed_partials(f, e["positions"], tiles, 0, e["max_depth"], gm, S, P, 1.0, n)
    v1, v2 = orc.differentiate_f32(f, e["positions"], tiles, 0, e["max_depth"], S, P, 2, 1.0, n)
    vo = orc.mixed_partials_f32(f, e["positions"], tiles, 0, e["max_depth"], S, P, 1.0, n)
    checked = 0
    for slot in tiles[:3]:
        for z in (0, 5, S - 1):
            for y in (0, 7, S - 1):
                for x in range(S):
                    o = ((slot * S + z) * S + y) * S + x
                    s = orc.thread_seed(gs, x, y, slot * S + z)
                    a, s = orc.cf3_pack_dither(*map(float, v2[o]), s)
                    b, s = orc.cf3_pack_dither(*map(float, v1[o]), s)
                    c, _ = orc.cf3_pack_dither(*map(float, vo[o]), orc.thread_seed(gm, x, y, slot * S + z))
                    assert (a, b, c) == (int(w2[o]), int(w1[o]), int(wo[o])), (slot, z, y, x)
                    checked += 1
    assert checked == 3 * 9 * S


def test_stage_and_constraints_equal_across_formats(orc):
    ops, n = meshes.uniform(1)
    t = orc.OracleTree(n, True)
    meshes.apply(t, ops)
    e = t.export()
    rng = np.random.default_rng(11)
    W = n * S ** 3
    flat = np.zeros(24)
    flat[[0, 1, 4, 6, 17]] = 1.0
    state = [(flat[v] + 0.05 * rng.standard_normal(n * T ** 3)).astype(np.float32) for v in range(24)]
    old = [(x + np.float32(1e-3)).astype(np.float32) for x in state]

    def words():
        # exponent field 0..24: values m * 2^(e-22) with m < 256 fit fp16 (max 255 * 4 = 1020);
        # e <= 7 keeps some below fp16's smallest normal 2^-14
        e_ = rng.integers(0, 25, W).astype(np.uint32)
        return (e_ << 27) | rng.integers(0, 1 << 27, W, dtype=np.uint32)

    sets = {"d1": [words() for _ in range(24)], "d2d": [words() for _ in range(11)],
            "d2o": [words() for _ in range(11)]}
    as_fmt = {0: lambda w: w, 1: lambda w: orc.cf3_unpack(w).astype(np.float32),
              2: lambda w: orc.cf3_unpack(w).astype(np.float16)}
    dec = orc.cf3_unpack(sets["d1"][0])
    assert np.array_equal(dec.astype(np.float16).astype(np.float32), dec)  # exact in fp16
    assert np.any((dec != 0) & (np.abs(dec) < 2.0 ** -14))  # fp16 subnormals present
    tiles = np.arange(1, 9, dtype=np.int32)
    dtd = np.array([0.25 / 2 ** d for d in range(9)])
    params = np.array(orc.BSSN_DEFAULT, np.float64)
    acc_in = [rng.standard_normal(n * T ** 3) for _ in range(24)]
    results = []
    for fmt in (0, 1, 2):
        arrs = {k: [np.ascontiguousarray(as_fmt[fmt](w)) for w in v] for k, v in sets.items()}
        acc_o = [np.zeros(n * T ** 3) for _ in range(24)]
        new_o = [np.zeros(n * T ** 3) for _ in range(24)]
        orc.lib().orc_bssn_stage_from_words_fmt(orc.ptrs(state), orc.ptrs(old), orc.ptrs(acc_in), orc.ptrs(acc_o),
                                                orc.ptrs(new_o), orc.ptrs(arrs["d1"]), orc.ptrs(arrs["d2d"]),
                                                orc.ptrs(arrs["d2o"]), e["positions"].ctypes.data, tiles.ctypes.data,
                                                len(tiles), 0, e["max_depth"], 1, dtd.ctypes.data,
                                                params.ctypes.data, S, P, fmt)
        cons = np.zeros(4 * len(tiles))
        orc.lib().orc_bssn_constraints_from_words_fmt(orc.ptrs(state), orc.ptrs(arrs["d1"]), orc.ptrs(arrs["d2d"]),
                                                      orc.ptrs(arrs["d2o"]), tiles.ctypes.data, len(tiles),
                                                      params.ctypes.data, S, P, cons.ctypes.data, fmt)
        results.append((acc_o, new_o, cons))
    # the cf3 entry point agrees with the format-aware one at fmt 0
    acc0 = [np.zeros(n * T ** 3) for _ in range(24)]
    new0 = [np.zeros(n * T ** 3) for _ in range(24)]
    orc.lib().orc_bssn_stage_from_words(orc.ptrs(state), orc.ptrs(old), orc.ptrs(acc_in), orc.ptrs(acc0),
                                        orc.ptrs(new0), orc.ptrs(sets["d1"]), orc.ptrs(sets["d2d"]),
                                        orc.ptrs(sets["d2o"]), None, e["positions"].ctypes.data, tiles.ctypes.data,
                                        len(tiles), 0, e["max_depth"], 1, dtd.ctypes.data, params.ctypes.data, S, P)
    for v in range(24):
        assert np.array_equal(acc0[v], results[0][0][v]) and np.array_equal(new0[v], results[0][1][v])
    for fmt in (1, 2):
        for v in range(24):
            assert np.array_equal(results[fmt][0][v], results[0][0][v]), (fmt, v)
            assert np.array_equal(results[fmt][1][v], results[0][1][v]), (fmt, v)
        assert np.array_equal(results[fmt][2], results[0][2]), fmt
    assert np.all(np.isfinite(results[0][2])) and results[0][2][0] > 0
