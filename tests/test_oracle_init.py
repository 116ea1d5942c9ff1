"""The wave test's initial data (__init_temp, kernel.cu:9-43) in the oracle,
pinned by the closed form: at every written cell the float restatement is
within float rounding of f = 10 / (|x - a|^2 + 10) - 10 / (|x + a|^2 + 10),
a = (-25, 1/2, 0), evaluated in float64 at the cell positions of the initial
data convention (kernel.cu:21-22, tests/fields.py); depth-0 tiles are written
everywhere, deeper tiles on their interior only."""
import numpy as np
import pytest

from tests import fields, meshes


def bumps(x, y, z):
    return 10 / ((x + 25) ** 2 + (y - .5) ** 2 + z * z + 10) - 10 / ((x - 25) ** 2 + (y + .5) ** 2 + z * z + 10)


@pytest.mark.parametrize("S,P", [(16, 3), (28, 2)])
def test_init_temp_closed_form(orc, S, P):
    T = S + 2 * P
    t = orc.OracleTree(25)
    ops, n = meshes.graded()
    meshes.apply(t, ops)
    e = t.export()
    outer = 100.0
    dx = float(np.float32(np.float32(outer) / np.float32(S)))
    old, new = orc.init_temp(e["positions"], fields.used_slots(e), n, S, P, outer, dx, 2.0)
    assert np.array_equal(old.view(np.uint32), new.view(np.uint32))
    want = 2.0 * fields.sample(e["positions"], n, S, P, outer, bumps, dtype=np.float64)
    o = old.reshape(n, T, T, T)
    w = want.reshape(n, T, T, T)
    for s in fields.used_slots(e):
        d = fields.depth_of(e, s)
        if d == 0:
            assert np.max(np.abs(o[s] - w[s])) < 2e-6
        else:
            inner = (slice(P, P + S),) * 3
            assert np.max(np.abs(o[s][inner] - w[s][inner])) < 2e-6, s
            ghost = np.ones((T, T, T), bool)
            ghost[inner] = False
            assert not o[s][ghost].any()
