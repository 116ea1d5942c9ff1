"""am_init_temp (kernel.cu:9-43) on the GPU against the oracle, bit for bit,
through the C ABI and through the SimpleWave driver."""
import ctypes as C

import numpy as np
import pytest

from tests import fields, meshes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("S,P", [(16, 3), (28, 2)])
def test_init_temp_bit_exact(am, orc, torch_cuda, S, P):
    torch = torch_cuda
    print(f"init_temp S={S} P={P}", flush=True)
    T = S + 2 * P
    g = am.geometry(S, P, 2 * P, 100.0, 0.4)
    ops, n = meshes.graded()
    tree, ot = am.Tree(n), orc.OracleTree(n)
    meshes.apply(tree, ops)
    meshes.apply(ot, ops)
    e = ot.export()
    old = torch.full((n * T ** 3,), 7.0, dtype=torch.float32, device="cuda")
    new = torch.full((n * T ** 3,), 7.0, dtype=torch.float32, device="cuda")
    m = tree.mesh(-1)
    am.check(am.lib().am_init_temp(old.data_ptr(), new.data_ptr(), C.byref(m), C.byref(g), 1.5, am.stream_handle()),
             "am_init_temp")
    torch.cuda.synchronize()
    wo, wn = orc.init_temp(e["positions"], fields.used_slots(e), n, S, P, g.outer_size, g.outer_dx, 1.5)
    written = np.zeros((n, T, T, T), bool)
    for s in fields.used_slots(e):
        if fields.depth_of(e, s) == 0:
            written[s] = True
        else:
            written[s, P:P + S, P:P + S, P:P + S] = True
    written = written.reshape(-1)
    for got, want in ((old, wo), (new, wn)):
        g_ = got.cpu().numpy()
        assert np.array_equal(g_[written].view(np.uint32), want[written].view(np.uint32))
        assert np.all(g_[~written] == 7.0)


def test_wave_sim_init_temp(am, orc, torch_cuda):
    S, P, n = 16, 3, 9
    g = am.geometry(S, P, 6, 100.0, 0.4)
    sim = am.WaveSim(g, n)
    sim.init_temp(1.0)
    root = orc.OracleTree(1).export()
    wo, _ = orc.init_temp(root["positions"], [0], 1, S, P, g.outer_size, g.outer_dx, 1.0)
    T = S + 2 * P
    for which in (0, 1):
        got = sim.get_field(which)
        assert np.array_equal(got[:T ** 3].view(np.uint32), wo.view(np.uint32))
        assert not got[T ** 3:].any()
