"""am_bssn_constraints_noalloc (the allocation-free variant of
am_bssn_constraints) against am_bssn_constraints reduced on the host, on
two calls in a row through the C ABI, the second over fewer tiles, so the
second call runs on the scratch array that the first left behind.  The
sums agree to FP64 summation order (the device reduces the per-workgroup
partials in another order than the host), the cell count and the maximum
exactly.  A call with more tiles than the scratch array holds is refused
before any launch."""
import ctypes as C

import numpy as np
import pytest

from tests import meshes

pytestmark = pytest.mark.gpu

S, P = 16, 3
T = S + 2 * P


def cf3_words(rng, n):
    """Random compressed_float3 words with exponents 12..17 (values up to
    about 8): finite, non-trivial derivative triples."""
    e = rng.integers(12, 18, n, dtype=np.uint32)
    return ((e << 27) | rng.integers(0, 1 << 27, n, dtype=np.uint32)).view(np.int32)


def setup(am, torch, n_slots, rng):
    flat = np.zeros(24, np.float32)
    flat[[0, 1, 4, 6, 17]] = 1.0
    state = torch.from_numpy((flat[:, None] + rng.uniform(-1e-2, 1e-2, (24, n_slots * T ** 3))).astype(np.float32))
    state = state.cuda()
    d1 = torch.from_numpy(cf3_words(rng, 24 * n_slots * S ** 3).reshape(24, -1)).cuda()
    d2d = torch.from_numpy(cf3_words(rng, 11 * n_slots * S ** 3).reshape(11, -1)).cuda()
    d2o = torch.from_numpy(cf3_words(rng, 11 * n_slots * S ** 3).reshape(11, -1)).cuda()
    f = am.BssnFields()
    for v in range(24):
        f.state[v] = state[v].data_ptr()
        f.d1[v] = d1[v].data_ptr()
    for h in range(11):
        f.d2diag[h] = d2d[h].data_ptr()
        f.d2off[h] = d2o[h].data_ptr()
    return f, (state, d1, d2d, d2o)


def host_sums(am, torch, f, p, m, g):
    nb = am.lib().am_bssn_constraint_blocks(C.byref(g), m.n_tiles)
    partial = torch.zeros(4 * nb, dtype=torch.float64, device="cuda")
    am.check(am.lib().am_bssn_constraints(C.byref(f), C.byref(p), C.byref(m), C.byref(g), partial.data_ptr(),
                                          am.stream_handle()), "am_bssn_constraints")
    q = partial.cpu().numpy().reshape(-1, 4)
    return np.array([q[:, 0].sum(), q[:, 1].sum(), q[:, 2].sum(), q[:, 3].max()])


def test_constraints_noalloc_matches_workspace_form(am, torch_cuda):
    torch = torch_cuda
    ops, n = meshes.uniform(2)
    tree = am.Tree(n, True)
    meshes.apply(tree, ops)
    g = am.geometry(S, P, 4, 64.0, 0.25)
    rng = np.random.default_rng(7)
    f, keep = setup(am, torch, n, rng)
    p = am.bssn_options().params
    out4 = torch.full((4,), float("nan"), dtype=torch.float64, device="cuda")
    full = tree.mesh(-1)
    tiles_all = np.frombuffer(C.string_at(full.tiles, 4 * full.n_tiles), np.int32).copy()
    for n_tiles in (full.n_tiles, 11):
        print(f"noalloc constraints: {n_tiles} tiles", flush=True)
        tl = torch.from_numpy(tiles_all[:n_tiles].copy()).cuda() if n_tiles < full.n_tiles else None
        m = tree.mesh(-1)
        if tl is not None:
            m.tiles, m.n_tiles = tl.data_ptr(), n_tiles
        want = host_sums(am, torch, f, p, m, g)
        am.check(am.lib().am_bssn_constraints_noalloc(C.byref(f), C.byref(p), C.byref(m), C.byref(g),
                                                      out4.data_ptr(), am.stream_handle()), "noalloc")
        torch.cuda.synchronize()
        got = out4.cpu().numpy()
        assert got[2] == want[2] == n_tiles * S ** 3
        assert got[3] == want[3] and want[3] > 0
        assert np.allclose(got[:2], want[:2], rtol=1e-12, atol=0), (got, want)
    # no tiles: zeros
    m = tree.mesh(-1)
    m.n_tiles = 0
    am.check(am.lib().am_bssn_constraints_noalloc(C.byref(f), C.byref(p), C.byref(m), C.byref(g), out4.data_ptr(),
                                                  am.stream_handle()), "noalloc, no tiles")
    torch.cuda.synchronize()
    assert np.all(out4.cpu().numpy() == 0)
    # beyond the scratch array: refused before any launch
    m = tree.mesh(-1)
    m.n_tiles = 4097
    rc = am.lib().am_bssn_constraints_noalloc(C.byref(f), C.byref(p), C.byref(m), C.byref(g), out4.data_ptr(),
                                              am.stream_handle())
    assert rc == -1  # AM_ERR_INVALID
    assert b"scratch" in am.lib().am_last_error_string()
