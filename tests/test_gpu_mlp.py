"""The learned prolongation's kernels on the GPU against the FP64 oracle
(oracle/mlp.py, itself pinned by central differences and Adam's closed
form).

* MFMA GEMM in all operand layouts the MLP uses (X W, dY W^T, X^T dZ) and
  both epilogues, to fp32 round-off (1e-5 of the largest |C|; K up to 1536).
* Forward activations and every parameter gradient of a 64-sample batch to
  2e-4 of each block's largest magnitude (fp32 sums of up to 1536 terms and
  erf-based GELU against FP64).
* Adam: first step equals p - lr g / (|g| + eps); twenty steps track the
  oracle to 1e-6 relative.
* Training pairs equal the oracle's: centre values, spreads and normalised
  stencils to the fp32 rounding of the restriction (4e-7 (1 + |mu|),
  divided by the spread for the stencils), denormalised targets to 1e-5 of
  the tile's largest value.
* Training on pairs of smooth fields lowers the loss, and on held-out pairs
  the network's error is below the tricubic baseline's.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp as M
from tests import fields as F
from tests import meshes

pytestmark = pytest.mark.gpu

S, P = 16, 3
T = S + 2 * P


def dev(torch, a, dtype=None):
    return torch.tensor(np.ascontiguousarray(a, dtype or np.float32), device="cuda")


@pytest.mark.parametrize("case", ["nn", "nt", "tn", "gelu", "dgelu"])
def test_gemm_layouts(am, torch_cuda, case):
    torch = torch_cuda
    rng = np.random.default_rng(1)
    M_, N_, K_ = {"nn": (128, 192, 1536), "nt": (256, 128, 192), "tn": (1536, 128, 256), "gelu": (64, 128, 1536),
                  "dgelu": (128, 128, 128)}[case]
    A = rng.standard_normal((M_, K_)) if case != "tn" else rng.standard_normal((K_, M_))
    B = rng.standard_normal((K_, N_)) if case != "nt" else rng.standard_normal((N_, K_))
    bias = rng.standard_normal(N_)
    aux = rng.standard_normal((M_, N_))
    dA, dB, dC, d2 = dev(torch, A), dev(torch, B), torch.zeros((M_, N_), device="cuda"), torch.zeros((M_, N_),
                                                                                                   device="cuda")
    dbias, daux = dev(torch, bias), dev(torch, aux)
    A64 = A if case != "tn" else A.T
    B64 = B if case != "nt" else B.T
    sam, sak = (K_, 1) if case != "tn" else (1, M_)
    sbk, sbn = (N_, 1) if case != "nt" else (1, K_)
    epi = {"gelu": 1, "dgelu": 2}.get(case, 0)
    am.check(am.lib().am_gemm_f32(dA.data_ptr(), sam, sak, dB.data_ptr(), sbk, sbn, dC.data_ptr(), N_, M_, N_, K_,
                                  dbias.data_ptr() if case in ("nn", "gelu") else None, epi,
                                  daux.data_ptr() if case == "dgelu" else None,
                                  d2.data_ptr() if case == "gelu" else None, am.stream_handle()), "gemm")
    torch.cuda.synchronize()
    ref = A64 @ B64
    if case in ("nn", "gelu"):
        ref = ref + bias
    if case == "dgelu":
        ref = ref * M.gelu_prime(aux)
    got = dC.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(got - ref)) < 1e-5 * np.max(np.abs(ref)) * np.sqrt(K_ / 16), case
    if case == "gelu":
        assert np.max(np.abs(d2.cpu().numpy() - M.gelu(ref))) < 1e-5 * np.max(np.abs(ref)) * np.sqrt(K_ / 16)


def test_forward_and_gradients_match_oracle(am, torch_cuda):
    from adaptivemesh_amd import learned as L
    torch = torch_cuda
    rng = np.random.default_rng(3)
    p = L.init_params(5)
    p[M.N_IN * M.N_HID:M.N_IN * M.N_HID + M.N_HID] = rng.normal(0, 0.1, M.N_HID)
    x = rng.uniform(-1, 1, (64, M.N_IN)).astype(np.float32)
    t = rng.normal(0, 0.05, (64, M.N_OUT)).astype(np.float32)
    net = L.Prolongation(params=p)
    part = net.loss_and_grads(dev(torch, x), dev(torch, t))
    torch.cuda.synchronize()
    f = M.forward(p.astype(np.float64), x)
    w = net.work(64)
    for k in ("z1", "h1", "z2", "h2", "y"):
        got = w[k].cpu().numpy()
        assert np.max(np.abs(got - f[k])) < 2e-4 * np.max(np.abs(f[k])), k
    loss, g = M.grads(p.astype(np.float64), x, t)
    assert abs(part.sum().item() / (64 * M.N_OUT) - loss) < 1e-4 * loss
    gg = net.grads.cpu().numpy()
    o = 0
    for s in M.SHAPES:
        n = int(np.prod(s))
        assert np.max(np.abs(gg[o:o + n] - g[o:o + n])) < 2e-4 * np.max(np.abs(g[o:o + n])), s
        o += n


def test_adam_matches_closed_form_and_oracle(am, torch_cuda):
    from adaptivemesh_amd import learned as L
    torch = torch_cuda
    rng = np.random.default_rng(4)
    n = 4096
    p0 = rng.standard_normal(n)
    pd, md, vd, gd = dev(torch, p0), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"), torch.zeros(
        n, device="cuda")
    pr, mr, vr = p0.astype(np.float32).astype(np.float64), np.zeros(n), np.zeros(n)
    for step in range(1, 21):
        g = rng.standard_normal(n).astype(np.float32)
        gd.copy_(torch.from_numpy(g).cuda())
        am.check(am.lib().am_mlp_adam(pd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(), n, step, 1e-2,
                                      0.9, 0.999, 1e-8, am.stream_handle()), "adam")
        torch.cuda.synchronize()
        if step == 1:
            want = pr - 1e-2 * g / (np.abs(g) + 1e-8)
            assert np.max(np.abs(pd.cpu().numpy() - want)) < 1e-6
        pr, mr, vr = M.adam(pr, g.astype(np.float64), mr, vr, step, lr=1e-2)
        assert np.max(np.abs(pd.cpu().numpy() - pr)) < 1e-6 * max(1.0, np.max(np.abs(pr))), step


def smooth_fields(e, n, outer, seed):
    rng = np.random.default_rng(seed)
    out = []
    for v in range(24):
        k = rng.uniform(0.5, 2.0, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        a = rng.uniform(0.1, 1.0) * 10.0 ** rng.uniform(-3, 0)
        c = rng.uniform(-1, 1)
        out.append(F.sample(e["positions"], n, S, P, outer,
                            lambda x, y, z, k=k, ph=ph, a=a, c=c: c + a * np.sin(k[0] * x + ph[0])
                            * np.cos(k[1] * y + ph[1]) * np.cos(k[2] * z + ph[2])))
    return out


def test_pairs_match_oracle(am, orc, torch_cuda):
    from adaptivemesh_amd import learned as L
    torch = torch_cuda
    ops, n = meshes.uniform(1)
    t = orc.OracleTree(n, True)
    meshes.apply(t, ops)
    e = t.export()
    f = smooth_fields(e, n, 8.0, 0)
    fd = [dev(torch, a) for a in f]
    g = am.geometry(S, P, 4, 8.0, 0.25)
    tiles = [2, 7]
    pairs = L.PairSet.from_fields([a.data_ptr() for a in fd], tiles, g)
    torch.cuda.synchronize()
    per = int(am.lib().am_mlp_pairs_per_tile(C.byref(g)))
    assert per == 343 and pairs.n == 686
    for i, s in enumerate(tiles):
        X, Tn, mu, sg = M.tile_pairs([a.astype(np.float64) for a in f], s, S, P)
        gx = pairs.x[i * per:(i + 1) * per].cpu().numpy()
        gt = pairs.t[i * per:(i + 1) * per].cpu().numpy()
        gmu = pairs.mu[i * per:(i + 1) * per].cpu().numpy()
        gsg = pairs.sigma[i * per:(i + 1) * per].cpu().numpy()
        # fp32 restriction rounds each coarse value by ~1e-7 |mu|: sigma and the
        # normalised stencil inherit that absolute error
        tol = 4e-7 * (1 + np.abs(mu))
        assert np.all(np.abs(gmu - mu) <= tol)
        assert np.all(np.abs(gsg - sg) <= 2 * tol)
        assert np.all(np.abs(gx - X).reshape(-1, 24, 64) <= (4 * tol / sg)[:, :, None] + 2e-6)
        scale = max(1.0, np.max(np.abs(mu)))
        assert np.max(np.abs(gt * gsg.repeat(8, 1) - Tn * sg.repeat(8, 1))) < 1e-5 * scale


def test_training_beats_tricubic_on_held_out_pairs(am, orc, torch_cuda):
    print("MLP: train on pairs of smooth fields, evaluate on held-out pairs", flush=True)
    from adaptivemesh_amd import learned as L
    torch = torch_cuda
    ops, n = meshes.uniform(2)
    t = orc.OracleTree(n, True)
    meshes.apply(t, ops)
    e = t.export()
    g = am.geometry(S, P, 4, 16.0, 0.25)
    leaves = list(range(9, n))
    sets = []
    for seed in range(3):
        f = smooth_fields(e, n, 16.0, 10 + seed)
        fd = [dev(torch, a) for a in f]
        sets.append(L.PairSet.from_fields([a.data_ptr() for a in fd], leaves, g))
        torch.cuda.synchronize()
    train, test = L.PairSet.cat(sets).split(8192)
    net = L.Prolongation(seed=0)
    hist = net.fit(train, steps=600, batch=2048, lr=1e-3, report_every=100, lr_final=1e-4)
    torch.cuda.synchronize()
    e_net, e_base = net.evaluate(test, batch=2048)
    print(f"  loss {hist[0][1]:.3e} -> {hist[-1][1]:.3e}; held-out normalised MSE: network {e_net:.3e}, "
          f"tricubic {e_base:.3e} (gain {1 - e_net / e_base:.1%})", flush=True)
    assert hist[-1][1] < 0.5 * hist[0][1]
    assert e_net < e_base
