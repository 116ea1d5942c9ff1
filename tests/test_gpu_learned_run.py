"""The learned prolongation through an evolution: pairs harvested from the
plain solver at twice the resolution, a brief training on the GPU (native
MLP kernels), then two coarse steps of the puncture run (upwinded advection
+ Kreiss-Oliger 0.3) with the tricubic fill and with the learned fill, gated
per tile at a restriction residual of 1e-5 and ungated (the network on
every face it covers).  Reports the Hamiltonian-constraint ratio and the
throughput ratio of each against the plain run; the learned runs must stay
finite and keep H_l2 within 5% of the plain run's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_learned_against_plain_two_coarse_steps(am, torch_cuda):
    from adaptivemesh_amd import learned as LN
    print("learned vs plain: harvest, train, 2 coarse steps each", flush=True)
    pairs = LN.harvest(am, snapshots=2)
    train, val = pairs.split(8192, seed=1)
    del pairs
    net = LN.Prolongation(seed=1)
    hist = net.fit(train, steps=400, batch=4096, lr=1e-3, lr_final=1e-4, seed=1, report_every=100)
    e_net, e_tri = net.evaluate(val)
    print(f"  {train.n} pairs, loss {hist[0][1]:.4f} -> {hist[-1][1]:.4f}; held-out MSE network {e_net:.4f}, "
          f"tricubic {e_tri:.4f}", flush=True)
    del train, val
    plain = LN.puncture_run(am, None, coarse_steps=2)
    plain = LN.puncture_run(am, None, coarse_steps=2)  # the second run is timed warm
    runs = {}
    for name, thr in (("gated", 1e-5), ("everywhere", float("inf"))):
        LN.puncture_run(am, net, threshold=thr, coarse_steps=2)
        r = LN.puncture_run(am, net, threshold=thr, coarse_steps=2)
        runs[name] = r
        h_ratio = r["H_l2"] / plain["H_l2"]
        t_ratio = r["zone_cycles_per_s"] / plain["zone_cycles_per_s"]
        print(f"  {name} (per-tile threshold {thr}): H_l2 plain {plain['H_l2']:.6e} learned {r['H_l2']:.6e} "
              f"(ratio {h_ratio:.5f}); throughput plain {plain['zone_cycles_per_s']:.4e} learned "
              f"{r['zone_cycles_per_s']:.4e} zone-cycles/s (ratio {t_ratio:.3f}); fill {r['learned']}", flush=True)
        assert np.isfinite(r["H_l2"]) and np.isfinite(r["M_l2"])
        assert h_ratio < 1.05
    assert runs["everywhere"]["learned"]["learned_faces"] > 0
    assert runs["everywhere"]["learned"]["fallback_faces"] == 0
