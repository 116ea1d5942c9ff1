#!/usr/bin/env python3
"""BSSN zone-cycles/s on one MI355X, with the roofline of the dominant kernel
and a CPU baseline.

Default workload (one GPU): the BSSN single-level mesh -- 512 tiles of 16^3
interior cells (128^3) at depth 3 with 3-cell padding, their 73 ancestors,
periodic outer faces, domain [-32, 32]^3, 24 fp32 fields, 1+log slicing and
Gamma-driver shift, flat metric plus seeded uniform noise of +-1e-6 on every
field, FD order 4, derivatives as compressed_float3 words, CFL 0.25.  One
step = one timestep() of the octree driver = 2^3 fine RK4 sub-steps (depth
d advances every 2^(3-d) sub-steps), so the default 12 steps are ~100 fine
RK4 steps.  A zone-cycle is one interior cell advanced by one full RK4 step
at its own depth, ancestors included.  Before timing, the autotuner picks
the launch shapes and the coarse step is recorded as a hipGraph (replays
are bit-identical to eager launches); --no-autotune / --no-graph turn them
off.  --workload puncture times the octree puncture run instead (its own
JSON line, not the headline).

Output: one JSON line (rank 0).  --gpus > 1 is refused: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BSSN zone-cycles/s + %HBM peak; learned vs plain ghost fill at equal ‖H‖₂"
HBM_PEAK_GBS = 8000.0
S, P, FD_ORDER, OUTER, CFL, DEPTH = 16, 3, 4, 64.0, 0.25, 3
NOISE = 1e-6
PUNCTURE_KO_SIGMA = 0.3

# algorithmic bytes per interior cell (each distinct element read once and
# written once per launch)
WORD = 4
Q = FD_ORDER // 2  # stencil half-width: the ghost layers the FD reads and the push writes
# HBM bytes per depth-3 k_bssn_stage launch (2,097,152 cells), averaged over
# the four stages, from rocprofv3 --pmc FETCH_SIZE (x 2.000, the copy and
# read probes' calibration) and WRITE_SIZE (x 1.000)
# (profiles/bssn_rk4_512tiles_hbm_counters.md), and the algorithmic bytes of
# the same launches
RHS_TRAFFIC_PER_LAUNCH = 1317664164
RHS_ALGORITHMIC_PER_LAUNCH = 1317011456
TRAFFIC_DIRECT_FACES = 6  # the mask the counters were read with


def face_stats(export, S, depth_filter=None):
    """Mean, over the advanced tiles weighted by how often each advances
    (2^depth), of the faces per tile whose neighbour has the same depth
    (pushed) and of those whose neighbour is coarser (interpolated)."""
    pos, bnd = export["positions"], export["boundaries"]
    eq = co = wsum = 0.0
    for s, p_ in enumerate(pos):
        p_ = int(p_)
        if p_ == 0xFFFFFFFF:
            continue
        d = p_ >> 27
        if depth_filter is not None and d != depth_filter:
            continue
        w = 2.0 ** d
        wsum += w
        for f in range(6):
            nd = int(bnd[s * 6 + f]) & 0xFFFFFFFF
            tgt = nd & 0xFFFFFF
            if tgt & 0x800000 or nd >> 28 in (0, 15):
                continue
            if (nd >> 24) & 15:
                co += w
            else:
                eq += w
    return eq / max(wsum, 1), co / max(wsum, 1)


def bytes_model(direct_faces, eq_faces, coarse_faces, upwind=False, q=Q):
    """Algorithmic bytes per interior cell and launch class.  eq_faces /
    coarse_faces: mean faces per advanced tile with an equal-depth / a
    coarser neighbour (face_stats).  direct_faces is the driver's axis mask
    (0, 2, 4 or 6) of the equal-depth faces that stages 0-2 do not push (the
    next stage's FD reads those ghost layers from the neighbours: the same
    distinct elements as the pushed copies).  q: the stencils' half-width,
    the ghost layers the FD reads and the push writes (fd_order / 2).
    upwind: the FD also writes the upwinded advection + Kreiss-Oliger term
    per field (read back by the stage kernel), and both the FD's stencils
    and the push reach P layers."""
    q = P if upwind else q
    adv = 24 * 4 if upwind else 0
    push = eq_faces * q * S * S * 24 * 4 / S ** 3  # equal-depth ghost layers written by the stage kernel
    push_early = push * (3 - bin(direct_faces & 7).count("1")) / 3
    return {
        # per stage, 24 fields: the interior and q ghost layers per face, (S^3 + 6 q S^2) * 4 per field; 24
        # gradient + 11 diag + 11 off-diag words (and the advection term)
        "fd": 24 * (S ** 3 + 6 * q * S * S) * 4 / S ** 3 + 46 * WORD + adv,
        # per RK4 step, the four stages summed: state 96 and 46 words 184 (and the advection term) read in every
        # stage; y_n 96 and acc 96 read in stages 1-3; new 96 written in every stage, acc 96 written in stages
        # 0-2; the pushed face ghosts (stage 3 pushes every equal-depth face)
        "rhs_stage": (96 + 184 + 192 + push_early) + (96 + 184 + 96 + 96 + 192 + push_early) * 2
                     + (96 + 184 + 96 + 96 + 96 + push) + 4 * adv,
        # per stage: P S^2 ghost cells per coarse face, one source read and one write, 24 fields
        "ghost": coarse_faces * 24 * P * S * S * 8 / S ** 3,
        # per step: interior read of new + interior write of old per field (tiles that a finer face reads
        # copy their ghost cells too; not counted)
        "copy_back": 24 * 2 * 4,
    }


WORKLOADS = {
    "bssn512": "BSSN single level: 512 tiles of 16^3 (128^3) + 73 ancestors, P=3, 24 fp32 fields, 1+log + "
               "Gamma-driver, flat + 1e-6 noise, cf3 derivative words (Christoffel symbols raised in registers, not "
               "stored), RK4, "
               "FD order 4, periodic, [-32,32]^3, CFL 0.25",
    "puncture": "BSSN octree: single puncture (M=1) at (2,2,2) on [-32,32]^3, depth-2 base of 64 tiles; a tile "
                "at depth 2/3/4 is refined while the box distance from the puncture to its cube is below "
                "16/10.5/8 (1/8/64/160/312/648 tiles to depth 5), P=3, 24 fp32 fields, upwinded advection + "
                "Kreiss-Oliger sigma=%g, cf3 derivative words (Christoffel symbols raised in registers), RK4 "
                "sub-cycled 2x per level, tricubic + quadratic-time coarse-face ghosts, FD order 4, CFL 0.25",
}
STAGES_PER_STEP = {"fd": 4, "rhs_stage": 1, "ghost": 4, "copy_back": 1}
# kernel classes launched by this workload (every face of the uniform periodic mesh has an equal-depth target,
# so no coarse-face ghost fill runs)
WORKLOAD_CLASSES = ("fd", "rhs_stage", "copy_back")


def build_mesh(sim):
    frontier, nxt = [0], 1
    for _ in range(DEPTH):
        nf = []
        for p in frontier:
            for c in range(8):
                assert sim.add_node(p, c) == nxt
                nf.append(nxt)
                nxt += 1
        frontier = nf
    return nxt


def initial_data(n_slots, T):
    """Flat space plus uniform noise in [-1e-6, 1e-6) on every field and
    cell: numpy PCG64 seeded with 20240 + field index."""
    flat = np.zeros(24, np.float32)
    flat[[0, 1, 4, 6, 17]] = 1.0  # chi, gt_xx, gt_yy, gt_zz, alpha
    out = []
    for v in range(24):
        rng = np.random.default_rng(20240 + v)
        out.append((flat[v] + rng.uniform(-NOISE, NOISE, n_slots * T ** 3)).astype(np.float32))
    return out


def cpu_baseline(n_slots, fields, export, reps=5, target_s=2.0):
    """The FP32 host implementation of the scheme (oracle/am_cpu_fp32.c,
    OpenMP on this host's cores): one RK4 step -- per stage, derivatives of
    order 4 straight from the fp32 state, the BSSN right-hand side and the
    RK4 update in one pass per cell, then the equal-depth face ghosts of the
    stage output refreshed from the neighbours -- over a sample of the 512
    depth-3 tiles sized so a step takes about target_s; one warm-up step,
    then the median of `reps` timed steps.  Only that scheme work is in the
    timed region (no allocation, no Python array work)."""
    from oracle import cpu_fp32 as CPU
    from oracle import oracle as orc
    pos, bnd = export["positions"], export["boundaries"]
    tiles_all = np.array([t for t in range(n_slots) if (int(pos[t]) >> 27) == DEPTH and int(pos[t]) != 0xFFFFFFFF],
                         np.int32)
    outer_dx = np.float32(OUTER) / np.float32(S)
    inv_dx = float(np.float32(2 ** DEPTH) / outer_dx)
    dt = float(np.float32(outer_dx * np.float32(CFL)) / 2 ** DEPTH)
    params = np.array(orc.BSSN_DEFAULT, np.float64)
    yn = [np.ascontiguousarray(f, np.float32) for f in fields]
    a = [x.copy() for x in yn]
    b = [x.copy() for x in yn]
    acc = [np.zeros_like(x) for x in yn]
    threads = int(CPU.lib().cpu32_max_threads())

    def step(tiles):
        t0 = time.perf_counter()
        CPU.rk4_step(yn, a, b, acc, bnd, tiles, S, P, Q, inv_dx, dt, params)
        return time.perf_counter() - t0

    t32 = step(tiles_all[:32])
    n = int(min(len(tiles_all), max(8, 32 * target_s / max(t32, 1e-4))))
    sample = np.ascontiguousarray(tiles_all[:n])
    step(sample)  # warm-up
    times = sorted(step(sample) for _ in range(reps))
    med = times[len(times) // 2]
    cells = n * S ** 3
    return {"value": cells / med, "unit": "zone-cycles/s", "cores": threads, "kind": "port", "dtype": "fp32",
            "sample": f"FP32 host implementation of the scheme (oracle/am_cpu_fp32.c, gcc -O3 x86-64-v3, OpenMP, "
                      f"{threads} threads): one RK4 step of {n} of the 512 depth-3 tiles ({cells} zone-cycles); "
                      f"timed region: 4 x (order-4 derivatives + BSSN RHS + RK4 update per cell, equal-depth ghost "
                      f"refresh of the sample); median of {reps} steps after one warm-up: {med:.3f} s "
                      f"(min {times[0]:.3f}, max {times[-1]:.3f})"}


def cpu_kernel_table(n_slots, fields, export, sub=64):
    """The reference's kernel bodies as the oracle restates them (gcc -O2,
    no contraction, bit-exact with the reference's float arithmetic) and the
    FP32 implementation's passes, timed on this host with 1 thread (the
    first `sub` depth-3 tiles) and with every thread (all 512), per RK4
    stage of the 24 fields: markdown rows of ms per call and ns per
    interior cell."""
    from oracle import cpu_fp32 as CPU
    from oracle import oracle as orc
    L, L32 = orc.lib(), CPU.lib()
    pos, bnd, ch = export["positions"], export["boundaries"], export["children"]
    tiles_all = np.array([t for t in range(n_slots) if (int(pos[t]) >> 27) == DEPTH and int(pos[t]) != 0xFFFFFFFF],
                         np.int32)
    used = np.array([t for t in range(n_slots) if int(pos[t]) != 0xFFFFFFFF], np.int32)
    outer_dx = float(np.float32(OUTER) / np.float32(S))
    inv_dx = float(np.float32(2 ** DEPTH) / np.float32(outer_dx))
    dt = float(np.float32(np.float32(outer_dx) * np.float32(CFL)) / 2 ** DEPTH)
    params = np.array(orc.BSSN_DEFAULT, np.float64)
    st = [np.ascontiguousarray(f, np.float32) for f in fields]
    W = n_slots * S ** 3
    w1 = [np.zeros(W, np.uint32) for _ in range(24)]
    w2 = [np.zeros(W, np.uint32) for _ in range(11)]
    w3 = [np.zeros(W, np.uint32) for _ in range(11)]
    old = [x.copy() for x in st]
    new = [x.copy() for x in st]
    out = [x.copy() for x in st]
    acc = [np.zeros_like(x) for x in st]
    d1 = [np.zeros(3 * len(tiles_all) * S ** 3, np.float32) for _ in range(24)]
    d2 = [np.zeros(6 * len(tiles_all) * S ** 3, np.float32) for _ in range(11)]
    hess = list(orc.HESS_VARS)

    def differentiate(t):
        h = 0
        for v in range(24):
            dd = w2[h].ctypes.data if v in hess else None
            L.orc_differentiate(st[v].ctypes.data, w1[v].ctypes.data, dd, pos.ctypes.data, t.ctypes.data, len(t), 0,
                                DEPTH, 11 + v, S, P, Q, outer_dx)
            h += v in hess

    def mixed(t):
        for h, v in enumerate(hess):
            L.orc_mixed_partials(st[v].ctypes.data, w3[h].ctypes.data, pos.ctypes.data, t.ctypes.data, len(t), 0,
                                 DEPTH, 7 + v, S, P, outer_dx)

    def copy_boundaries(t):
        for v in range(24):
            L.orc_copy_boundaries(old[v].ctypes.data, new[v].ctypes.data, None, bnd.ctypes.data, t.ctypes.data,
                                  len(t), 0, DEPTH, 1, 0, S, P, 1.0, 0)

    def copy_new_to_old(t):
        for v in range(24):
            L.orc_copy_new_to_old(old[v].ctypes.data, new[v].ctypes.data, pos.ctypes.data, ch.ctypes.data,
                                  t.ctypes.data, len(t), 0, DEPTH, S, P)

    def fp32_derivatives(t):
        L32.cpu32_derivatives(CPU.ptrs(st), CPU.ptrs(d1), CPU.ptrs(d2), t.ctypes.data, len(t), S, P, Q, inv_dx)

    def fp32_stage(t):
        L32.cpu32_bssn_stage(CPU.ptrs(st), CPU.ptrs(st), CPU.ptrs(acc), CPU.ptrs(out), t.ctypes.data, len(t), S, P, Q,
                             inv_dx, dt, 1, params.ctypes.data, 0.0, 0)

    rows = [("`__differentiate` (AMR.h:1152-1219): 24 gradients + 11 Hessian diagonals, dithered words",
             differentiate, "tiles"),
            ("`__mixed_partials` (AMR.h:1232-1266): 11 fields, dithered words", mixed, "tiles"),
            ("`__copy_boundaries` (AMR.h:702-762): 24 fields, quadratic formula, depth-3 faces", copy_boundaries,
             "tiles"),
            ("`__copy_new_to_old` (AMR.h:961-996): 24 fields, every slot", copy_new_to_old, "used"),
            ("FP32 implementation: derivatives of the 24 fields (gradients + 11 Hessians, fp32 out)",
             fp32_derivatives, "tiles"),
            ("FP32 implementation: one RK4 stage (derivatives + RHS + update)", fp32_stage, "tiles")]
    n_all = int(L32.cpu32_max_threads())
    lines = [f"| kernel body | 1 thread: ms ({sub} tiles) | 1 thread: ns/cell | {n_all} threads: ms (512 tiles) | "
             f"{n_all} threads: ns/cell |", "|---|---|---|---|---|"]
    for name, fn, dom in rows:
        res = []
        for threads, t in ((1, tiles_all[:sub]), (n_all, tiles_all)):
            if dom == "used":
                t = used if threads == n_all else used[np.isin(used, tiles_all[:sub]) | (used < 73)]
            t = np.ascontiguousarray(t, np.int32)
            L.orc_set_threads(threads)
            L32.cpu32_set_threads(threads)
            fn(t)  # warm-up
            times = []
            for _ in range(3):
                t0 = time.perf_counter()
                fn(t)
                times.append(time.perf_counter() - t0)
            ms = 1e3 * sorted(times)[1]
            cells = len([x for x in t if (int(pos[x]) >> 27) == DEPTH]) * S ** 3
            res += [f"{ms:.1f}", f"{1e6 * ms / cells:.2f}"]
        lines.append(f"| {name} | " + " | ".join(res) + " |")
    L.orc_set_threads(n_all)
    L32.cpu32_set_threads(n_all)
    return "\n".join(lines)


def copy_probe(mib=1024, reps=10):
    """A plain device-to-device copy of `mib` MiB, and a sum over the same
    buffer: the reference points for the HBM byte counters (the copy reads
    and writes exactly that many bytes, the sum's main kernel only reads
    them) and the achievable copy and read bandwidths."""
    import torch
    n = mib * 2 ** 20 // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    out = {}
    for name, fn, moved in (("copy_probe", lambda: b.copy_(a), 8 * n), ("read_probe", lambda: a.sum(), 4 * n)):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n if moved == 8 * n else 0,
                     "avg_us": round(1e3 * ms, 2), "GBps": round(moved / (ms * 1e-3) / 1e9, 1)}
    print(json.dumps(out), flush=True)


def timed_run(args, fd_order, steps, autotune=True):
    """Builds the workload's driver at stencil order fd_order, autotunes it
    (unless fixed by the flags), records the coarse step as a graph, runs
    args.warmup untimed and `steps` timed coarse steps, then times every
    kernel class with HIP events over one more eager step.  Returns the
    driver, the zone-cycles and seconds of the timed steps, the kernel table,
    the roofline of the dominant kernel, the whole-step HBM figures and the
    autotuner's report."""
    import adaptivemesh_amd as am
    import torch
    from adaptivemesh_amd import workloads as WL

    q = fd_order // 2
    T = S + 2 * P
    if args.workload == "bssn512":
        g = am.geometry(S, P, fd_order, OUTER, CFL)
        n_slots = 1 + 8 + 64 + 512
        sim = am.BssnSim(g, n_slots, am.bssn_options(periodic=True))
        assert sim.push_layers == q, "the byte model assumes the driver's default push"
        assert build_mesh(sim) == n_slots
        fields = initial_data(n_slots, T)
    else:
        ops, n_slots = WL.puncture_ops(OUTER)
        g = am.geometry(S, P, fd_order, OUTER, CFL)
        sim = am.BssnSim(g, n_slots, am.bssn_options(periodic=False, interp=am.AM_INTERP_TRICUBIC))
        for p_, c in ops:
            sim.add_node(p_, c)
        sim.set_dissipation(args.ko_sigma, True)
        fields = WL.puncture_fields(sim.export()["positions"], S, P, OUTER)
    for v in range(24):
        sim.set_field(0, v, fields[v])
        sim.set_field(1, v, fields[v])
    sim.fill_ghosts()
    sim.sync()
    tune = None
    if args.z_block or args.no_autotune or not autotune:
        sim.set_launch(args.z_block, 0, 0)
    else:
        tune = sim.autotune(reps=3)
    if args.direct_faces is not None:  # after the autotuner, which also picks a mask
        sim.set_direct_faces(args.direct_faces)
    if not args.no_graph:
        sim.use_graph(True)

    for _ in range(args.warmup):
        sim.timestep()
    sim.sync()
    torch.cuda.synchronize()
    z0 = sim.zone_cycles
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.timestep()
    sim.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    zones = sim.zone_cycles - z0
    elapsed = t1 - t0
    classes = WORKLOAD_CLASSES if args.workload == "bssn512" else WORKLOAD_CLASSES + ("ghost",)
    eq_faces, coarse_faces = face_stats(sim.export(), S)
    BYTES_ = bytes_model(sim.direct_faces, eq_faces, coarse_faces, upwind=sim.dissipation[1], q=q)
    bytes_per_zone = sum(BYTES_[k] * STAGES_PER_STEP[k] for k in classes)
    whole_gbs = zones * bytes_per_zone / elapsed / 1e9

    roofline = None
    kernels = {}
    if not args.no_profile:
        # per-launch HIP events on the driver's stream over one more step,
        # launched eagerly (a replayed graph has no per-kernel events)
        sim.use_graph(False)
        sim.profile(True)
        zp = sim.zone_cycles
        sim.timestep()
        sim.sync()
        zp = sim.zone_cycles - zp
        times = sim.kernel_times()
        sim.profile(False)
        for k, (ms, cnt) in times.items():
            if cnt == 0:
                continue
            kernels[k] = {"ms_total": round(ms, 4), "launches": cnt, "avg_us": round(1e3 * ms / cnt, 2)}
            if k in classes:
                total_bytes = zp * BYTES_[k] * STAGES_PER_STEP[k]
                kernels[k].update(GBps=round(total_bytes / (ms * 1e-3) / 1e9, 1),
                                  bytes_per_zone_cycle=round(BYTES_[k] * STAGES_PER_STEP[k], 1))
        dom = max((k for k in kernels if "GBps" in kernels[k]), key=lambda k: kernels[k]["ms_total"])
        d = kernels[dom]
        measured = (dom == "rhs_stage" and args.workload == "bssn512" and fd_order == FD_ORDER
                    and sim.direct_faces == TRAFFIC_DIRECT_FACES)
        roofline = {"bound": "hbm", "kernel": dom, "achieved": d["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(d["GBps"] / HBM_PEAK_GBS, 4),
                    "traffic": RHS_TRAFFIC_PER_LAUNCH if measured else None}
        if args.workload != "bssn512":
            roofline["note"] = (f"byte model from this mesh's faces: {eq_faces:.2f} equal-depth and "
                                f"{coarse_faces:.2f} coarser neighbours per advanced tile")
        if roofline["traffic"]:
            roofline["traffic_note"] = ("HBM bytes per depth-3 launch (mean of the four stages) from rocprofv3 "
                                        "FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000; "
                                        f"{RHS_TRAFFIC_PER_LAUNCH / RHS_ALGORITHMIC_PER_LAUNCH:.4f}x the algorithmic "
                                        "bytes of the same launches")
    hbm = {"achieved": round(whole_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(whole_gbs / HBM_PEAK_GBS, 4), "bytes_per_zone_cycle": round(bytes_per_zone)}
    return {"sim": sim, "fields": fields, "n_slots": n_slots, "zones": zones, "elapsed": elapsed, "tune": tune,
            "kernels": kernels, "roofline": roofline, "hbm": hbm}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed coarse steps (default 12; 64 for the puncture, 64 M of evolution)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel event timing pass")
    ap.add_argument("--z-block", type=int, default=0, help="fixed FD z-block (skips the autotuner)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="bssn512")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--ko-sigma", type=float, default=PUNCTURE_KO_SIGMA,
                    help="Kreiss-Oliger strength of the puncture run (it always upwinds the advection terms)")
    ap.add_argument("--direct-faces", type=int, default=None, choices=(0, 2, 4, 6),
                    help="axis mask (4: z) of the equal-depth faces RK4 stages 0-2 leave to the next FD "
                         "(default: the autotuner's choice)")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a recorded step")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-versus-plain ghost fill comparison (training + two puncture runs)")
    ap.add_argument("--learned-steps", type=int, default=3000, help="Adam steps of the learned prolongation")
    ap.add_argument("--no-order6", action="store_true",
                    help="skip the second measurement with the reference's order-6 stencil (3 ghost layers)")
    ap.add_argument("--cpu-kernels", action="store_true",
                    help="print only a markdown table of the reference's kernel bodies (the oracle's restatement) and "
                         "the FP32 implementation's passes timed on this host, 1 thread and all threads")
    ap.add_argument("--copy-probe", action="store_true",
                    help="time only a plain 1 GiB device copy and a 1 GiB sum (the calibrations for HBM byte "
                         "counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU: sharding this path over several GPUs is a design only; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.steps is None:
        args.steps = 64 if args.workload == "puncture" else 12
    if args.copy_probe:
        copy_probe()
        return
    if args.cpu_kernels:  # host cores only: the mesh from the oracle's octree
        from oracle import oracle as orc
        n_slots = 1 + 8 + 64 + 512
        ot = orc.OracleTree(n_slots, True)
        assert build_mesh(ot) == n_slots
        print(cpu_kernel_table(n_slots, initial_data(n_slots, S + 2 * P), ot.export()), flush=True)
        return

    import adaptivemesh_amd as am  # fails loudly without the HIP library

    R = timed_run(args, FD_ORDER, args.steps)
    sim, fields, n_slots, tune = R["sim"], R["fields"], R["n_slots"], R["tune"]
    zones, elapsed = R["zones"], R["elapsed"]
    value = zones / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    cons = sim.constraints()
    finite = all(np.isfinite(v) for v in cons.values())
    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload] % ((args.ko_sigma,) if "%g" in WORKLOADS[args.workload]
                                                           else ()),
                   "tiles": n_slots, "zone_cycles_per_step": zones // max(args.steps, 1),
                   "fine_rk4_steps_per_step": 2 ** sim.counter()[1],
                   "launch": {"z_block": sim.get_launch()[0], "fd_grid_limit": sim.get_launch()[1],
                              "rhs_grid_limit": sim.get_launch()[2], "fd_prefetch_planes": sim.get_launch()[3],
                              "push_layers": sim.push_layers,
                              "direct_faces": sim.direct_faces, "overlap_depths": sim.overlap_depths,
                              "autotuned": tune is not None,
                              "autotune_trials": tune["trials"] if tune else None,
                              "graph_per_coarse_step": not args.no_graph}},
        "hbm_whole_step": R["hbm"],
        "roofline": R["roofline"],
        "kernels": R["kernels"],
        "constraints_after": {k: float(f"{v:.6g}") for k, v in cons.items()},
        "valid": finite,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline and args.workload == "bssn512":
        result["cpu_baseline"] = cpu_baseline(n_slots, fields, sim.export())
        result["cpu_baseline"]["value"] = round(result["cpu_baseline"]["value"], 1)
    if not args.no_order6 and args.workload == "bssn512" and not args.cpu_kernels:
        # the reference's own stencil at P = 3: order 2P = 6 (AMR.h:1144-1150,
        # 1171), reading and pushing all three ghost layers
        R6 = timed_run(args, 6, max(2, args.steps // 2))
        d6 = R6["roofline"]
        result["fd_order_6"] = {
            "workload": "as config.workload with FD order 6 (the reference's order at P=3: three ghost layers read "
                        "by the stencils and pushed)",
            "value": round(R6["zones"] / R6["elapsed"], 1), "unit": "zone-cycles/s",
            "steps": max(2, args.steps // 2), "ms_per_step": round(1e3 * R6["elapsed"] / max(2, args.steps // 2), 3),
            "push_layers": R6["sim"].push_layers, "direct_faces": R6["sim"].direct_faces,
            "hbm_whole_step": R6["hbm"], "roofline": d6, "kernels": R6["kernels"]}
        del R6
    if not args.no_learned and args.workload == "bssn512":
        # the second half of the metric: learned versus plain (tricubic)
        # coarse-face ghost fill on the octree puncture run, same resolution
        from adaptivemesh_amd import learned as LN
        _, rep = LN.train_and_compare(am, train_steps=args.learned_steps)
        pl, le, la = rep["plain"], rep["learned"], rep["learned_all"]
        g6 = lambda x: float(f"{x:.6g}")
        result["learned_vs_plain"] = {
            "workload": WORKLOADS["puncture"] % PUNCTURE_KO_SIGMA + "; 2 coarse steps; learned fill gated per "
                        "tile (a tile whose largest restriction residual exceeds the threshold keeps tricubic)",
            "H_l2_plain": g6(pl["H_l2"]), "H_l2_learned": g6(le["H_l2"]),
            "H_l2_ratio": round(le["H_l2"] / pl["H_l2"], 5),
            "M_l2_plain": g6(pl["M_l2"]), "M_l2_learned": g6(le["M_l2"]),
            "zone_cycles_per_s_plain": round(pl["zone_cycles_per_s"], 1),
            "zone_cycles_per_s_learned": round(le["zone_cycles_per_s"], 1),
            "zone_cycles_per_s_ratio": round(le["zone_cycles_per_s"] / pl["zone_cycles_per_s"], 4),
            "residual_threshold": rep["threshold"],
            "learned_fill": le["learned"],
            "without_threshold": {"H_l2": g6(la["H_l2"]), "H_l2_ratio": round(la["H_l2"] / pl["H_l2"], 5),
                                  "M_l2": g6(la["M_l2"]), "zone_cycles_per_s": round(la["zone_cycles_per_s"], 1),
                                  "zone_cycles_per_s_ratio": round(la["zone_cycles_per_s"] / pl["zone_cycles_per_s"],
                                                                   4),
                                  "learned_fill": la["learned"]},
            "training": {k: rep[k] for k in ("pairs", "train_steps", "batch", "loss_start", "loss_end", "harvest_s",
                                             "train_s")},
            "prolongation_mse": {"heldout_network": rep["heldout_mse_network"],
                                 "heldout_tricubic": rep["heldout_mse_tricubic"],
                                 "unseen_network": rep["unseen_mse_network"],
                                 "unseen_tricubic": rep["unseen_mse_tricubic"]},
        }
    print(json.dumps(result), flush=True)
    if not finite:
        print("bench.py: the constraint norms after the timed steps are not finite; the run is invalid",
              file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
