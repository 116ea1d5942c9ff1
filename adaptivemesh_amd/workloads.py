"""Host-side builders of the standard meshes and initial data (numpy).

* single-level mesh: every octant refined to one depth (the 512-tile BSSN
  workload at depth 3);
* puncture mesh: a uniform base depth, then each tile whose centre lies
  within radius / 2^k of the origin refined, k levels above the base;
* data: flat space with seeded noise; a single puncture (Brill-Lindquist
  conformal factor psi = 1 + M / (2 r), r regularised at r_min, chi =
  psi^-4, gt = delta, pre-collapsed lapse alpha = psi^-2, everything else 0).

Tile centres follow the octree position encoding: child offset bit 0/1/2
selects the +x/+y/+z half (AMR.h:33-43).
"""
from __future__ import annotations

import numpy as np

FLAT = np.zeros(24, np.float32)
FLAT[[0, 1, 4, 6, 17]] = 1.0  # chi, gt_xx, gt_yy, gt_zz, alpha


def uniform_ops(depth: int):
    """add_node calls for a uniform octree of `depth` (1 + 8 + ... slots)."""
    ops, nxt, frontier = [], 1, [0]
    for _ in range(depth):
        nf = []
        for p in frontier:
            for c in range(8):
                ops.append((p, c))
                nf.append(nxt)
                nxt += 1
        frontier = nf
    return ops, nxt


def puncture_ops(outer: float = 64.0, base_depth: int = 2, levels: int = 3, radius: float = 16.0):
    """add_node calls of the puncture mesh and the slot count.  A tile at
    base_depth + k (k < levels) is refined when |centre| < radius / 2^k."""
    ops, nxt = [], 1
    centre = {0: np.zeros(3)}
    depth = {0: 0}
    frontier = [0]
    d = 0
    while frontier:
        nf = []
        for p in frontier:
            k = d - base_depth
            refine = d < base_depth or (k < levels and np.linalg.norm(centre[p]) < radius / 2 ** k)
            if not refine:
                continue
            w = outer / 2 ** d
            for c in range(8):
                off = np.array([1 if c & 1 else -1, 1 if c & 2 else -1, 1 if c & 4 else -1]) * (w / 4)
                ops.append((p, c))
                centre[nxt] = centre[p] + off
                depth[nxt] = d + 1
                nf.append(nxt)
                nxt += 1
        frontier = nf
        d += 1
    return ops, nxt


def cell_coords(position: int, S: int, P: int, outer: float):
    """(x, y, z) of every padded cell of the tile at `position`, each shaped
    (T, T, T) and indexed [z, y, x]."""
    T = S + 2 * P
    d = position >> 27
    c = np.zeros(3)
    for i in range(d):
        o = (position >> (3 * i)) & 7
        c += np.array([1 if o & 1 else -1, 1 if o & 2 else -1, 1 if o & 4 else -1]) * 0.25 / 2 ** i
    c *= outer
    h = outer / S / 2 ** d
    r = (np.arange(T) + 0.5 - T / 2) * h
    z, y, x = np.meshgrid(c[2] + r, c[1] + r, c[0] + r, indexing="ij")
    return x, y, z


def puncture_fields(positions, S: int, P: int, outer: float, mass: float = 1.0, r_min: float = 1e-3):
    """24 float32 arrays (slots * T^3) of single-puncture data on every used
    slot (unused slots hold flat space)."""
    n = len(positions)
    T = S + 2 * P
    out = [np.full((n, T, T, T), FLAT[v], np.float32) for v in range(24)]
    for s, pos in enumerate(positions):
        pos = int(pos)
        if pos == 0xFFFFFFFF:
            continue
        x, y, z = cell_coords(pos, S, P, outer)
        r = np.maximum(np.sqrt(x * x + y * y + z * z), r_min)
        psi = 1.0 + mass / (2.0 * r)
        out[0][s] = psi ** -4
        out[17][s] = psi ** -2
    return [a.reshape(-1) for a in out]


def flat_noise_fields(n_slots: int, S: int, P: int, amplitude: float = 1e-6, seed: int = 20240):
    """Flat space plus uniform noise in [-amplitude, amplitude) on every field
    and cell: numpy PCG64 seeded with seed + field index."""
    T = S + 2 * P
    out = []
    for v in range(24):
        rng = np.random.default_rng(seed + v)
        out.append((FLAT[v] + rng.uniform(-amplitude, amplitude, n_slots * T ** 3)).astype(np.float32))
    return out


def gauge_wave_fields(positions, S: int, P: int, outer: float, amplitude: float = 0.1, t: float = 0.0,
                      axis: int = 0, dtype=np.float32):
    """The harmonic gauge wave (Alcubierre et al. 2004 testbed) as BSSN data:
    ds^2 = -H dt^2 + H dw^2 + (the other two axes)^2 with
    H = 1 - A sin(2 pi (w - t) / outer), w the coordinate along `axis`, so one
    wavelength spans the domain.  alpha = sqrt(H), beta = B = 0,
    chi = H^(-1/3), gt_ww = H^(2/3), the other diagonal gt = H^(-1/3),
    K = K_ww / H with K_ww = -d_t H / (2 sqrt H), At_ww = (2/3) chi K_ww, the
    other diagonal At = -chi K / 3, Gt^w = (2/3) H^(-5/3) d_w H.  Computed in
    float64, returned as `dtype`."""
    n = len(positions)
    T = S + 2 * P
    out = [np.full((n, T, T, T), FLAT[v], np.float64) for v in range(24)]
    k = 2 * np.pi / outer
    diag = (1, 4, 6)  # gt_xx, gt_yy, gt_zz
    adiag = (8, 11, 13)
    for s, pos in enumerate(positions):
        pos = int(pos)
        if pos == 0xFFFFFFFF:
            continue
        w = cell_coords(pos, S, P, outer)[axis]
        ph = k * (w - t)
        H = 1 - amplitude * np.sin(ph)
        dtH = amplitude * k * np.cos(ph)
        dwH = -amplitude * k * np.cos(ph)
        chi = H ** (-1 / 3)
        Kww = -dtH / (2 * np.sqrt(H))
        K = Kww / H
        out[0][s] = chi
        for a in range(3):
            out[diag[a]][s] = H ** (2 / 3) if a == axis else chi
            out[adiag[a]][s] = (2 / 3) * chi * Kww if a == axis else -chi * K / 3
        out[7][s] = K
        out[14 + axis][s] = (2 / 3) * H ** (-5 / 3) * dwH
        out[17][s] = np.sqrt(H)
    return [a.reshape(-1).astype(dtype) for a in out]
