"""Host-side builders of the standard meshes and initial data (numpy).

* single-level mesh: every octant refined to one depth (the 512-tile BSSN
  workload at depth 3);
* puncture mesh: a uniform base depth, then each tile whose box lies within
  radii[k] of the puncture refined, k levels above the base;
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


# The puncture mesh: the puncture off every tile corner, at (2, 2, 2) (in
# units of M = 1), and a tile at depth 2 + k refined while the distance from
# the puncture to its box is below PUNCTURE_RADII[k].  On [-32, 32]^3 this
# gives 1/8/64/160/312/648 tiles at depths 0-5 (1,193 slots): uneven counts
# per level near the 64/160/320/640 of the octree workload.  A puncture at
# the origin would give only the symmetric shells (64, 256 or 448 tiles at
# depth 3).
PUNCTURE_CENTRE = (2.0, 2.0, 2.0)
PUNCTURE_RADII = (16.0, 10.5, 8.0)


def puncture_ops(outer: float = 64.0, base_depth: int = 2, radii=PUNCTURE_RADII, centre=PUNCTURE_CENTRE):
    """add_node calls of the puncture mesh and the slot count: a uniform base
    at base_depth, then a tile at base_depth + k (k < len(radii)) is refined
    when the distance from `centre` to its box is below radii[k] (lengths in
    the units of outer, which the radii and centre scale with: they are
    given for outer = 64)."""
    scale = outer / 64.0
    c0 = np.asarray(centre, float) * scale
    ops, nxt = [], 1
    lo = {0: np.full(3, -outer / 2)}
    frontier = [0]
    d = 0
    while frontier:
        nf = []
        w = outer / 2 ** d
        for p in frontier:
            k = d - base_depth
            if d < base_depth:
                refine = True
            elif k < len(radii):
                gap = np.maximum(0.0, np.maximum(lo[p] - c0, c0 - (lo[p] + w)))
                refine = np.linalg.norm(gap) < radii[k] * scale
            else:
                refine = False
            if not refine:
                continue
            for c in range(8):
                off = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1], float) * (w / 2)
                ops.append((p, c))
                lo[nxt] = lo[p] + off
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


def puncture_fields(positions, S: int, P: int, outer: float, mass: float = 1.0, r_min: float = 1e-3,
                    centre=PUNCTURE_CENTRE):
    """24 float32 arrays (slots * T^3) of single-puncture data on every used
    slot (unused slots hold flat space), the puncture at `centre` (given for
    outer = 64 and scaled with outer, as in puncture_ops)."""
    c0 = np.asarray(centre, float) * (outer / 64.0)
    n = len(positions)
    T = S + 2 * P
    out = [np.full((n, T, T, T), FLAT[v], np.float32) for v in range(24)]
    for s, pos in enumerate(positions):
        pos = int(pos)
        if pos == 0xFFFFFFFF:
            continue
        x, y, z = cell_coords(pos, S, P, outer)
        x, y, z = x - c0[0], y - c0[1], z - c0[2]
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
