"""Field helpers for the tests: analytic fields sampled on padded tiles."""
import numpy as np


def cell_coords(positions, slot, S, P, outer_size):
    """Physical coordinates (x, y, z) of every padded cell of `slot`, shape
    (T, T, T) each, indexed [z, y, x] (kernel.cu:21-22)."""
    T = S + 2 * P
    pos = int(positions[slot])
    d = pos >> 27
    c = np.zeros(3)
    for i in range(d):
        o = (pos >> (3 * i)) & 7
        c += np.array([1 if o & 1 else -1, 1 if o & 2 else -1, 1 if o & 4 else -1]) * 0.25 / 2 ** i
    c *= outer_size
    h = outer_size / S / 2 ** d
    r = (np.arange(T) + 0.5 - T / 2) * h
    z, y, x = np.meshgrid(c[2] + r, c[1] + r, c[0] + r, indexing="ij")
    return x, y, z


def sample(positions, n_slots, S, P, outer_size, fn, slots=None, dtype=np.float32):
    """fn(x, y, z) on every padded cell of every used slot."""
    T = S + 2 * P
    out = np.zeros((n_slots, T, T, T), dtype)
    for s in range(n_slots) if slots is None else slots:
        if int(positions[s]) == 0xFFFFFFFF:
            continue
        x, y, z = cell_coords(positions, s, S, P, outer_size)
        out[s] = fn(x, y, z)
    return out.reshape(-1)


def interior(a, n_slots, S, P):
    T = S + 2 * P
    return a.reshape(n_slots, T, T, T)[:, P:P + S, P:P + S, P:P + S]


def used_slots(exp):
    return [i for i, p in enumerate(exp["positions"]) if int(p) != 0xFFFFFFFF]


def depth_of(exp, s):
    return int(exp["positions"][s]) >> 27


def outer_push_layers(exp, n_slots, tiles, S, P, layers, axes=7):
    """Mask (n_slots * T^3) of the ghost cells that a push of the innermost
    `layers` ghost layers leaves alone: the outer P - layers layers of each
    face slab of `tiles` whose neighbour has the same depth, on the axes in
    `axes` (bits 1 x, 2 y, 4 z)."""
    T = S + 2 * P
    m = np.zeros((n_slots, T, T, T), bool)  # [slot, z, y, x]
    inner = slice(P, P + S)
    for s in tiles:
        for g in range(6):
            nd = int(exp["boundaries"][s * 6 + g]) & 0xFFFFFFFF
            tgt = nd & 0xFFFFFF
            if tgt & 0x800000:
                tgt -= 1 << 24
            own = nd >> 28
            if tgt < 0 or own == 15 or own == 0 or (nd >> 24) & 15 or not (axes >> (g % 3)) & 1:
                continue
            band = slice(P + S + layers, T) if g < 3 else slice(0, P - layers)
            idx = [inner, inner, inner]  # x, y, z
            idx[g % 3] = band
            m[s, idx[2], idx[1], idx[0]] = True
    return m.reshape(-1)


def fd_input(orc, e, n, tiles, state_in, old, stage, direct, S, P, frac):
    """The fields an RK4 stage's FD reads: with direct faces (an axis mask),
    stages 1-3 take the ghost cells across equal-depth faces of those axes
    from the neighbours' interiors (the equal-depth copies of the oracle's
    fill at time `frac`); otherwise the stage input as it lies."""
    if not direct or stage == 0:
        return state_in
    eq = outer_push_layers(e, n, tiles, S, P, 0, direct)
    out = []
    for v in range(len(state_in)):
        filled = state_in[v].copy()
        orc.copy_boundaries(old[v].copy(), filled, None, e["boundaries"], tiles, 0, e["max_depth"], True, 0, S, P,
                            frac, equal_copy=True)
        x = state_in[v].copy()
        x[eq] = filled[eq]
        out.append(x)
    return out


def unpushed_layers(e, n, tiles, S, P, layers, direct, stage):
    """Mask of the equal-depth ghost cells an RK4 stage leaves alone: the
    outer P - layers layers of every face, and with direct faces in stages
    0-2 every layer of the faces of those axes."""
    if stage == 3 or not direct:
        return outer_push_layers(e, n, tiles, S, P, layers)
    return outer_push_layers(e, n, tiles, S, P, 0, direct) | outer_push_layers(e, n, tiles, S, P, layers, 7 & ~direct)
