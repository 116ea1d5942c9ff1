"""FP64 restatement of the learned prolongation (the checker).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports it.

The network and its training pairs have no counterpart in the reference.
They are written here from their definitions, to check the GPU kernels:

* pairs: fine padded tiles (24 fields) are restricted to coarse cells (mean
  of the 2x2x2 children, the restriction of AMR.h:975-995); for each coarse
  cell c whose 4^3 stencil c-1..c+2 lies inside the tile's padded region,
  the input is that stencil, per field shifted by the centre value mu and
  scaled by sigma = max |stencil - mu| + 1e-3 |mu| + 1e-7 (the floor keeps a
  field that is flat to rounding from turning its rounding noise into O(1)
  inputs); the target is the
  8 fine children minus the tricubic baseline (4-point Lagrange through the
  stencil at offsets -1/4, +1/4 per axis), divided by sigma;
* MLP 1536 -> 128 -> 128 -> 192 with exact GELU (z Phi(z)), parameters in
  one array W1[1536][128] b1 W2[128][128] b2 W3[128][192] b3;
* MSE loss mean((y - t)^2) over batch x 192, its gradients by backprop;
* Adam with bias correction.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

N_IN, N_HID, N_OUT = 1536, 128, 192
SHAPES = [(N_IN, N_HID), (N_HID,), (N_HID, N_HID), (N_HID,), (N_HID, N_OUT), (N_OUT,)]
N_PARAMS = sum(int(np.prod(s)) for s in SHAPES)


def unpack(params):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(params[o:o + n], np.float64).reshape(s))
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(p, np.float64).ravel() for p in parts])


def init_params(seed=0):
    """He-style uniform init (fan-in), zero biases; W3 scaled down so the
    untrained network starts near the tricubic baseline."""
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (N_IN, N_HID)) * math.sqrt(6.0 / N_IN)
    w2 = rng.uniform(-1, 1, (N_HID, N_HID)) * math.sqrt(6.0 / N_HID)
    w3 = rng.uniform(-1, 1, (N_HID, N_OUT)) * math.sqrt(6.0 / N_HID) * 0.1
    return pack([w1, np.zeros(N_HID), w2, np.zeros(N_HID), w3, np.zeros(N_OUT)])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_prime(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(params, x):
    w1, b1, w2, b2, w3, b3 = unpack(params)
    x = np.asarray(x, np.float64)
    z1 = x @ w1 + b1
    h1 = gelu(z1)
    z2 = h1 @ w2 + b2
    h2 = gelu(z2)
    y = h2 @ w3 + b3
    return dict(z1=z1, h1=h1, z2=z2, h2=h2, y=y)


def loss(params, x, t):
    y = forward(params, x)["y"]
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def grads(params, x, t):
    """Loss and its gradient with respect to the flat parameter array."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    x = np.asarray(x, np.float64)
    f = forward(params, x)
    B = x.shape[0]
    dy = 2.0 * (f["y"] - np.asarray(t, np.float64)) / (B * N_OUT)
    dw3 = f["h2"].T @ dy
    db3 = dy.sum(0)
    dz2 = (dy @ w3.T) * gelu_prime(f["z2"])
    dw2 = f["h1"].T @ dz2
    db2 = dz2.sum(0)
    dz1 = (dz2 @ w2.T) * gelu_prime(f["z1"])
    dw1 = x.T @ dz1
    db1 = dz1.sum(0)
    return float(np.mean((f["y"] - t) ** 2)), pack([dw1, db1, dw2, db2, dw3, db3])


def adam(params, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new params, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


# ------------------------------------------------------------- pairs
def lag4(t):
    """4-point Lagrange weights through nodes -1, 0, 1, 2 at t."""
    nodes = (-1.0, 0.0, 1.0, 2.0)
    w = []
    for i, xi in enumerate(nodes):
        p = 1.0
        for m, xm in enumerate(nodes):
            if m != i:
                p *= (t - xm) / (xi - xm)
        w.append(p)
    return np.array(w)


W_CHILD = np.stack([lag4(-0.25), lag4(0.25)])  # [child offset j][stencil node]


def coarse_range(S, P):
    """Coarse cells c (per axis, in the tile's coarse index) whose stencil
    c-1..c+2 restricts from fine cells inside the padded tile."""
    return range(max(0, -((P - 2) // 2)), (S + P - 6) // 2 + 1)


def tile_pairs(fine, slot, S, P):
    """(X, T, mu, sigma) for every coarse cell of one tile: fine is a list of
    24 padded arrays (slots * T^3), X (n, 1536), T (n, 192), mu/sigma (n, 24)."""
    T_ = S + 2 * P
    cr = list(coarse_range(S, P))
    xs, ts, mus, sgs = [], [], [], []
    f = [np.asarray(a, np.float64).reshape(-1, T_, T_, T_)[slot] for a in fine]
    for cz in cr:
        for cy in cr:
            for cx in cr:
                X = np.zeros((24, 4, 4, 4))  # [v][z][y][x]
                Y = np.zeros((24, 2, 2, 2))
                for v in range(24):
                    for k in range(4):
                        for j in range(4):
                            for i in range(4):
                                z0 = 2 * (cz - 1 + k) + P
                                y0 = 2 * (cy - 1 + j) + P
                                x0 = 2 * (cx - 1 + i) + P
                                X[v, k, j, i] = f[v][z0:z0 + 2, y0:y0 + 2, x0:x0 + 2].mean()
                    z0, y0, x0 = 2 * cz + P, 2 * cy + P, 2 * cx + P
                    Y[v] = f[v][z0:z0 + 2, y0:y0 + 2, x0:x0 + 2]
                mu = X[:, 1, 1, 1]
                sg = np.abs(X - mu[:, None, None, None]).reshape(24, -1).max(1) + 1e-3 * np.abs(mu) + 1e-7
                base = np.einsum("ai,bj,ck,vkji->vcba", W_CHILD, W_CHILD, W_CHILD, X)  # [v][jz][jy][jx]
                xs.append(((X - mu[:, None, None, None]) / sg[:, None, None, None]).reshape(-1))
                ts.append(((Y - base) / sg[:, None, None, None]).reshape(-1))
                mus.append(mu)
                sgs.append(sg)
    return np.array(xs), np.array(ts), np.array(mus), np.array(sgs)


def children_from(X, Yn, mu, sigma):
    """Fine children (n, 24, 8) from normalized stencils and outputs."""
    n = X.shape[0]
    Xr = X.reshape(n, 24, 4, 4, 4) * sigma[:, :, None, None, None] + mu[:, :, None, None, None]
    base = np.einsum("ai,bj,ck,nvkji->nvcba", W_CHILD, W_CHILD, W_CHILD, Xr).reshape(n, 24, 8)
    return base + Yn.reshape(n, 24, 8) * sigma[:, :, None]


# ------------------------------------------------------- learned ghost fill
def _read_pos(w, rel, S, P):
    """AMR.h:736-745 in float32, for one axis (rel_scale 1/2)."""
    f = np.float32
    r = (f(w) - f(P) + f(.5)) * (f(1) / f(S)) - f(.5)
    r = r * f(.5) + f(rel)
    return (r + f(.5)) * f(S) + f(P) - f(.5)


def _slab_cells(d, S, P):
    """(w0, w1, w2) arrays of the ghost slab cells of face d, in the kernel's
    order."""
    c = np.arange(S * S * P)
    if d in (0, 3):
        l, a, b = c % P, (c // P) % S, c // (P * S)
        return (l + P + S if d == 0 else l), a + P, b + P
    if d in (1, 4):
        a, l, b = c % S, (c // S) % P, c // (S * P)
        return a + P, (l + P + S if d == 1 else l), b + P
    a, b, l = c % S, (c // S) % S, c // (S * S)
    return a + P, b + P, (l + P + S if d == 2 else l)


def learned_fill(old, new, older, boundaries, positions, tiles, substep, max_depth, frac, params, threshold, S, P,
                 orc):
    """The learned stage fill on float32 arrays (24 fields, in place on
    `new`): the tricubic fill (the float restatement, equal-depth faces
    copied) everywhere, then, on faces one level coarser than a non-root
    target, whose coarse cells form blocks of 64 samples (z, y, x order), the
    MLP's children in FP64 for every cell whose stencil avoids the target's
    edge and corner ghosts, unless the block's largest restriction residual
    exceeds `threshold`.  Returns ((blocks learned, blocks kept tricubic,
    cells tricubic, cells evaluated), the tricubic-only result, the flat
    indices of the cells given learned values)."""
    T = S + 2 * P
    C = T ** 3
    for v in range(24):
        orc.copy_boundaries(old[v].copy(), new[v], None if older is None else older[v], boundaries, tiles, substep,
                            max_depth, True, 2, S, P, frac, equal_copy=True)
    tric = [a.copy() for a in new]
    stats = [0, 0, 0, 0]
    written = []
    bnd = np.asarray(boundaries, np.uint64)
    O = np.stack([np.asarray(a, np.float64) for a in old])
    N = np.stack([np.asarray(a, np.float64) for a in new])
    Q = None if older is None else np.stack([np.asarray(a, np.float64) for a in older])
    si = np.arange(64)
    ox, oy, oz = si & 3, (si >> 2) & 3, si >> 4
    for node in tiles:
        for d in range(6):
            bb = int(bnd[node * 6 + d])
            nd = bb & 0xFFFFFFFF
            tgt = nd & 0xFFFFFF
            tgt = tgt - (1 << 24) if tgt & (1 << 23) else tgt
            cd = -1 if (nd >> 28) == 15 else nd >> 28
            dd = (nd >> 24) & 15
            if tgt <= 0 or cd <= 0 or dd != 1 or substep & ((1 << (max_depth - cd)) - 1):
                continue
            rel = orc.cf3_unpack(bb >> 32)[0]
            ty = (float((substep >> (max_depth - cd)) & 1) + frac) * 0.5
            tc = (0.5 * ty * (ty - 1), (1 - ty) * (1 + ty), 0.5 * ty * (ty + 1)) if older is not None else \
                (0.0, 1 - ty, ty)
            n = d % 3
            clo, ncl = [], []
            for k in range(3):
                wlo = P if k != n else (P + S if d < 3 else 0)
                whi = P + S - 1 if k != n else (2 * P + S - 1 if d < 3 else P - 1)
                a = int(np.floor(_read_pos(wlo, rel[k], S, P) + np.float32(.5)))
                z = int(np.floor(_read_pos(whi, rel[k], S, P) + np.float32(.5)))
                clo.append(a)
                ncl.append(z - a + 1)
            NS = ncl[0] * ncl[1] * ncl[2]
            sa = np.arange(NS)
            cx = clo[0] + sa % ncl[0]
            cy = clo[1] + (sa // ncl[0]) % ncl[1]
            cz = clo[2] + sa // (ncl[0] * ncl[1])
            px = cx[:, None] - 1 + ox[None, :]
            py = cy[:, None] - 1 + oy[None, :]
            pz = cz[:, None] - 1 + oz[None, :]
            outs = sum(((p_ < P) | (p_ >= P + S)).astype(int) for p_ in (px, py, pz))
            inb = (px >= 0) & (px < T) & (py >= 0) & (py < T) & (pz >= 0) & (pz < T) & (outs <= 1)
            valid = inb.all(1)
            r = tgt * C + (np.clip(pz, 0, T - 1) * T + np.clip(py, 0, T - 1)) * T + np.clip(px, 0, T - 1)
            X = O[:, r] * tc[1] + N[:, r] * tc[2]
            if Q is not None:
                X = X + Q[:, r] * tc[0]
            X = np.where(inb[None], X, 0.0).transpose(1, 0, 2)  # (NS, 24, 64)
            mu = X[:, :, 21]
            sg = np.abs(X - mu[:, :, None]).max(2) + 1e-3 * np.abs(mu) + 1e-7
            Xn = (X - mu[:, :, None]) / sg[:, :, None]
            y = forward(params, Xn.reshape(NS, -1))["y"].reshape(NS, 24, 8)
            base = np.einsum("ai,bj,ck,nvkji->nvcba", W_CHILD, W_CHILD, W_CHILD,
                             Xn.reshape(NS, 24, 4, 4, 4)).reshape(NS, 24, 8)
            ch = base + y
            m = ch.mean(2)
            rs = np.abs(m * sg).max(1)
            ch = ch - m[:, :, None]
            stats[3] += NS
            stats[2] += int((~valid).sum())
            passed = np.zeros(NS, bool)
            for b0 in range(0, NS, 64):
                blk = slice(b0, min(NS, b0 + 64))
                r = rs[blk][valid[blk]].max() if valid[blk].any() else 0.0
                if r > threshold:
                    stats[1] += 1
                else:
                    stats[0] += 1
                    passed[blk] = True
            w = _slab_cells(d, S, P)
            cl, child = [], np.zeros(S * S * P, int)
            for k in range(3):
                rp = np.array([_read_pos(int(x), rel[k], S, P) for x in range(T)], np.float32)[w[k]]
                c_ = np.floor(rp + np.float32(.5)).astype(int)
                cl.append(c_)
                child |= (rp > c_).astype(int) << k
            sidx = (cl[0] - clo[0]) + ncl[0] * ((cl[1] - clo[1]) + ncl[1] * (cl[2] - clo[2]))
            keep = valid[sidx] & passed[sidx]
            wn = node * C + (w[2] * T + w[1]) * T + w[0]
            written.append(wn[keep])
            for v in range(24):
                val = mu[sidx, v] + sg[sidx, v] * ch[sidx, v, child]
                new[v][wn[keep]] = val[keep].astype(np.float32)
    return tuple(stats), tric, (np.concatenate(written) if written else np.zeros(0, np.int64))
