"""Mesh recipes shared by the tests: sequences of AMR::add_node calls."""


def uniform(depth):
    """Every octant refined down to `depth` (1 + 8 + ... + 8^depth slots)."""
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


def nested(levels):
    """The reference BSSN_test mesh (kernel.cu:71-83): the 8 root children,
    then each refined towards the centre `levels - 1` more times."""
    ops, nxt = [], 1
    children = []
    for c in range(8):
        children.append(nxt)
        ops.append((0, c))
        nxt += 1
    for _ in range(1, levels):
        for c in range(8):
            ops.append((children[c], (~c) & 7))
            children[c] = nxt
            nxt += 1
    return ops, nxt


def graded():
    """Depth-1 base, one octant refined to depth 2, one of its children to
    depth 3: coarse-fine faces at every level."""
    ops, n = uniform(1)
    for c in range(8):
        ops.append((1, c))
    for c in range(8):
        ops.append((9 + 7, c))
    return ops, n + 16


def apply(tree, ops):
    return [tree.add_node(p, c) for p, c in ops]
