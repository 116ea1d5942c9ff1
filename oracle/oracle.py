"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None
_P = C.c_void_p
_i32, _u32, _f32 = C.c_int32, C.c_uint32, C.c_float

_SIGS = {
    "orc_num_threads": (C.c_int, []),
    "orc_set_threads": (None, [C.c_int]),
    "orc_prng_int": (_i32, [C.POINTER(_i32)]),
    "orc_prng_float": (_f32, [C.POINTER(_i32)]),
    "orc_thread_seed": (_i32, [_i32, _u32, _u32, _u32]),
    "orc_cf3_pack": (_u32, [_f32, _f32, _f32]),
    "orc_cf3_pack_dither": (_u32, [_f32, _f32, _f32, C.POINTER(_i32)]),
    "orc_cf3_unpack": (None, [_u32, _P]),
    "orc_cf3_pack_n": (None, [_P, _P, C.c_int64]),
    "orc_cf3_unpack_n": (None, [_P, _P, C.c_int64]),
    "orc_morton_sum": (_u32, [_u32, _u32, C.c_int]),
    "orc_morton_diff": (_u32, [_u32, _u32, C.c_int]),
    "orc_reverse_bits": (_u32, [_u32]),
    "orc_tree_create": (_P, [C.c_int, C.c_int]),
    "orc_tree_destroy": (None, [_P]),
    "orc_find_node_idx": (_u32, [_P, _u32, _u32]),
    "orc_yield_boundary": (C.c_uint64, [_P, C.c_int, _u32]),
    "orc_add_node": (C.c_int, [_P, C.c_int, C.c_int]),
    "orc_remove_node": (C.c_int, [_P, C.c_int]),
    "orc_increment_timer": (None, [_P]),
    "orc_tree_export": (None, [_P, _P, _P, _P, _P, _P, _P]),
    "orc_tree_set_timer": (None, [_P, _u32]),
    "orc_differentiate": (None, [_P, _P, _P, _P, _P, C.c_int, _u32, _u32, _i32, C.c_int, C.c_int, C.c_int, _f32]),
    "orc_mixed_partials": (None, [_P, _P, _P, _P, C.c_int, _u32, _u32, _i32, C.c_int, C.c_int, _f32]),
    "orc_differentiate_f32": (None, [_P, _P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int, C.c_int, C.c_int, _f32]),
    "orc_mixed_partials_f32": (None, [_P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int, C.c_int, _f32]),
    "orc_copy_boundaries": (None, [_P, _P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int, C.c_int, C.c_int, C.c_int,
                                   _f32, C.c_int]),
    "orc_copy_new_to_old": (None, [_P, _P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int, C.c_int]),
    "orc_copy_to_child": (None, [_P, _u32, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_rasterise": (None, [_P, _P, _u32, _u32, C.c_int, C.c_int, C.c_float, C.c_int, _P]),
    "orc_encode_rgba8": (None, [_P, _P, C.c_int64]),
    "orc_encode_rgba8_f32": (None, [_P, _P, C.c_int64]),
    "orc_encode_rgba8_cf3": (None, [_P, _P, C.c_int64]),
    "orc_init_temp": (None, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]),
    "orc_wave_stage": (None, [_P, _P, _P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int, C.c_int, C.c_int, _f32]),
    "orc_deriv_trace": (None, [_P, _P, _P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int, C.c_int]),
    "orc_christoffel": (None, [_P, _P, _P, C.c_int, _P, _P, C.c_int, _u32, _u32, _i32, C.c_int, C.c_int]),
}


def build() -> Path:
    """Compile liboracle.so with the Makefile next to this file."""
    r = subprocess.run(["make", "-C", str(HERE), "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _i32arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


# ------------------------------------------------------------- primitives
def prng_ints(seed: int, n: int) -> list[int]:
    s = C.c_int32(seed)
    return [lib().orc_prng_int(C.byref(s)) for _ in range(n)]


def prng_floats(seed: int, n: int) -> list[float]:
    s = C.c_int32(seed)
    return [lib().orc_prng_float(C.byref(s)) for _ in range(n)]


def thread_seed(gs: int, x: int, y: int, z: int) -> int:
    return lib().orc_thread_seed(gs, x, y, z)


def cf3_pack(x: float, y: float, z: float) -> int:
    return lib().orc_cf3_pack(x, y, z)


def cf3_pack_dither(x: float, y: float, z: float, seed: int) -> tuple[int, int]:
    s = C.c_int32(seed)
    w = lib().orc_cf3_pack_dither(x, y, z, C.byref(s))
    return w, s.value


def cf3_unpack(w) -> np.ndarray:
    w = np.ascontiguousarray(np.atleast_1d(np.asarray(w, dtype=np.uint32)))
    out = np.empty((w.size, 3), np.float32)
    lib().orc_cf3_unpack_n(w.ctypes.data, out.ctypes.data, w.size)
    return out


def cf3_pack_array(xyz: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float32).reshape(-1, 3))
    out = np.empty(a.shape[0], np.uint32)
    lib().orc_cf3_pack_n(a.ctypes.data, out.ctypes.data, a.shape[0])
    return out


# ----------------------------------------------------------------- octree
class OracleTree:
    def __init__(self, max_slots: int, periodic: bool = False):
        self.max_slots = max_slots
        self.h = lib().orc_tree_create(max_slots, 1 if periodic else 0)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_tree_destroy(self.h)
            self.h = None

    def add_node(self, parent: int, offset: int) -> int:
        return lib().orc_add_node(self.h, parent, offset)

    def remove_node(self, node: int) -> int:
        return lib().orc_remove_node(self.h, node)

    def increment_timer(self):
        lib().orc_increment_timer(self.h)

    def set_timer(self, t: int):
        lib().orc_tree_set_timer(self.h, t)

    def find_node_idx(self, packed: int, max_depth: int = 0xFFFFFFFF) -> int:
        return lib().orc_find_node_idx(self.h, packed, max_depth)

    def yield_boundary(self, node: int, face: int) -> int:
        return lib().orc_yield_boundary(self.h, node, face)

    def export(self) -> dict:
        n = self.max_slots
        out = dict(parent=np.empty(n, np.int32), children=np.empty(8 * n, np.int32),
                   positions=np.empty(n, np.uint32), boundaries=np.empty(6 * n, np.uint64),
                   lifetime=np.empty(n, np.uint32), scalars=np.empty(4, np.int32))
        lib().orc_tree_export(self.h, *(a.ctypes.data for a in out.values()))
        s = out["scalars"]
        out.update(max_depth=int(s[0]), first_free=int(s[1]), final_used=int(s[2]), timer=int(s[3]))
        return out


# ---------------------------------------------------------------- kernels
def differentiate(field, positions, tiles, substep, max_depth, seed, S, P, q, outer_dx, n_slots, with_d2=True):
    d1 = np.zeros(n_slots * S ** 3, np.uint32)
    d2 = np.zeros(n_slots * S ** 3, np.uint32) if with_d2 else None
    t = _i32arr(tiles)
    lib().orc_differentiate(_p(field), _p(d1), _p(d2), _p(positions), _p(t), t.size, substep, max_depth, seed,
                            S, P, q, outer_dx)
    return d1, d2


def mixed_partials(field, positions, tiles, substep, max_depth, seed, S, P, outer_dx, n_slots):
    off = np.zeros(n_slots * S ** 3, np.uint32)
    t = _i32arr(tiles)
    lib().orc_mixed_partials(_p(field), _p(off), _p(positions), _p(t), t.size, substep, max_depth, seed, S, P,
                             outer_dx)
    return off


def differentiate_f32(field, positions, tiles, substep, max_depth, S, P, q, outer_dx, n_slots, with_d2=True):
    """differentiate() without the codec: (n_slots * S^3, 3) float32 arrays
    of the gradient and the Hessian diagonal."""
    d1 = np.zeros((n_slots * S ** 3, 3), np.float32)
    d2 = np.zeros((n_slots * S ** 3, 3), np.float32) if with_d2 else None
    t = _i32arr(tiles)
    lib().orc_differentiate_f32(_p(field), _p(d1), _p(d2), _p(positions), _p(t), t.size, substep, max_depth, S, P,
                                q, outer_dx)
    return d1, d2


def mixed_partials_f32(field, positions, tiles, substep, max_depth, S, P, outer_dx, n_slots):
    """mixed_partials() without the codec: (n_slots * S^3, 3) float32 (xy, xz, yz)."""
    off = np.zeros((n_slots * S ** 3, 3), np.float32)
    t = _i32arr(tiles)
    lib().orc_mixed_partials_f32(_p(field), _p(off), _p(positions), _p(t), t.size, substep, max_depth, S, P,
                                 outer_dx)
    return off


def copy_boundaries(old, new, older, boundaries, tiles, substep, max_depth, copy_only_new, mode, S, P, frac=1.0,
                    equal_copy=False):
    """In place on old/new (float32 arrays).  equal_copy: equal-depth faces
    copy the neighbour cell exactly (the RK4 driver's rule)."""
    t = _i32arr(tiles)
    lib().orc_copy_boundaries(_p(old), _p(new), _p(older), _p(boundaries), _p(t), t.size, substep, max_depth,
                              1 if copy_only_new else 0, mode, S, P, frac, 1 if equal_copy else 0)


def copy_new_to_old(old, new, positions, children, tiles, substep, max_depth, S, P):
    t = _i32arr(tiles)
    lib().orc_copy_new_to_old(_p(old), _p(new), _p(positions), _p(children), _p(t), t.size, substep, max_depth,
                              S, P)


def copy_to_child(buf, offset, child, parent, mode, S, P):
    lib().orc_copy_to_child(_p(buf), offset, child, parent, mode, S, P)


def rasterise(field, children, width, height, S, P, z_slice=1e-9, boxes=True):
    """(height, width, 4) float32 raster of a z-slice (AMR.h:1313-1353)."""
    f = np.ascontiguousarray(field, dtype=np.float32)
    ch = np.ascontiguousarray(children, dtype=np.int32)
    out = np.empty((height, width, 4), np.float32)
    lib().orc_rasterise(_p(f), _p(ch), width, height, S, P, z_slice, 1 if boxes else 0, _p(out))
    return out


def encode_rgba8(raster):
    """RGBA8 words of a float4 raster (image_process.h:14-21)."""
    r = np.ascontiguousarray(raster, dtype=np.float32)
    out = np.empty(r.shape[:-1], np.uint32)
    lib().orc_encode_rgba8(_p(r), _p(out), out.size)
    return out


def encode_rgba8_f32(img):
    """Grey RGBA8 words of a float image (image_process.h:31-38)."""
    a = np.ascontiguousarray(img, dtype=np.float32)
    out = np.empty(a.shape, np.uint32)
    lib().orc_encode_rgba8_f32(_p(a), _p(out), out.size)
    return out


def encode_rgba8_cf3(words):
    """RGB RGBA8 words of compressed_float3 words (image_process.h:22-29)."""
    a = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.empty(a.shape, np.uint32)
    lib().orc_encode_rgba8_cf3(_p(a), _p(out), out.size)
    return out


def init_temp(positions, tiles, n_slots, S, P, outer_size, outer_dx, mul=1.0):
    """(old, new) after __init_temp (kernel.cu:9-43); untouched cells are 0."""
    T = S + 2 * P
    old = np.zeros(n_slots * T ** 3, np.float32)
    new = np.zeros(n_slots * T ** 3, np.float32)
    pos = np.ascontiguousarray(positions, dtype=np.uint32)
    t = _i32arr(tiles)
    lib().orc_init_temp(_p(old), _p(new), _p(pos), _p(t), len(t), S, P, outer_size, outer_dx, mul)
    return old, new


def wave_stage(old, d1, d2, new, positions, tiles, substep, max_depth, corrector, S, P, outer_dt):
    t = _i32arr(tiles)
    lib().orc_wave_stage(_p(old), _p(d1), _p(d2), _p(new), _p(positions), _p(t), t.size, substep, max_depth,
                         1 if corrector else 0, S, P, outer_dt)


def _ptr_table(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def deriv_trace(hd, ho, metric6, positions, tiles, substep, max_depth, S, P, n_slots):
    out = np.zeros(n_slots * S ** 3, np.float32)
    t = _i32arr(tiles)
    lib().orc_deriv_trace(_p(hd), _p(ho), _ptr_table(metric6), _p(out), _p(positions), _p(t), t.size, substep,
                          max_depth, S, P)
    return out


def christoffel(metric6, dg6, compat, positions, tiles, substep, max_depth, seed, S, P, n_slots):
    outs = [np.zeros(n_slots * S ** 3, np.uint32) for _ in range(6)]
    t = _i32arr(tiles)
    lib().orc_christoffel(_ptr_table(metric6), _ptr_table(dg6), _ptr_table(outs), compat, _p(positions), _p(t),
                          t.size, substep, max_depth, seed, S, P)
    return outs


# ------------------------------------------------------------------- BSSN
_SIGS_BSSN = {
    "orc_bssn_point": (None, [_P, _P, _P, _P, _P, _P, _P]),
    "orc_bssn_stage_from_words": (None, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int,
                                         _P, _P, C.c_int, C.c_int]),
    "orc_bssn_constraints_from_words": (None, [_P, _P, _P, _P, _P, _P, C.c_int, _P, C.c_int, C.c_int, _P]),
    "orc_bssn_stage_from_words_fmt": (None, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int,
                                             _P, _P, C.c_int, C.c_int, C.c_int]),
    "orc_bssn_constraints_from_words_fmt": (None, [_P, _P, _P, _P, _P, C.c_int, _P, C.c_int, C.c_int, _P,
                                                   C.c_int]),
    "orc_bssn64_stage": (None, [_P, _P, _P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int, _P, _P, _P, C.c_int, C.c_int,
                                C.c_int]),
    "orc_bssn64_constraints": (None, [_P, _P, _P, C.c_int, _P, _P, C.c_int, C.c_int, C.c_int, _P]),
    "orc_fd64_float_coefficients": (None, [C.c_int]),
    "orc_bssn_point_adv": (None, [_P, _P, _P, _P, _P, _P]),
    "orc_bssn_stage_from_words_adv": (None, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int, _u32, _u32,
                                             C.c_int, _P, _P, C.c_int, C.c_int]),
    "orc_bssn64_set_dissipation": (None, [C.c_double, C.c_int]),
    "orc_fd_adv": (None, [_P, _P, _P, _P, _P, _P, _P, C.c_int, _u32, _u32, C.c_int, C.c_int, _f32, _f32, C.c_int]),
    "orc_ghost64": (None, [_P, _P, _P, C.c_int, _P, _P, C.c_int, _u32, _u32, _f32, C.c_int, C.c_int, C.c_int,
                           C.c_int]),
    "orc_copy_back64": (None, [_P, _P, _P, C.c_int, _P, _P, _P, C.c_int, _u32, _u32, C.c_int, C.c_int]),
}
_SIGS.update(_SIGS_BSSN)

# eta, shift_coeff, chi_floor, advect_shift, slicing (0 = 1+log, 1 = harmonic)
BSSN_DEFAULT = (2.0, 0.75, 1e-4, 1, 0)
HESS_VARS = (0, 1, 2, 3, 4, 5, 6, 17, 18, 19, 20)


def _params(p):
    return np.ascontiguousarray(np.asarray(p, dtype=np.float64))


def bssn_point(u, d1, dd, G=None, params=BSSN_DEFAULT):
    """RHS (24) and constraints (H, M_x, M_y, M_z) at one point in FP64."""
    u = np.ascontiguousarray(u, np.float64)
    d1 = np.ascontiguousarray(d1, np.float64)
    dd = np.ascontiguousarray(dd, np.float64)
    G = None if G is None else np.ascontiguousarray(G, np.float64)
    p = _params(params)
    rhs = np.zeros(24)
    hm = np.zeros(4)
    lib().orc_bssn_point(_p(u), _p(d1), _p(dd), _p(G), _p(p), _p(rhs), _p(hm))
    return rhs, hm


def bssn_point_adv(u, d1, dd, advin, params=BSSN_DEFAULT):
    """RHS (24) at one point with the advection of every field given
    (advin[24]: upwinded advection plus dissipation, as the FD writes it)."""
    u, d1, dd, a = (np.ascontiguousarray(x, np.float64) for x in (u, d1, dd, advin))
    p = _params(params)
    rhs = np.zeros(24)
    lib().orc_bssn_point_adv(_p(u), _p(d1), _p(dd), _p(p), _p(a), _p(rhs))
    return rhs


def fd_adv(field, shift3, positions, tiles, substep, max_depth, S, P, outer_dx, sigma, advect, n_slots):
    """The FD's advection + dissipation output of one padded float32 field
    (interior-only float32 array), shift3: the three padded shift fields."""
    out = np.zeros(n_slots * S ** 3, np.float32)
    t = _i32arr(tiles)
    f = np.ascontiguousarray(field, np.float32)
    b = [np.ascontiguousarray(x, np.float32) for x in shift3]
    lib().orc_fd_adv(_p(f), _p(b[0]), _p(b[1]), _p(b[2]), _p(out), _p(positions), _p(t), t.size, substep, max_depth,
                     S, P, outer_dx, sigma, 1 if advect else 0)
    return out


def set_dissipation(sigma, upwind=True):
    """Upwinded advection with Kreiss-Oliger dissipation of strength sigma in
    the FP64 tile scheme (bssn64_stage, Bssn64Sim); upwind=False: centred."""
    lib().orc_bssn64_set_dissipation(float(sigma), 1 if upwind else 0)


def ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[None if a is None else a.ctypes.data for a in arrs])


# ------------------------------------------------------ FP64 scheme driver
def bssn64_stage(state, old, acc, new, positions, tiles, substep, max_depth, stage, inv_dx, dtd,
                 params=BSSN_DEFAULT, S=16, P=3, q=2):
    """One FP64 RK4 stage with exact derivatives on the listed tiles (lists of
    24 float64 padded arrays; acc/new written on interior cells)."""
    t = _i32arr(tiles)
    # the arrays must outlive the call: bind them before taking addresses
    ix, dt, pa = np.asarray(inv_dx, np.float64), np.asarray(dtd, np.float64), _params(params)
    lib().orc_bssn64_stage(ptrs(state), ptrs(old), ptrs(acc), ptrs(new), _p(positions), _p(t), t.size, substep,
                           max_depth, stage, _p(ix), _p(dt), _p(pa), S, P, q)


def ghost64(old, new, older, boundaries, tiles, substep, max_depth, frac, mode, S=16, P=3, equal_copy=True):
    t = _i32arr(tiles)
    lib().orc_ghost64(ptrs(old), ptrs(new), None if older is None else ptrs(older), len(new), _p(boundaries), _p(t),
                      t.size, substep, max_depth, frac, mode, S, P, 1 if equal_copy else 0)


def copy_back64(old, new, older, positions, children, tiles, substep, max_depth, S=16, P=3):
    t = _i32arr(tiles)
    lib().orc_copy_back64(ptrs(old), ptrs(new), None if older is None else ptrs(older), len(new), _p(positions),
                          _p(children), _p(t), t.size, substep, max_depth, S, P)


def bssn64_constraints(state, positions, tiles, inv_dx, params=BSSN_DEFAULT, S=16, P=3, q=2):
    """{rms H, rms |M|, max |H|, cells} over the listed tiles, exact derivatives."""
    t = _i32arr(tiles)
    out = np.zeros(4 * t.size)
    ix, pa = np.asarray(inv_dx, np.float64), _params(params)
    lib().orc_bssn64_constraints(ptrs(state), _p(positions), _p(t), t.size, _p(ix), _p(pa), S, P, q, _p(out))
    o = out.reshape(-1, 4)
    c = o[:, 2].sum()
    return np.sqrt(o[:, 0].sum() / c), np.sqrt(o[:, 1].sum() / c), o[:, 3].max(), c


class Bssn64Sim:
    """The RK4 octree driver's schedule in FP64 with exact derivatives: per
    sub-step copy-back (keeping older), then every active depth coarse first,
    four stages each with the stage-time ghost fill (equal-depth faces
    copied, coarser faces interpolated), then the counter.  The product's
    am_bssn_sim runs the same schedule in fp32 on dithered words."""

    def __init__(self, tree_export, outer_dx, cfl, mode=0, params=BSSN_DEFAULT, S=16, P=3, q=2):
        e = tree_export
        self.e, self.S, self.P, self.q, self.mode, self.params = e, S, P, q, mode, params
        self.n = len(e["positions"])
        T = S + 2 * P
        self.cells = self.n * T ** 3
        self.old = [np.zeros(self.cells) for _ in range(24)]
        self.new = [np.zeros(self.cells) for _ in range(24)]
        self.acc = [np.zeros(self.cells) for _ in range(24)]
        self.inp = [np.zeros(self.cells) for _ in range(24)]
        self.older = [np.zeros(self.cells) for _ in range(24)] if mode == 2 else None
        odx = np.float32(outer_dx)
        self.inv_dx = [float(np.float32(2 ** d) / odx) for d in range(9)]
        odt = np.float32(odx * np.float32(cfl))
        self.dtd = [float(odt / np.float32(2 ** d)) for d in range(9)]
        used = [i for i, p in enumerate(e["positions"]) if int(p) != 0xFFFFFFFF]
        self.all = used
        self.by_depth = {}
        for i in used:
            self.by_depth.setdefault(int(e["positions"][i]) >> 27, []).append(i)
        self.max_depth = e["max_depth"]
        self.timer = 0
        self.leaves = [i for i in used if all(c == -1 for c in e["children"][8 * i:8 * i + 8])]

    def set_state(self, fields):
        for v in range(24):
            self.old[v][:] = fields[v]
            self.new[v][:] = fields[v]

    def fill_ghosts(self):
        for d in range(1, self.max_depth + 1):
            if d in self.by_depth:
                ghost64(self.old, self.new, None, self.e["boundaries"], self.by_depth[d], 0, self.max_depth, 1.0,
                        self.mode, self.S, self.P, True)

    def substep(self):
        e, md, t = self.e, self.max_depth, self.timer
        copy_back64(self.old, self.new, self.older, e["positions"], e["children"], self.all, t, md, self.S, self.P)
        for d in range(md + 1):
            if d not in self.by_depth or (t & ((1 << (md - d)) - 1)) != 0:
                continue
            tiles = self.by_depth[d]
            for s in range(4):
                st = self.old if s == 0 else self.new
                if s == 0:
                    bssn64_stage(st, self.old, self.acc, self.new, e["positions"], tiles, t, md, s, self.inv_dx,
                                 self.dtd, self.params, self.S, self.P, self.q)
                else:
                    # the stage reads neighbouring cells of its input while
                    # writing new: give it a copy of the listed tiles
                    idx = np.asarray(tiles)
                    for v in range(24):
                        self.inp[v].reshape(self.n, -1)[idx] = self.new[v].reshape(self.n, -1)[idx]
                    bssn64_stage(self.inp, self.old, self.acc, self.new, e["positions"], tiles, t, md, s,
                                 self.inv_dx, self.dtd, self.params, self.S, self.P, self.q)
                if d > 0:
                    ghost64(self.old, self.new, self.older, e["boundaries"], tiles, t, md, (.5, .5, 1., 1.)[s],
                            self.mode, self.S, self.P, True)
        self.timer = (t + 1) % (1 << md)

    def constraints(self):
        return bssn64_constraints(self.new, self.e["positions"], self.leaves, self.inv_dx, self.params, self.S,
                                  self.P, self.q)
