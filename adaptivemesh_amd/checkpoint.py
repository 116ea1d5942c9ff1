"""Reader of the drivers' checkpoint files (the format of am_checkpoint_header
in include/adaptivemesh_host.h), for inspecting a saved run on any machine:
no GPU and no native library needed.

    header, sections = read("run.ckpt")
    header["timer"], sections["positions"], sections["old"][var, slot]

Field sections come back shaped (n_vars, used slots, T, T, T) for the RK4
drivers and (used slots, T, T, T) for SimpleWave, with T = S + 2P; the
SimpleWave derivative words (max_slots, S, S, S).
"""
from __future__ import annotations

import numpy as np

MAGIC = b"AMCKPT\x00\x01"
KINDS = {1: "simplewave", 2: "bssn", 3: "scalar_wave"}

HEADER = np.dtype([
    ("magic", "V8"), ("kind", "<u4"),
    ("size_domain", "<i4"), ("padding_domain", "<i4"), ("fd_order", "<i4"),
    ("outer_size", "<f4"), ("outer_dx", "<f4"), ("outer_dt", "<f4"),
    ("max_slots", "<u4"), ("periodic", "<u4"), ("max_depth", "<i4"),
    ("first_free_slot", "<u4"), ("final_used_slot", "<u4"), ("timer", "<u4"),
    ("seed", "<i4"), ("n_vars", "<u4"), ("flags", "<u4"), ("n_sections", "<u4"),
])
assert HEADER.itemsize == 76

DTYPES = {"parent": "<i4", "children": "<i4", "positions": "<u4", "boundaries": "<u8", "lifetime": "<u4",
          "old": "<f4", "new": "<f4", "older": "<f4", "d1": "<u4", "d2diag": "<u4"}


class CheckpointError(ValueError):
    pass


def read(path) -> tuple[dict, dict]:
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < HEADER.itemsize or data[:8] != MAGIC:
        raise CheckpointError(f"{path}: not a checkpoint file")
    h = np.frombuffer(data, HEADER, count=1)[0]
    header = {k: (h[k].item() if k != "magic" else bytes(h[k])) for k in HEADER.names}
    header["kind_name"] = KINDS.get(header["kind"], "unknown")
    S, P = header["size_domain"], header["padding_domain"]
    T = S + 2 * P
    used, n, nv = header["final_used_slot"] + 1, header["max_slots"], header["n_vars"]
    at, sections = HEADER.itemsize, {}
    for _ in range(header["n_sections"]):
        if len(data) - at < 24:
            raise CheckpointError(f"{path}: truncated section table")
        name = data[at:at + 16].rstrip(b"\x00").decode()
        size = int(np.frombuffer(data, "<u8", count=1, offset=at + 16)[0])
        at += 24
        if size > len(data) - at:
            raise CheckpointError(f"{path}: truncated section {name}")
        a = np.frombuffer(data, DTYPES.get(name, "u1"), count=size // np.dtype(DTYPES.get(name, "u1")).itemsize,
                          offset=at)
        if name in ("old", "new", "older"):
            a = a.reshape((used, T, T, T) if header["kind"] == 1 else (nv, used, T, T, T))
        elif name in ("d1", "d2diag"):
            a = a.reshape(n, S, S, S)
        sections[name] = a
        at += size
    if at != len(data):
        raise CheckpointError(f"{path}: trailing bytes after the last section")
    return header, sections
