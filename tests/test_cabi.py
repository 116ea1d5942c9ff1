"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares; structs agree in size with the Python mirrors."""
import ctypes as C
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(am_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(am):
    L = am.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers all of them
    assert not sorted(names - set(am.exported_symbols()))


STRUCTS = {"Geometry": "am_geometry", "Mesh": "am_mesh", "FdBatch": "am_fd_batch", "FieldBatch": "am_field_batch",
           "BssnParams": "am_bssn_params", "BssnFields": "am_bssn_fields", "BssnSimOptions": "am_bssn_sim_options",
           "SwaveFields": "am_swave_fields", "SwaveSimOptions": "am_swave_sim_options",
           "TuneReport": "am_tune_report"}


def test_struct_sizes(am, tmp_path):
    """Every ctypes mirror has the size the C compiler gives the header's
    struct (a C program built from include/*.h prints them)."""
    import subprocess
    from pathlib import Path
    inc = Path(__file__).resolve().parents[1] / "include"
    src = tmp_path / "sizes.c"
    body = "".join(f'    printf("%zu\\n", sizeof({c}));\n' for c in STRUCTS.values())
    src.write_text('#include <stdio.h>\n#include "adaptivemesh.h"\n#include "adaptivemesh_host.h"\n'
                   f"int main(void) {{\n{body}    return 0;\n}}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-I", str(inc), str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    for (py, c), n in zip(STRUCTS.items(), sizes):
        assert C.sizeof(getattr(am, py)) == n, (py, c, C.sizeof(getattr(am, py)), n)


def test_geometry_init_matches_reference_constants(am):
    g = am.geometry(28, 2, 0, 100.0, 0.4)  # reference defaults, AMR.h:10-17
    import numpy as np
    assert g.fd_order == 4
    assert np.float32(g.outer_dx) == np.float32(100.0) / np.float32(28)
    assert np.float32(g.outer_dt) == np.float32(g.outer_dx) * np.float32(0.4)
    import pytest
    with pytest.raises(am.AMError):
        am.geometry(16, 3, 8)
    with pytest.raises(am.AMError):
        am.geometry(16, 2, 6)


def test_error_paths_without_gpu(am):
    L = am.lib()
    g = am.geometry(16, 3, 6)
    # null mesh / batch rejected before any device work
    assert L.am_fd_batched(None, None, C.byref(g), None) == -1
    assert L.am_ghost_fill_batched(None, None, C.byref(g), 1, 0, None) == -1
    assert b"null" in L.am_last_error_string()
    bad = am.geometry(16, 3, 6)
    bad.size_domain = 20
    m = am.Mesh()
    m.positions = 1
    m.tiles = 1
    m.n_tiles = 1
    b = am.FdBatch()
    b.n_vars = 1
    b.field[0] = 16
    b.d1[0] = 16
    assert L.am_fd_batched(C.byref(b), C.byref(m), C.byref(bad), None) == -2
    b.n_vars = 33
    assert L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), None) == -7
    b.n_vars = 1
    b.field[0] = 18
    assert L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), None) == -6
