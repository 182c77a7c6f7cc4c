"""The C-ABI library loads and exports every function the public headers
declare; the struct layouts the binding mirrors match the headers."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
hmat_amd = pytest.importorskip("hmat_amd")


def declared_functions():
    names = set()
    for h in ("hmat.h", "hmat_mi355x.h"):
        src = open(os.path.join(ROOT, "include", "hmat", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = "\n".join(l for l in src.splitlines() if not l.lstrip().startswith("#"))
        for m in re.finditer(r"HMAT_API[^;(]*?\b(\w+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(hmat_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_interface_table_is_filled():
    it = hmat_amd.interface(hmat_amd.HMAT_DOUBLE_PRECISION)
    assert it.value_type == hmat_amd.HMAT_DOUBLE_PRECISION
    for name, typ in hmat_amd.hmat_interface_t._fields_:
        if name in ("value_type", "internal"):
            continue
        v = getattr(it, name)
        ptr = ctypes.cast(v, ctypes.c_void_p).value if not isinstance(v, int) else v
        assert ptr, f"slot {name} is NULL"


def _c_layout(tmp_path, struct, fields):
    src = tmp_path / "layout.c"
    body = "\n".join(f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields)
    src.write_text(f"""#include <stdio.h>
#include <stddef.h>
#include "hmat/hmat_mi355x.h"
int main(void) {{ printf("%zu\\n", sizeof({struct})); {body} return 0; }}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]


@pytest.mark.parametrize("cls,struct", [
    ("hmat_interface_t", "hmat_interface_t"),
    ("hmat_assemble_context_t", "hmat_assemble_context_t"),
    ("hmat_info_t", "hmat_info_t"),
    ("hmat_mi355x_kernel_t", "hmat_mi355x_kernel_t"),
    ("hmat_solve_context_t", "struct hmat_solve_context_t"),
])
def test_struct_layouts_match_header(tmp_path, cls, struct):
    py = getattr(hmat_amd, cls)
    fields = [f for f, _ in py._fields_]
    c = _c_layout(tmp_path, struct, fields)
    assert c[0] == ctypes.sizeof(py)
    assert c[1:] == [getattr(py, f).offset for f in fields]


def test_interface_slot_count():
    # 56 function-pointer slots + value_type + internal (hmat.h:717-1103)
    assert len(hmat_amd.hmat_interface_t._fields_) == 58


def test_version_and_errors_without_gpu():
    lib = hmat_amd.load()
    assert lib.hmat_get_version().startswith(b"hmat-mi355x")
