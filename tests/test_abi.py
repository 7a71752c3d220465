"""The C-ABI library loads and exports every entry point include/lbm_abi.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from taichi_lbm import _hip as H
from taichi_lbm._build import LIB_PATH

HEADER = Path(__file__).resolve().parent.parent / "include" / "lbm_abi.h"


def declared():
    txt = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(lbm_\w+)\s*\(", txt, flags=re.M)))


def test_header_lists_entry_points():
    names = declared()
    assert "lbm_run" in names and "lbm_finish" in names and len(names) >= 18
    assert sorted(H.exported_symbols()) == names


def test_library_exports_every_declared_symbol():
    assert LIB_PATH.exists(), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(str(LIB_PATH))
    for name in declared():
        assert hasattr(lib, name), name


def test_abi_version_and_error_string():
    lib = H.lib()
    assert lib.lbm_abi_version() == H.ABI_VERSION
    assert isinstance(lib.lbm_last_error(), bytes)


def test_struct_layout_matches_header(tmp_path):
    # every ctypes field sits at the offset the C compiler gives it
    structs = {"lbm_neem_rule": H.NeemRule, "lbm_lattice": H.Lattice, "lbm_fields": H.Fields,
               "lbm_gfields": H.GFields, "lbm_ib_particles": H.IBParticles}
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_invalid_arguments_fail_cleanly():
    # argument validation happens before any HIP call
    lib = H.lib()
    L = H.Lattice()
    L.nx, L.ny, L.nc = 4, 4, 3           # 3 components: unsupported
    L.plane_stride = 16
    F = H.Fields()
    rc = lib.lbm_run(ctypes.byref(L), ctypes.byref(F), None, 1, 0, 1, None)
    assert rc == -1
    assert b"num_components" in lib.lbm_last_error()
    L.nc = 1
    rc = lib.lbm_run(ctypes.byref(L), ctypes.byref(F), None, 1, 0, 1, None)
    assert rc == -1 and b"missing" in lib.lbm_last_error()
    with pytest.raises(H.LBMError):
        H.check(rc, "lbm_run")


def test_no_alloc_entry_point_refuses_a_state_beyond_its_workspace():
    # the size check comes before any HIP call (nothing runs on a device)
    lib = H.lib()
    L = H.Lattice()
    L.nx, L.ny, L.nc = 4096, 2048, 1
    L.plane_stride = 4096 * 2048
    L.tau_mode = H.TAU_POWER_LAW
    L.tau_min, L.tau_max = 0.51, 2.0
    F = H.Fields()
    for name in ("tau", "code"):
        setattr(F, name, 256)
    F.f[0] = F.f[1] = F.macro[0] = F.macro[1] = 256
    out = ctypes.c_int32(-1)
    rc = lib.lbm_run_blocked_noalloc(ctypes.byref(L), ctypes.byref(F), None, 2, 0, 1, 0, ctypes.byref(out), None)
    assert rc == -1
    assert b"exceeds the workspace" in lib.lbm_last_error()
    assert out.value == -1
