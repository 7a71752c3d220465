"""Builds the HIP library (gfx950) in-tree: taichi_lbm/_lib/liblbm_hip.so.

hipcc cross-compiles for gfx950 without a GPU present, so this runs anywhere
ROCm is installed.  Each .hip file compiles to an object in parallel, then one
shared-library link.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent                       # taichi-lbm_amd/
CSRC = ROOT / "csrc"
INCLUDE = ROOT.parent / "include"
LIB_DIR = PKG_DIR / "_lib"
LIB_PATH = LIB_DIR / "liblbm_hip.so"
STAMP_PATH = LIB_DIR / "liblbm_hip.stamp"
SOURCES = ["lbm_fused.hip", "lbm_granular.hip", "lbm_reduce.hip", "lbm_ibm.hip"]
ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # multiply-adds fuse within one source expression only: the per-node
    # arithmetic shared by the one-step, two-step and batched kernels then
    # rounds the same wherever it is inlined (their fields are bit-identical)
    "-ffp-contract=on",
    # packed fp32 ops off: hipcc must not emit v_pk_{add,mul,fma}_f32
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def source_hash() -> str:
    """Hash of every source the library is built from (HIP, headers, C-ABI header)."""
    import hashlib
    h = hashlib.sha256()
    files = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.h")) + [INCLUDE / "lbm_abi.h"]
    for f in files:
        h.update(f.name.encode())
        h.update(f.read_bytes())
    return h.hexdigest()


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return exe


def _filter(out: str) -> str:
    return "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)


def build(verbose: bool = False) -> Path:
    """Compile every HIP source for gfx950 and link the shared library."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    obj_dir = LIB_DIR / "obj"
    obj_dir.mkdir(exist_ok=True)
    cc = hipcc()
    deps = [CSRC / s for s in SOURCES]

    def compile_one(src: Path) -> Path:
        obj = obj_dir / (src.stem + ".o")
        cmd = [cc, *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        msg = _filter(r.stdout + r.stderr)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{msg}")
        if verbose and msg.strip():
            print(msg, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=len(deps)) as ex:
        objs = list(ex.map(compile_one, deps))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{_filter(r.stdout + r.stderr)}")
    os.replace(tmp, LIB_PATH)
    STAMP_PATH.write_text(source_hash() + "\n")
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
