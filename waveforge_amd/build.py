"""Build recipe for libwaveforge_hip.so (gfx950 only).

``python -m waveforge_amd.build`` compiles every ``csrc/*.hip`` file with
hipcc into ``waveforge_amd/libwaveforge_hip.so`` (in-tree, so the built
library travels with the source tree). Flags:

* ``--offload-arch=gfx950``: MI355X only; no other target, no CUDA path.
* ``-ffp-contract=off``: one rounding per fp32 operation, so results match the
  reference's eager tensor ops bit for bit.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed fp32 VALU
  instructions (v_pk_*_f32) are generated.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
INCLUDE = REPO_DIR / "include"
LIB_NAME = "libwaveforge_hip.so"
LIB_PATH = PKG_DIR / LIB_NAME


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libwaveforge_hip.so)")


def compile_flags() -> list[str]:
    return [
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
    ]


def build(verbose: bool = True, jobs: int = 8) -> Path:
    """Compile all HIP sources and link the shared library; returns its path."""
    hipcc = hipcc_path()
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources in {CSRC}")
    obj_dir = PKG_DIR / "_build"
    obj_dir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in sources:
        obj = obj_dir / (src.stem + ".o")
        objs.append(obj)
        cmd = [hipcc, *compile_flags(), "-c", str(src), "-o", str(obj)]
        if verbose:
            print("[waveforge_amd.build]", " ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB_PATH.with_suffix(".so.tmp")
    link = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print("[waveforge_amd.build]", " ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def _wait(item) -> None:
    src, proc = item
    out, _ = proc.communicate()
    text = out.decode(errors="replace")
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{text}")
    # The host pass does not know the packed-fp32 feature flag; that note is harmless.
    lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if lines:
        print("\n".join(lines), flush=True)


if __name__ == "__main__":
    path = build(verbose="-q" not in sys.argv)
    print(path)
