"""Builds libfrost_amd.so in-tree with hipcc for gfx950.

    python -m frost_amd.build            # incremental
    python -m frost_amd.build --force    # rebuild everything

Every HIP source is compiled with -ffp-contract=off (the kernels follow the
CPU oracle's operation order, so no FMA contraction) and with packed FP32
instructions disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent
REPO = ROOT.parent
CSRC = ROOT / "csrc"
OUT_DIR = ROOT / "_lib"
OBJ_DIR = OUT_DIR / "obj"
LIB = OUT_DIR / "libfrost_amd.so"
ARCH = os.environ.get("FROST_AMD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-result", f"-I{REPO / 'include'}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]


def sources() -> list[Path]:
    return sorted([p for p in CSRC.iterdir() if p.suffix in (".hip", ".cpp")])


def headers() -> list[Path]:
    return sorted(CSRC.glob("*.hpp")) + [REPO / "include" / "frost_amd.h"]


def _compile(src: Path, force: bool) -> Path:
    obj = OBJ_DIR / (src.name + ".o")
    newest_dep = max([src.stat().st_mtime] + [h.stat().st_mtime for h in headers()])
    if not force and obj.exists() and obj.stat().st_mtime >= newest_dep:
        return obj
    cmd = [HIPCC] + COMMON
    if src.suffix == ".hip":
        cmd += ["-x", "hip"] + DEVICE
    cmd += ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed on {src.name}")
    warn = "\n".join(l for l in res.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if warn:
        sys.stderr.write(warn + "\n")
    return obj


def build(force: bool = False) -> Path:
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    newest = max(o.stat().st_mtime for o in objs)
    if force or not LIB.exists() or LIB.stat().st_mtime < newest:
        tmp = LIB.with_suffix(".so.tmp")
        cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp)] + \
            [str(o) for o in objs]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc link failed")
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    print(build(a.force))
