"""Builds libpjzfdtd.so (all HIP sources under csrc/) for gfx950, in-tree.

Usage:  python -m pjz_amd.build      (or pjz_amd.build.build() from Python)

Every HIP translation unit is compiled with hipcc for gfx950 only, with packed
FP32 instruction selection disabled (-packed-fp32-ops), then linked into
pjz_amd/libpjzfdtd.so next to this file.
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpjzfdtd.so")
ARCH = "gfx950"

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unknown-warning-option",
]


def hipcc():
  for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
    if cand and os.path.exists(cand):
      return cand
  raise RuntimeError("hipcc not found: the gfx950 kernels cannot be built")


def sources():
  return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
  obj = os.path.join(BUILD, os.path.basename(src) + ".o")
  cmd = [hipcc()] + HIPCC_FLAGS + ["-c", src, "-o", obj]
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
  return obj


def build(verbose=False):
  """Compiles every csrc/*.hip and links libpjzfdtd.so; returns its path."""
  os.makedirs(BUILD, exist_ok=True)
  srcs = sources()
  with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
    objs = list(ex.map(_compile, srcs))
  tmp = LIB + ".tmp"
  cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
  r = subprocess.run(cmd, capture_output=True, text=True)
  if r.returncode != 0:
    raise RuntimeError(f"link failed:\n{r.stderr}")
  os.replace(tmp, LIB)
  if verbose:
    print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
  return LIB


if __name__ == "__main__":
  build(verbose=True)
