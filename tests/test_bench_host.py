"""bench.py's host-side contract, without a GPU: more than one GPU is refused
with a message (the stepper is a single-GPU kernel), and the per-cell
algorithmic byte count matches a count made by hand from the data each cell
moves per pass."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("gpus", ["2", "8"])
def test_more_than_one_gpu_is_refused(gpus):
  r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", gpus],
                     capture_output=True, text=True, timeout=120)
  assert r.returncode == 2
  assert "one GPU" in r.stderr
  assert r.stdout.strip() == ""


@pytest.mark.parametrize("zz,lo,hi,expect", [
    # 512x512x128, CPML 16 + 16: E, H read and written (48 B), cb read (12 B);
    # psi of 32 of 128 cells, 4 components read and written (8 B a cell on
    # average); the absorber's 3 values per column over 128 cells.
    (128, 16, 16, 48 + 12 + 32 * 32 / 128 + 12 / 128),
    (96, 16, 16, 48 + 12 + 32 * 32 / 96 + 12 / 96),
    (64, 0, 0, 48 + 12 + 12 / 64),
])
def test_algorithmic_bytes_per_cell(zz, lo, hi, expect):
  assert bench.algorithmic_bytes_per_cell(zz, lo, hi) == pytest.approx(expect)
  # the figures bench.py printed for the two heights it reports
  if (zz, lo, hi) == (128, 16, 16):
    assert round(bench.algorithmic_bytes_per_cell(zz, lo, hi), 4) == 68.0938
  if (zz, lo, hi) == (96, 16, 16):
    assert round(bench.algorithmic_bytes_per_cell(zz, lo, hi), 4) == 70.7917
