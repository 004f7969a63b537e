"""The LDS-ring kernels issue their global -> LDS copies from inline asm that
sets M0 (the LDS address) and lists it as clobbered.  The compiler warns
that it does not preserve reserved registers such as M0 across an asm
statement, so the copies are only correct while nothing else in these
kernels keeps a value in M0: no indirect register indexing (movrel), no
messages, no GWS or add-TID LDS operations and no LDS copies of the
compiler's own.  This compiles the two sources that use the copies to gfx950
assembly (no GPU needed) and checks exactly that: every M0 write is one of
the asm pairs, immediately followed by its copy, and no other instruction
names or implicitly reads M0.
"""

import os
import re
import subprocess

import pytest

from pjz_amd import build as B

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pjz_amd", "csrc")
SOURCES = ["kernels_stream2_ring.hip", "kernels_adjoint_stream2.hip"]

# Mnemonic fragments of instructions that read M0 without naming it.
IMPLICIT_M0 = ("movrel", "sendmsg", "_gws_", "addtid", "s_ttracedata")


def device_asm(src, tmp_path):
  try:
    hipcc = B.hipcc()
  except RuntimeError:
    pytest.skip("hipcc not found")
  out = tmp_path / (src + ".s")
  flags = [f for f in B.HIPCC_FLAGS if f not in ("-fPIC", "-fvisibility=hidden")]
  r = subprocess.run([hipcc] + flags + ["--cuda-device-only", "-S", "-I", CSRC,
                                        os.path.join(CSRC, src), "-o", str(out)],
                     capture_output=True, text=True)
  assert r.returncode == 0, r.stderr[-2000:]
  return out.read_text()


def instructions(text):
  """Instruction lines (mnemonic first), without labels, directives and comments."""
  out = []
  for line in text.splitlines():
    s = line.split(";")[0].strip()
    if not s or s.startswith(".") or s.endswith(":"):
      continue
    out.append(s)
  return out


def check_m0(ins, src):
  m0 = re.compile(r"\bm0\b")
  writes = copies = 0
  for i, s in enumerate(ins):
    op = s.split()[0]
    assert not any(f in op for f in IMPLICIT_M0), f"{src}: {s!r} reads M0"
    if op.startswith("global_load_lds") or (op.startswith("buffer_load") and " lds" in s):
      copies += 1
      assert i > 0 and re.fullmatch(r"s_mov_b32 m0, s\d+", ins[i - 1]), \
          f"{src}: LDS copy {s!r} not preceded by its own M0 write (got {ins[i - 1]!r})"
    if m0.search(s):
      writes += 1
      assert re.fullmatch(r"s_mov_b32 m0, s\d+", s), f"{src}: {s!r} uses M0"
      assert i + 1 < len(ins) and ins[i + 1].startswith("global_load_lds"), \
          f"{src}: M0 write {s!r} not followed by an LDS copy"
  assert copies > 0 and writes == copies, (src, writes, copies)


@pytest.mark.parametrize("src", SOURCES)
def test_m0_only_feeds_the_lds_copies(src, tmp_path):
  check_m0(instructions(device_asm(src, tmp_path)), src)


GOOD = """\
\t;;#ASMSTART
\ts_mov_b32 m0, s4
\tglobal_load_lds_dwordx4 v[2:3], off
\t;;#ASMEND
\tv_add_f32_e32 v1, v2, v3
"""


@pytest.mark.parametrize("bad", [
    "\ts_mov_b32 m0, s7\n\tv_movrels_b32_e32 v1, v2\n",        # indirect indexing
    "\ts_mov_b32 m0, s7\n\ts_sendmsg sendmsg(MSG_DEALLOC_VGPRS)\n",
    "\tv_readlane_b32 s1, v2, m0\n",                              # M0 named elsewhere
    "\ts_mov_b32 m0, -1\n",                                       # a write of the compiler's
    "\tglobal_load_lds_dword v[2:3], off\n",                      # a copy with no M0 write
])
def test_checker_flags_other_m0_uses(bad):
  """The checker passes the asm pair alone and fails each planted misuse."""
  check_m0(instructions(GOOD), "good")
  with pytest.raises(AssertionError):
    check_m0(instructions(GOOD + bad), "planted")
