"""The three edits a maintainer makes to the reference to select the
library (a "HIP" arm in dg_optr_kernel_setup, the halo map made public, the
step loop handed to dgk_run), applied by oracle/patch_reference.py to a copy of the reference's
own sources and compiled against libdgk.so and mod_dgk_hip into
oracle/_ref/scale-dg_extraction_hip (built where the reference tree is
present; the program travels like the library).

CPU: the patched program links the library and, with the reference's own
operator types, still prints the reference program's lines.  GPU: with
DGOptrKernel_OptType = "HIP" it prints, character for character, the lines
of the unpatched reference program (tests/golden/golden.json "stock")."""
import os
import subprocess

import pytest

import cases as CS
from conftest import REPO, say

PROG = os.path.join(REPO, "oracle", "_ref", "scale-dg_extraction_hip")
OPDIR = os.path.join(REPO, "operator_data")

pytestmark = pytest.mark.skipif(not os.path.exists(PROG),
                                reason="patched reference program not built (no reference tree)")


def run(tmp_path, cfg, opt):
    os.symlink(OPDIR, tmp_path / "operator_data")
    conf = tmp_path / "input.conf"
    conf.write_text(CS.namelist_text(cfg, opt=opt))
    return subprocess.run([PROG, str(conf)], cwd=tmp_path, capture_output=True, text=True,
                          timeout=600)


def lines_of(r):
    return [l for l in r.stdout.splitlines() if l[:8].strip().isdigit()]


def test_links_the_library():
    r = subprocess.run(["ldd", PROG], capture_output=True, text=True, timeout=60)
    assert "libdgk.so" in r.stdout and "not found" not in r.stdout


@pytest.mark.parametrize("name", sorted(CS.STOCK_RUNS))
def test_reference_types_unchanged(tmp_path, golden, name):
    cfg = CS.STOCK_RUNS[name]
    r = run(tmp_path, cfg, cfg["opt"])
    assert r.returncode == 0, r.stderr
    assert lines_of(r) == golden["stock"][name]["lines"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CS.STOCK_RUNS))
def test_hip_type_prints_the_reference_lines(gpu, tmp_path, golden, name):
    cfg = CS.STOCK_RUNS[name]
    say(f"[patched reference] {name} HIP")
    r = run(tmp_path, cfg, "HIP")
    assert r.returncode == 0, r.stderr
    assert lines_of(r) == golden["stock"][name]["lines"]
