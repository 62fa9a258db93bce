"""Host checks of the certificate's arithmetic (cert_tuv_math.h), compiled for
the host with hipcc and run on the CPU: the per-pair split sequence (the
packed running-ratio builder must equal the level-by-level replay of the split
rule on every input, including ties, powers of two, subnormals, inf and NaN)
and the box's corner range (equal to the eight corners evaluated one by
one)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc is not installed")
def test_split_sequence_matches_replay(tmp_path):
    exe = tmp_path / "split_sequence_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", f"-I{ROOT}/p2c-ccd_amd/csrc", f"-I{ROOT}/include",
                    str(ROOT / "tests/native/split_sequence_check.hip"), "-o", str(exe)], check=True,
                   stderr=subprocess.DEVNULL)
    out = subprocess.run([str(exe)], check=False, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 mismatches" in out.stdout


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc is not installed")
def test_corner_range_matches_definition(tmp_path):
    """The kernels' corner range (least and greatest corner per t-plane from
    the min / max of the u and v terms) equals the min and max over the eight
    corners evaluated one by one, the definition the oracle states."""
    exe = tmp_path / "corner_range_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", f"-I{ROOT}/p2c-ccd_amd/csrc", f"-I{ROOT}/include",
                    str(ROOT / "tests/native/corner_range_check.hip"), "-o", str(exe)], check=True,
                   stderr=subprocess.DEVNULL)
    out = subprocess.run([str(exe)], check=False, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "8000000 evaluations, 0 mismatches" in out.stdout


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc is not installed")
def test_jump_walk_matches_depth_first_walk(tmp_path):
    """The stream kernel's jump walk (the climb to the next box in one jump,
    from the pending-sibling marks and the level masks) evaluates the same
    boxes as the plain depth-first walk and finds the same least done t_lo,
    on uniform, near-contact, coplanar and scaled pairs, past level 64."""
    exe = tmp_path / "jump_walk_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", f"-I{ROOT}/p2c-ccd_amd/csrc", f"-I{ROOT}/include",
                    str(ROOT / "tests/native/jump_walk_check.hip"), "-o", str(exe)], check=True,
                   stderr=subprocess.DEVNULL)
    out = subprocess.run([str(exe)], check=False, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 mismatches" in out.stdout
