"""The host jump-ahead for std::mt19937 (forkerrenderer_amd/csrc/mt_jump.cpp):
the characteristic polynomial found by Berlekamp-Massey and the jump
polynomials t^(c 624 K) mod phi for the two chunk lengths K = 256 and 512
(asked for in turn, one table extended after the other), checked against
brute-force generation of the word sequence for a seed the polynomial was
not derived from (tests/mt_jump_check.cpp, built with the host compiler)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "forkerrenderer_amd", "csrc")


def test_jump_polynomials_against_brute_force(tmp_path):
    exe = str(tmp_path / "mt_jump_check")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "mt_jump_check.cpp"),
                        os.path.join(CSRC, "mt_jump.cpp"), "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("build failed:\n" + r.stdout + r.stderr)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
