"""The raster's correctly rounded division (div_rn, div_rn.h: two
Markstein corrections from RN(1/b)) equals IEEE single division bit for bit
on 60 million quotients over the raster's operand ranges and whole binades
(tests/emu/div_rn_check.cpp), under AddressSanitizer/UBSan; a planted
weakening (no correction step) is caught.  CPU only."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "emu", "div_rn_check.cpp")
INC = os.path.join(ROOT, "webgpu-ocean_amd", "csrc")


def build(tmp_path, *flags):
    exe = str(tmp_path / "div_rn_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", *flags, "-I" + INC, SRC,
                    "-o", exe], check=True)
    return exe


def test_div_rn_equals_ieee_division(tmp_path):
    exe = build(tmp_path, "-fsanitize=address,undefined", "-fno-sanitize-recover=all")
    out = subprocess.run([exe, "20000000"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert " 0 mismatches" in out.stdout


def test_uncorrected_product_is_caught(tmp_path):
    exe = build(tmp_path, "-DNO_CORRECTION")
    out = subprocess.run([exe, "2000000"], capture_output=True, text=True)
    assert out.returncode == 1 and "mismatch" in out.stdout
