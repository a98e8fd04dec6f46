"""Accuracy of the kinetics' exp/log (ember_amd/csrc/emx_math.h) on the host.

The device kernels evaluate Arrhenius rates, the 1/Kc standard-state factors, Troe
falloff and log T with emx_exp / emx_log.  This compiles the same header for the host
with g++ and measures both against long double over the argument ranges the kinetics
use (tests/cpp/math_accuracy.cpp).
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exp_log_accuracy(tmp_path):
    exe = str(tmp_path / "math_accuracy")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "ember_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "math_accuracy.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    vals = {}
    for line in out.splitlines():
        k, v = line.split()[:2]
        vals[k] = float(v)
    assert vals["special"] == 1.0
    assert vals["exp_max_ulp"] <= 1.5, out
    assert vals["log_max_ulp"] <= 2.0, out
    assert vals["exp_denormal_ulp"] <= 1.0, out
    assert vals["log_denormal_ulp"] <= 1.0, out
