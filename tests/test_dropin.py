"""The drop-in C++ classes (include/zldsp_fft_cfft.hpp, zldsp_fft_rfft.hpp):
the README usage of the reference compiles against them, and, on the GPU, the
README-style program's results match a double-precision DFT."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "dropin", "_build", "dropin_example")


def test_dropin_headers_compile():
    src = os.path.join(ROOT, "tests", "dropin", "dropin_example.cpp")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_dropin_header_rejects_double():
    code = '#include "zldsp_fft_cfft.hpp"\nint main() { zldsp::fft::CFFT<double> c(4); return 0; }\n'
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-x", "c++", "-I", os.path.join(ROOT, "include"), "-"],
                       input=code, capture_output=True, text=True)
    assert r.returncode != 0 and "CFFT<float>" in r.stderr


@pytest.mark.gpu
def test_dropin_program_on_gpu():
    assert os.path.exists(BIN), "build the drop-in example first (zldsp_fft_amd/build.py)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "DROPIN OK" in r.stdout
