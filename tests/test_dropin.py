"""The drop-in C++ classes (include/zldsp_fft_cfft.hpp, zldsp_fft_rfft.hpp):
the README usage of the reference compiles against them, and, on the GPU, the
README-style program's results match a double-precision DFT, for CFFT<float>,
RFFT<float>, CFFT<double> and RFFT<double>."""
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


def test_dropin_header_rejects_other_types():
    """float and double compile (the reference's two instantiations); other
    element types fail at compile time with a message."""
    inc = os.path.join(ROOT, "include")
    for t, ok in (("float", True), ("double", True), ("long double", False), ("int", False)):
        code = ('#include "zldsp_fft_rfft.hpp"\nint main() { zldsp::fft::CFFT<%s> c(4); '
                'zldsp::fft::RFFT<%s> r(4); return 0; }\n' % (t, t))
        r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-x", "c++", "-I", inc, "-"],
                           input=code, capture_output=True, text=True)
        assert (r.returncode == 0) == ok, (t, r.stderr[-400:])
        if not ok:
            assert "supports float and double" in r.stderr


@pytest.mark.gpu
def test_dropin_program_on_gpu():
    assert os.path.exists(BIN), "build the drop-in example first (zldsp_fft_amd/build.py)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "DROPIN OK" in r.stdout


def test_c_abi_header_is_plain_c():
    """include/zlfft.h compiles as C99 and as C++ (the C-ABI the drop-in
    classes and FFI callers use), and a C caller can name every entry point."""
    import re
    inc = os.path.join(ROOT, "include")
    hdr = open(os.path.join(inc, "zlfft.h")).read()
    names = sorted(set(re.findall(r"\b(zlfft_[a-z0-9_]+)\s*\(", re.sub(r"/\*.*?\*/", "", hdr, flags=re.S))))
    code = '#include "zlfft.h"\n' + "typedef void (*fn_t)(void);\n" + "".join(f"fn_t p_{n} = (fn_t)&{n};\n" for n in names) + "int main(void) { return 0; }\n"
    for lang, std in (("c", "-std=c99"), ("c++", "-std=c++17")):
        r = subprocess.run(["gcc" if lang == "c" else "g++", std, "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                            "-pedantic", "-x", lang, "-I", inc, "-"], input=code, capture_output=True, text=True)
        assert r.returncode == 0, (lang, r.stderr[-600:])
