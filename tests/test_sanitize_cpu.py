"""The host layer under AddressSanitizer (with LeakSanitizer) and
UndefinedBehaviorSanitizer, on the CPU: build.py compiles the C-ABI, plan,
host-pipeline and ragged-planning units with -Xarch_host -fsanitize=... into
zldsp_fft_amd/_build/asan/libzlfft.so, and tests/sanitize/host_abi_driver.cpp
walks their argument checks, plan creation and clean-up (without a GPU every
creation fails at the device), ragged validation and item planning, and the
tuning-file parser fed a malformed file.  Any sanitizer report fails the test."""
import os
import subprocess

from zldsp_fft_amd import build as b


def test_host_layer_under_asan_and_ubsan(tmp_path):
    if not os.path.exists(b.DRIVER_BIN):
        b.build()
    tune = tmp_path / "tune.txt"
    tune.write_text("# malformed\nragged_logr 9 4\nragged_logr 14 7\nragged_logr -3 4\nragged_logr 9\nbogus\n"
                    + "7" * 300 + "\nragged_logr 12 16 extra\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([b.DRIVER_BIN, str(tune)], capture_output=True, text=True, timeout=300, env=env)
    out = p.stdout + p.stderr
    assert "ERROR: AddressSanitizer" not in out and "ERROR: LeakSanitizer" not in out, out[-4000:]
    assert "runtime error" not in out, out[-4000:]
    assert p.returncode == 0, out[-4000:]
    assert "0 failed checks" in p.stdout
