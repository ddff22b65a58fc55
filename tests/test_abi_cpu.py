"""CPU-only checks of the C-ABI boundary: the library loads, exports every
function include/zlfft.h declares, and rejects invalid arguments before any
GPU work (the reference has no error paths; its invalid inputs are UB)."""
import ctypes
import os
import re

import pytest

import zldsp_fft_amd as zf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(ROOT, "include", "zlfft.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(zlfft_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    lib = zf.load_library()
    names = _declared()
    assert "zlfft_plan_create" in names and len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(zf.EXPORTS)


def test_library_exports_only_the_c_abi():
    """Built with hidden visibility: the dynamic symbol table holds exactly the
    zlfft_* functions that include/zlfft.h declares."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", zf.lib_path], capture_output=True, text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if len(line.split()) >= 3}
    assert syms == set(_declared()), sorted(syms ^ set(_declared()))[:10]


def test_version():
    assert zf.load_library().zlfft_version() >= 100


def test_constants_match_header():
    hdr = open(os.path.join(ROOT, "include", "zlfft.h")).read()
    consts = dict(re.findall(r"#define (ZLFFT_[A-Z0-9_]+) (\d+)", hdr))
    assert int(consts["ZLFFT_AOS"]) == zf.AOS and int(consts["ZLFFT_SOA"]) == zf.SOA
    assert int(consts["ZLFFT_FORWARD"]) == zf.FORWARD and int(consts["ZLFFT_BACKWARD"]) == zf.BACKWARD
    assert int(consts["ZLFFT_KIND_CFFT"]) == zf.KIND_CFFT and int(consts["ZLFFT_KIND_RFFT"]) == zf.KIND_RFFT
    assert int(consts["ZLFFT_MAX_CFFT_ORDER"]) == zf.MAX_CFFT_ORDER
    assert int(consts["ZLFFT_MAX_CFFT_ORDER_F64"]) == zf.MAX_CFFT_ORDER_F64
    assert int(consts["ZLFFT_PREC_F64"]) == zf.PREC_F64
    assert int(consts["ZLFFT_PREC_BF16_MFMA_SPLIT3"]) == zf.PREC_BF16_MFMA_SPLIT3


def _create(kind, prec, order):
    lib = zf.load_library()
    h = ctypes.c_void_p()
    rc = lib.zlfft_plan_create(ctypes.byref(h), kind, prec, order)
    return rc, h, lib.zlfft_last_error().decode()


@pytest.mark.parametrize("kind,prec,order,code", [
    (zf.KIND_RFFT, 0, 0, 1),           # RFFT order 0: the reference underflows rfft_order - 1
    (zf.KIND_CFFT, 0, 29, 1),          # beyond the supported maximum
    (zf.KIND_RFFT, 0, 30, 1),
    (7, 0, 10, 5),                     # unknown kind
    (zf.KIND_CFFT, 9, 10, 6),          # unknown precision
    (zf.KIND_CFFT, 6, 28, 1),          # FP64 above its maximum order (27)
    (zf.KIND_RFFT, 6, 29, 1),
    (zf.KIND_RFFT, 2, 12, 6),          # MFMA precisions are CFFT-only
])
def test_plan_create_rejects_invalid_arguments(kind, prec, order, code):
    rc, h, msg = _create(kind, prec, order)
    assert rc == code and not h.value and msg


def test_null_arguments_are_reported():
    lib = zf.load_library()
    assert lib.zlfft_plan_create(None, 0, 0, 4) == 2
    assert lib.zlfft_plan_destroy(None) == 2
    assert lib.zlfft_plan_get_size(None) == 0 and lib.zlfft_plan_get_order(None) == 0
    assert lib.zlfft_cfft_exec(None, 0, 0, None, None, 0, None, None, 1, None) == 2
    assert lib.zlfft_rfft_forward(None, None, 0, None, None) == 2
    assert "NULL" in lib.zlfft_last_error().decode()


def test_python_mirror_validates_order():
    with pytest.raises(ValueError):
        zf.CFFT(-1)
    with pytest.raises(zf.ZlfftError) as e:
        zf.RFFT(0)
    assert e.value.code == 1


def test_ragged_desc_numpy_image_matches_struct():
    """RAGGED_DESC_DTYPE (bulk descriptor tables) has the layout of
    zlfft_ragged_desc: same size and field offsets as the ctypes mirror."""
    import ctypes
    d = zf.RAGGED_DESC_DTYPE
    assert d.itemsize == ctypes.sizeof(zf.RaggedDesc) == 48
    for name, _ in zf.RaggedDesc._fields_:
        assert d.fields[name][1] == getattr(zf.RaggedDesc, name).offset, name


def test_copy_reference_library_exports():
    """libzlprobe.so (the plain-copy reference that bench.py times beside its
    lines) loads without a GPU and exports its two functions, nothing else;
    invalid arguments are rejected before any launch."""
    import subprocess
    path = os.path.join(ROOT, "zldsp_fft_amd", "libzlprobe.so")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if len(line.split()) >= 3}
    assert syms == {"zlprobe_copy", "zlprobe_chunk_bytes"}, syms
    lib = ctypes.CDLL(path)
    lib.zlprobe_chunk_bytes.restype = ctypes.c_size_t
    assert lib.zlprobe_chunk_bytes() == 32768
    assert lib.zlprobe_copy(None, None, ctypes.c_size_t(1 << 20), 0, 2, None) == -1
    buf = ctypes.create_string_buffer(64)
    assert lib.zlprobe_copy(buf, buf, ctypes.c_size_t(100), 0, 2, None) == -1  # less than one chunk
    assert lib.zlprobe_copy(buf, buf, ctypes.c_size_t(1 << 20), 7, 2, None) == -1  # unknown structure
