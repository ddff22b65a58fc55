"""The plain-copy reference kernels bench.py times beside its lines
(libzlprobe.so): both forms copy every byte, and nothing past the whole
chunks they are given."""
import ctypes
import os

import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("structure,per_cu", [(0, 1), (1, 1), (1, 2), (1, 3)])
def test_copy_reference_copies_exactly(structure, per_cu):
    lib = ctypes.CDLL(os.path.join(ROOT, "zldsp_fft_amd", "libzlprobe.so"))
    lib.zlprobe_chunk_bytes.restype = ctypes.c_size_t
    chunk = lib.zlprobe_chunk_bytes()
    nchunks = 1237  # more chunks than resident workgroups, not a multiple of them
    src = torch.rand(nchunks * chunk // 4 + 1000, device="cuda")
    dst = torch.full_like(src, -1.0)
    nbytes = nchunks * chunk + 17  # the partial chunk is not copied
    print(f"structure={structure} per_cu={per_cu} chunks={nchunks}", flush=True)
    rc = lib.zlprobe_copy(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), ctypes.c_size_t(nbytes),
                          structure, per_cu, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert rc == 0
    m = nchunks * chunk // 4
    assert torch.equal(dst[:m], src[:m])
    assert bool((dst[m:] == -1.0).all())
