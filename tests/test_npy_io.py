"""The .npy reader and writer of the C++ facade (miluphpc_amd/host/npy_io.cpp), on the CPU:
arrays numpy writes (every supported element type, vectors and N x 3 matrices, empty and
odd sizes) are read and written back by the facade and must come back unchanged."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
DRIVER = REPO / "miluphpc_amd" / "bin" / "mhpc_rhs_driver"


@pytest.mark.parametrize("dtype,shape", [("<f8", (17,)), ("<f8", (5, 3)), ("<i4", (1000,)), ("|i1", (33,)),
                                         ("<u8", (7,)), ("<i8", (2, 3)), ("<f8", (0,)), ("<f8", (1,))])
def test_npy_round_trip(tmp_path, dtype, shape):
    rng = np.random.default_rng(3)
    a = (rng.normal(size=shape) * 1e3).astype(dtype) if np.dtype(dtype).kind == "f" else \
        rng.integers(-100 if np.dtype(dtype).kind == "i" else 0, 100, size=shape).astype(dtype)
    if np.dtype(dtype) == np.uint64:
        a[...] = np.uint64(0xFFFFFFFFFFFFFFF0)
    src, dst = tmp_path / "a.npy", tmp_path / "b.npy"
    np.save(src, a)
    res = subprocess.run([str(DRIVER), "npycopy", str(src), str(dst)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    b = np.load(dst, allow_pickle=False)
    assert b.dtype == a.dtype and b.shape == a.shape
    assert np.array_equal(a, b)
    # the written header is the standard one: 64-byte aligned data, version 1.0
    raw = dst.read_bytes()
    assert raw[:8] == b"\x93NUMPY\x01\x00"
    hlen = raw[8] | (raw[9] << 8)
    assert (10 + hlen) % 64 == 0


def test_npy_rejects_unsupported(tmp_path):
    src = tmp_path / "a.npy"
    np.save(src, np.zeros(4, np.float32))
    res = subprocess.run([str(DRIVER), "npycopy", str(src), str(tmp_path / "b.npy")], capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 4 and "not supported" in res.stderr
    np.save(src, np.asfortranarray(np.zeros((3, 4))))
    res = subprocess.run([str(DRIVER), "npycopy", str(src), str(tmp_path / "b.npy")], capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 4 and "Fortran" in res.stderr
