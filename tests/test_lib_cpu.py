"""CPU-side checks of the C-ABI library: it loads, exports every function
include/airbox_hip.h declares, parses STL like the reference, and fails
loudly (no fallback) when no gfx950 GPU is present."""
import ctypes

import numpy as np
import pytest

from airboxcalc_amd import abx
from airboxcalc_amd import geometry as G
from oracle import oracle as O


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abx.lib_path())
    names = abx.declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    L = abx.load()
    for n in abx.declared_functions():
        assert getattr(L, n).argtypes is not None or n in ("abx_last_error", "abx_version"), n


@pytest.mark.parametrize("ascii_", [False, True])
def test_parse_stl_matches_oracle(tmp_path, ascii_):
    tris = G.airbox_triangles(G.AirboxSpec(LX=120, LY=60, LZ=90, inlet_d=30, outlet_d=24, outlet_h=20, pleat=16))
    path = str(tmp_path / "b.stl")
    (G.write_stl_ascii if ascii_ else G.write_stl_binary)(path, tris)
    data = open(path, "rb").read()
    a = abx.parse_stl(data)
    b = O.parse_stl(data)
    assert a.dtype == np.float32 and np.array_equal(a, b)
    if not ascii_:
        assert np.array_equal(a, tris.ravel())


def test_parse_stl_error_text():
    with pytest.raises(abx.AbxError, match="STL enthält keine Dreiecke."):
        abx.parse_stl(b"solid x\nfacet normal 0 0 1\nendsolid\n")
    with pytest.raises(O.OracleError, match="STL enthält keine Dreiecke."):
        O.parse_stl(b"solid x\nfacet normal 0 0 1\nendsolid\n")


def test_device_open_fails_loudly_without_gpu():
    # the library's own probe decides (torch's device count can read 0 where HIP works)
    try:
        d = abx.Device(0)
    except abx.AbxError as e:
        assert "GPU" in str(e)
        return
    d.close()
    pytest.skip("a GPU is present")


def test_gated_convolution_rejects_bad_gate_before_any_launch():
    """abx_conv3d_forward_gated validates its gate before touching a device:
    NULL, a stride below the channel count and an unaligned pointer are
    errors with a message (the tensors here are never dereferenced)."""
    L = abx.load()
    p = ctypes.c_void_p(1 << 20)                      # 16-byte aligned, never dereferenced
    args = (None, p, 16, 1, 2, 4, 16, 16, p, None, 16, 3, 0, 0, p, 16)
    assert L.abx_conv3d_forward_gated(*args, None, 16) == -1
    assert b"gate is NULL" in L.abx_last_error()
    assert L.abx_conv3d_forward_gated(*args, p, 8) == -1
    assert b"gate stride" in L.abx_last_error()
    assert L.abx_conv3d_forward_gated(*args, ctypes.c_void_p((1 << 20) + 8), 16) == -1
    assert b"gate stride" in L.abx_last_error()
