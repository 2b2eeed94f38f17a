"""The C-ABI library loads on the CPU and exports every entry point that
include/ymrt.h declares (no GPU calls are made)."""
import ctypes
import os
import re

import numpy as np

import ymrt_amd

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ymrt.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ymrt_[a-z_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(ymrt_amd.EXPORTS)


def test_library_exports_every_symbol():
    L = ctypes.CDLL(ymrt_amd.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_abi_version_and_error_string():
    L = ymrt_amd.lib()
    assert L.ymrt_abi_version() == 8
    assert isinstance(L.ymrt_last_error(), bytes)


def test_save_png_pixels(tmp_path):
    """ymrt_save_png writes what SavePngImage writes (src/FileIO.cpp:167-194):
    extension replaced by .png, folders created, clamp(v,0,1)*255+0.5
    truncated, bottom row first in the image (flipped in the file)."""
    import struct
    import zlib
    rng = np.random.default_rng(4)
    img = rng.uniform(-0.2, 1.2, (5, 7, 4)).astype(np.float32)
    img[0, 0] = (0.5 / 255, 1.5 / 255, 254.5 / 255, 1.0)    # rounding boundaries
    ymrt_amd.save_png(tmp_path / "a" / "b" / "frame.exr", img)
    data = (tmp_path / "a" / "b" / "frame.png").read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    w, h, depth, ctype = struct.unpack(">IIBB", data[16:26])
    assert (w, h, depth, ctype) == (7, 5, 8, 6)
    p, idat = 8, b""
    while p < len(data):
        n = struct.unpack(">I", data[p:p + 4])[0]
        if data[p + 4:p + 8] == b"IDAT":
            idat += data[p + 8:p + 8 + n]
        p += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + w * 4)
    assert np.all(raw[:, 0] == 0)
    got = raw[:, 1:].reshape(h, w, 4)[::-1]
    expect = (np.clip(img, 0.0, 1.0) * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)
    assert np.array_equal(got, expect)
    assert tuple(got[0, 0]) == (1, 2, 255, 255)
