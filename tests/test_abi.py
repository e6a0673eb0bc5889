"""The C-ABI library loads and exports every entry point include/*.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(sa_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    import stipple_amd
    so = stipple_amd.library_path()
    assert os.path.exists(so), "libstipple_amd.so not built"
    lib = ctypes.CDLL(so)
    names = declared()
    assert len(names) >= 8
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_loads():
    import stipple_amd
    L = stipple_amd.lib()
    assert L is not None


def test_engine_refusal_without_device_is_an_error():
    """No silent fallback: with no visible device the engines return an
    error code and a message."""
    import torch
    if torch.cuda.is_available():
        return
    import stipple_amd
    L = stipple_amd.lib()
    buf = ctypes.create_string_buffer(256)
    rc = L.sa_init(0, buf, len(buf))
    assert rc != 0 and len(buf.value) > 0
