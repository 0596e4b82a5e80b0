"""The C-ABI library loads without a GPU and exports every entry point the header declares;
the Python front end binds exactly those."""
import ctypes
import os
import re

from femocs_amd import fgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "femocs_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(fgpu_\w+)\s*\(", text))


def test_library_loads_and_exports_header():
    assert os.path.exists(fgpu.LIB_PATH), "build with `python -m femocs_amd.build`"
    lib = ctypes.CDLL(fgpu.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    bound = {name for name, _, _ in fgpu.API}
    assert bound == declared_symbols()


def test_no_gpu_context_fails_cleanly():
    import torch  # noqa: F401  (only to know whether a GPU is visible)
    if torch.cuda.is_available():
        return
    lib = fgpu.load()
    h = ctypes.c_void_p()
    assert lib.fgpu_ctx_create(0, ctypes.byref(h)) != 0
