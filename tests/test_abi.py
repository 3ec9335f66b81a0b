"""CPU checks of the C ABI boundary: the library loads and exports every entry
point that include/meshoid_hip.h declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "meshoid_hip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mhip_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "mhip_knn_hsml" in names and "mhip_apply_D" in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol():
    from meshoid_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTED) == declared()
    assert lib.mhip_version() == 10000


def test_context_creation_without_gpu_fails_cleanly():
    import torch

    from meshoid_amd import _lib

    if torch.cuda.is_available():
        return
    ctx = ctypes.c_void_p()
    rc = _lib.lib.mhip_ctx_create(0, None, ctypes.byref(ctx))
    assert rc in (_lib.ERR_HIP, _lib.ERR_VALUE)


def test_no_oracle_import_in_product():
    pkg = os.path.join(ROOT, "meshoid_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\w+)", src, re.M)
