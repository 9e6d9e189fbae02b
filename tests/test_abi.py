"""The C-ABI library loads and exports every symbol include/sofa_amd.h declares (no GPU
needed: no compute call is made)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(sofa_amd_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "sofa_amd_fem_add_dforce" in names and "sofa_amd_engine_step" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from sofa_amd import capi
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) == declared_symbols()


def test_module_identity_without_gpu():
    from sofa_amd import capi
    L = capi.lib()
    assert L.sofa_amd_module_name() == b"SofaMI355X"
    assert L.sofa_amd_module_version()


def test_product_does_not_import_oracle():
    """No file of the package references the oracle."""
    pkg = os.path.join(ROOT, "sofa_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "import oracle" not in text and "from oracle" not in text and "liboracle" not in text, fn
