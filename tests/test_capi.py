"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/aboria_hip.h (no compute calls: no GPU here)."""
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
            names |= set(re.findall(r"\b(abh_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for required in ["abh_context_create", "abh_cells_update", "abh_search_create",
                     "abh_search_fill", "abh_sparse_op_apply"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    from aboria_amd import _capi
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    from aboria_amd import _capi
    missing = sorted(declared_symbols() - set(_capi.SIGNATURES))
    assert not missing, missing


def test_status_strings_without_gpu():
    from aboria_amd import _capi
    lib = _capi.lib()
    assert lib.abh_status_string(0) == b"ok"
    assert lib.abh_status_string(4) == b"output buffer too small"


def test_cpp_layer_compiles_and_links():
    """include/aboria_hip.hpp compiles with plain g++ (no HIP headers) and
    the example links against libaboria_hip.so."""
    import __graft_entry__ as ge
    out = ge.build_cpp_example()
    assert os.path.exists(out)


def test_search_rejects_host_buffers():
    """Search.fill / run / run_fused check their tensors before any call into
    the library: a host tensor raises ValueError (no GPU needed)."""
    import pytest
    import torch
    import aboria_amd as ab
    s = ab.Search.__new__(ab.Search)
    s.handle = None
    with pytest.raises(ValueError):
        s.fill(torch.empty(16, dtype=torch.int32))
    with pytest.raises(ValueError):
        s.run(None, torch.empty((4, 3), dtype=torch.float64), 0.1)
    with pytest.raises(ValueError):
        s.run_fused(None, torch.empty((4, 3), dtype=torch.float64), 0.1,
                    torch.empty(16, dtype=torch.int32))
