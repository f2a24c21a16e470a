"""The C-ABI library loads and exports every entry point include/mcpgpu/mcpgpu.h declares.
No compute calls here (no GPU); only calls that touch no device are made."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mcpgpu" / "mcpgpu.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(mcpg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("mcpg_bspline_mi_f32", "mcpg_ap_mi_i32", "mcpg_maxmin_f32", "mcpg_mcp_f32", "mcpg_ratio_f32",
                 "mcpg_bspline_mi_backward_f32"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import mcpnet_amd
    lib = mcpnet_amd.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared()) == set(mcpnet_amd.EXPORTS)


def test_status_strings_and_no_device_error():
    import mcpnet_amd
    lib = mcpnet_amd.load()
    assert lib.mcpg_status_string(0) == b"ok"
    assert lib.mcpg_status_string(-3) == b"non-finite input"
    h = ctypes.c_void_p()
    st = lib.mcpg_create(ctypes.byref(h), 0)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        assert st == -6   # MCPG_ERR_NO_DEVICE, reported, not crashed
        with pytest.raises(mcpnet_amd.MCPGError):
            mcpnet_amd.Handle(0)
    else:
        assert st == 0
        lib.mcpg_destroy(h)
    assert lib.mcpg_create(None, 0) == -1


def test_null_handle_is_rejected():
    import mcpnet_amd
    lib = mcpnet_amd.load()
    assert lib.mcpg_maxmin_f32(None, 0, None, 0, None, 0, None, 1, 1, 1, 0, None, 0) == -1
    assert lib.mcpg_bspline_mi_f32(None, 0, None, 1, 1, 1, 10, 3, 0, None, 1) == -1
