"""The C-ABI library loads and exports every entry point include/ declares
(no GPU calls). CPU only."""
import ctypes
import os
import re

import openvkl_amd as vkl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for d, _, files in os.walk(os.path.join(ROOT, "include")):
        for f in files:
            if f.endswith(".h"):
                text = open(os.path.join(d, f)).read()
                for m in re.finditer(r"OPENVKL_INTERFACE\s+[^;{]*?\b(vkl[A-Za-z0-9_]+)\s*\(", text):
                    names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 79
    lib = ctypes.CDLL(vkl.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_resolve():
    L = vkl.lib()
    assert L.vklComputeSampleN is not None and L.vklIterateIntervalsBatch is not None


def test_struct_layouts_match_reference():
    # openvkl.h:175-180, 315-320; types.h:56-69
    assert ctypes.sizeof(vkl.VKLInterval) == 20
    assert ctypes.sizeof(vkl.VKLHit) == 12
    assert ctypes.sizeof(vkl.VKLObject) == 16
    assert ctypes.sizeof(vkl.WIDE[16][2]) == 16 * 5 * 4


def test_no_non_inlined_device_functions():
    """The device code has no non-inlined device functions: on gfx950 the
    compiler's branch relaxation in a large callee reused the return-address
    register pair s[30:31] for long jumps, and the callee's return then
    re-entered its own body (a kernel that never ends). Every device function
    is inlined (kernels have no return address)."""
    import glob
    import os
    import re
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "openvkl_amd", "csrc")
    for path in glob.glob(os.path.join(here, "*.hip")) + glob.glob(os.path.join(here, "*.hh")):
        text = re.sub(r"//.*", "", open(path).read())
        assert "__noinline__" not in text, path
        assert "noinline" not in text, path


def test_transfer_function_rows_are_checked():
    """Transfer functions reach the kernels as float32 (r, g, b, opacity)
    rows with their row count: a flat array or a float64 tensor is converted
    (the kernels read 16 bytes per row), and a size that is not a multiple of
    4 raises before any call."""
    import numpy as np
    import pytest
    import torch

    from openvkl_amd import _tf_rows

    flat = torch.arange(32, dtype=torch.float64)
    rows = _tf_rows(flat)
    assert rows.dtype == torch.float32 and tuple(rows.shape) == (8, 4) and rows.is_contiguous()
    assert torch.equal(rows.reshape(-1), flat.float())
    assert _tf_rows(np.arange(12, dtype=np.float64)).shape == (3, 4)
    assert _tf_rows(np.zeros((5, 4), np.float32)).shape == (5, 4)
    for bad in (torch.zeros(6), np.zeros(7, np.float32), torch.zeros(0)):
        with pytest.raises(ValueError):
            _tf_rows(bad)
