"""The C-ABI library loads and exports every entry point include/luna_hip.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "luna_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("lh_create", "lh_destroy", "lh_rhs", "lh_rhs_device", "lh_solve", "lh_batch_solve",
              "lh_strerror"):
        assert s in syms


def test_library_exports_all_declared(capi):
    L = capi.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s
    assert set(capi.EXPORTED) <= set(declared_symbols())


def test_strerror_and_arg_checks(capi):
    L = capi.lib()
    assert L.lh_strerror(capi.LH_E_PPT_RANGE).decode().startswith("field strength")
    h = ctypes.c_void_p()
    assert L.lh_create(ctypes.byref(h), None) == capi.LH_E_ARG
    d = capi.LhDesc()
    d.grid_kind = 7  # unsupported
    assert L.lh_create(ctypes.byref(h), ctypes.byref(d)) == capi.LH_E_ARG


def test_struct_sizes_match_header(capi):
    # layout of the descriptor as a C compiler sees it
    import subprocess
    import tempfile
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "luna_hip.h"\nint main(){printf("%zu %zu %zu %zu %zu",' \
          'sizeof(lh_desc), sizeof(lh_rk_opts), sizeof(lh_stats), offsetof(lh_desc, qdht_inv), offsetof(lh_desc, ppt_x));}'
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "s")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c])
        out = subprocess.check_output([exe]).decode().split()
    sizes = [int(x) for x in out]
    assert sizes == [ctypes.sizeof(capi.LhDesc), ctypes.sizeof(capi.LhRkOpts), ctypes.sizeof(capi.LhStats),
                     capi.LhDesc.qdht_inv.offset, capi.LhDesc.ppt_x.offset]
