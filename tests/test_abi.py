"""C ABI: the library loads, exports every function include/simplexmf.h
declares, and reports errors through return codes + mf_last_error (CPU only;
no device calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "simplexmf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol(smf):
    names = declared()
    assert len(names) >= 25
    L = C.CDLL(smf.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python mirror binds all of them
    assert set(names) <= set(smf._SIGNATURE_NAMES)


def test_error_codes(smf):
    m = smf.Mesh.kuhn_cube(2)
    with pytest.raises(smf.MFError, match="degree"):
        smf.Operator(m, 4, host_only=True)
    with pytest.raises(smf.MFError, match="precision"):
        smf.Operator(m, 2, precision=7, host_only=True)
    bad = smf.Mesh(m.vertices, np.where(m.cells == 0, 10 ** 6, m.cells))
    with pytest.raises(smf.MFError, match="missing vertex"):
        smf.Operator(bad, 1, host_only=True)
    op = smf.Operator(m, 2, host_only=True)
    L = smf.lib()
    # device entry points refuse a host-only operator
    rc = L.mf_vmult(op.handle, C.c_void_p(16), C.c_void_p(32), None)
    assert rc != 0 and b"host_only" in L.mf_last_error()
    rc = L.mf_vmult(op.handle, C.c_void_p(16), C.c_void_p(16), None)
    assert rc != 0 and b"alias" in L.mf_last_error()
    assert L.mf_vmult(None, None, None, None) != 0


def test_inverted_cell_rejected(smf):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    with pytest.raises(smf.MFError, match="non-positive"):
        smf.Operator(smf.Mesh(v, np.array([[0, 2, 1, 3]], np.int32)), 1, host_only=True)
