"""CPU checks of the boundary: the library loads and exports every entry
point include/mpf.h declares (no device calls)."""
import os
import re

import pytest

from mpf import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="mpf.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpf_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_table():
    from mpf import _nnlib
    assert sorted(_lib.EXPORTS) == _declared()
    assert sorted(_nnlib.NN_EXPORTS) == _declared("mpf_nn.h")


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in _declared() + _declared("mpf_nn.h"):
        assert hasattr(L, name), name
    from mpf import _nnlib
    _nnlib.lib()        # argtypes bind without a device


def test_abi_version_and_errors_without_gpu():
    L = _lib.lib()
    assert L.mpf_abi_version() == 3
    # argument validation happens before any device call
    import ctypes as C
    cfg = _lib.Config()
    cfg.nx = cfg.ny = cfg.nz = 4
    cfg.dim = 5
    h = C.c_void_p()
    rc = L.mpf_create(C.byref(cfg), C.byref(h))
    assert rc == _lib.MPF_EINVAL
    assert b"dim" in L.mpf_last_error()
    with pytest.raises(_lib.MPFError):
        _lib.check(rc)


def test_struct_layouts_match_header(tmp_path):
    """The ctypes structures have the sizes and field offsets that a C
    compiler gives the header's structures (a drift here would corrupt every
    call that passes them)."""
    import ctypes as C
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"mpf_scalars": _lib.Scalars, "mpf_config": _lib.Config, "mpf_stats": _lib.Stats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "mpf.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
