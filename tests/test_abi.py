"""The C ABI: library loads, exports every declared entry point, the ctypes
struct matches the C layout, and calls fail cleanly without a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

from tests.common import ROOT
from tisphi_amd import _lib

HEADER = os.path.join(ROOT, "include", "tisphi.h")


def declared():
    txt = open(HEADER).read()
    return set(re.findall(r"^\s*(?:int|const char\*|void)\s+(tsph_\w+)\s*\(", txt, re.M))


def test_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 20
    L = _lib.load()
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) == names
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (tsph_\w+)", out))
    assert names <= exported
    assert not (exported - names), f"undeclared exports: {exported - names}"


def test_abi_version():
    assert _lib.load().tsph_abi_version() == 1


def test_scene_desc_layout_matches_c(tmp_path):
    fields = [f for f, _ in _lib.SceneDesc._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){",
           'printf("%zu\\n", sizeof(tsph_scene_desc));']
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof(tsph_scene_desc, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(_lib.SceneDesc)
    for f, off in zip(fields, vals[1:]):
        assert getattr(_lib.SceneDesc, f).offset == off, f


def test_field_and_pass_ids_match_header():
    txt = open(HEADER).read()
    for name, (fid, _, _) in _lib.FIELDS.items():
        m = re.search(r"TSPH_F_" + name.upper() + r"\s*=\s*(\d+)", txt)
        assert m and int(m.group(1)) == fid, name
    for name, pid in _lib.PASSES.items():
        m = re.search(r"TSPH_PASS_" + name.upper() + r"\s*=\s*(\d+)", txt)
        assert m and int(m.group(1)) == pid, name


def test_argument_errors_without_device():
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.tsph_engine_create(9, 2, 10, 10, 1, ctypes.byref(h)) == _lib.ERR_INVALID_ARG
    assert b"method" in L.tsph_last_error()
    assert L.tsph_engine_create(1, 4, 10, 10, 1, ctypes.byref(h)) == _lib.ERR_INVALID_ARG
    assert L.tsph_engine_destroy(None) == _lib.ERR_INVALID_ARG
    n = ctypes.c_int(-1)
    assert L.tsph_device_count(ctypes.byref(n)) == 0 and n.value >= 0


@pytest.mark.skipif(_lib.device_count() > 0, reason="checks the no-device path")
def test_engine_refuses_to_run_without_gpu():
    with pytest.raises(RuntimeError, match="GPU"):
        _lib.Engine(1, 2, 100, 100)
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.tsph_engine_create(1, 2, 10, 10, 1, ctypes.byref(h)) == _lib.ERR_HIP
