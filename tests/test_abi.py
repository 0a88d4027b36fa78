"""The C-ABI library builds, loads and exports every entry point include/tp.h
declares.  CPU only: no HIP call is made."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tp.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(tp_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def _lib_path():
    from tp_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        _lib.build()
    return _lib.LIB_PATH


def test_header_lists_entry_points():
    names = _declared()
    assert "tp_energy" in names and "tp_c3" in names and "tp_anneal" in names
    from tp_amd import _lib
    assert sorted(_lib.EXPORTS) == names


def test_library_exports_every_declared_symbol():
    path = _lib_path()
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(tp_\w+)$", out, flags=re.M))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing


def test_abi_version_and_loader():
    lib = ctypes.CDLL(_lib_path())
    lib.tp_abi_version.restype = ctypes.c_int
    assert lib.tp_abi_version() == 3
    from tp_amd import _lib
    assert _lib.lib().tp_abi_version() == 3


def test_gfx950_code_object_present():
    path = _lib_path()
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", path], capture_output=True, text=True)
    if out.returncode == 0 and out.stdout.strip():
        assert "gfx950" in out.stdout
    else:
        data = open(path, "rb").read()
        assert b"gfx950" in data
