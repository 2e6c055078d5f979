"""The C-ABI library loads and exports every function that include/*.h
declares (no device calls: CPU only)."""

import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(amigo_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = _declared()
    assert "amigo_ldl_factor" in names and "amigo_ipm_residual" in names and "amigo_elements_eval" in names
    assert len(names) > 40


@pytest.mark.parametrize("name", _declared())
def test_library_exports(name):
    lib = ctypes.CDLL(os.path.join(ROOT, "amigo_amd", "libamigo_amd.so"))
    assert hasattr(lib, name), name


def test_python_binding_missing_library_fails_loudly(tmp_path, monkeypatch):
    from amigo_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(Exception):
        _lib.lib()
