"""The C++ drop-in adapter and pybind stub of INTEGRATION.md compile against
the reference's headers (include/sparse_ldl.h, sparse_cholesky.h,
csr_matrix.h, vector.h) and this repository's include/amigo_amd.h, so the
documented boundary is one that builds.  CPU only; skipped where the
reference tree or the compilers are absent."""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/include"
MPI_INC = "/opt/conda/include"


def _block(name):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(rf"<!-- {name}:begin -->\s*```cpp\n(.*?)```\s*<!-- {name}:end -->", text, flags=re.S)
    assert m, f"INTEGRATION.md has no {name} block"
    return m.group(1)


def _compile(tmp_path, source, extra=()):
    if not os.path.isdir(REF_INC) or not os.path.exists(os.path.join(MPI_INC, "mpi.h")):
        pytest.skip("reference headers or MPI headers not present")
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    src = tmp_path / "adapter.cpp"
    src.write_text(source)
    cmd = [gxx, "-std=c++17", "-fsyntax-only", "-fopenmp", "-I", os.path.join(ROOT, "include"),
           "-I", REF_INC, "-I", MPI_INC, *extra, str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_adapter_compiles_against_the_reference_headers(tmp_path):
    src = _block("adapter") + """
template class amigo::SparseLDLGpu<double>;
template class amigo::SparseCholeskyGpu<double>;
"""
    _compile(tmp_path, src)


def test_pybind_stub_compiles(tmp_path):
    pybind11 = pytest.importorskip("pybind11")
    import sysconfig

    src = "#include <pybind11/pybind11.h>\n#include <pybind11/stl.h>\n" + _block("adapter") + _block("pybind")
    _compile(tmp_path, src, ("-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"]))
