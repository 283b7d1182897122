"""The C-ABI libraries load on a CPU-only host and export every function the
public headers declare (no compute calls are made here)."""
import ctypes as C
import os
import re

import pytest

from sundials_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include", "sunamd")

# header -> library that must export its functions
HEADERS = {
    "sundials_core.h": "core",
    "cvode_sunamd.h": "core",
    "massaction.h": "core",
    "nvector_mi355.h": "hip",
    "sunmatrix_mi355block.h": "hip",
    "massaction_mi355.h": "hip",
    "batched_bdf.h": "hip",
    "sundials_iterative.h": "core",
    "sunlinsol_spgmr.h": "core",
    "tridiag_mi355.h": "hip",
    "learned_predictor.h": "hip",
    "sunamd_profiler.h": "hip",
    "arkode_sunamd.h": "core",
    "imex_adr.h": "core",
    "kinsol_sunamd.h": "core",
    "kin_fp.h": "hip",
    "kin_newton.h": "hip",
}

# declared in a header whose functions live in the other library
SPLIT = {"batched_bdf.h": {"SUNAMD_BatchCreate", "SUNAMD_BatchFree"},
         "imex_adr.h": {"SUNAMD_AdrSolve_Mi355"}}

_DECL = re.compile(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
    text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
    text = re.sub(r"struct[^{;]*\{.*?\};", "", text, flags=re.S)
    names = set()
    for stmt in text.split(";"):
        s = " ".join(stmt.split())
        m = re.match(r"^(?:extern \"C\" \{ )?[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", s)
        if m:
            names.add(m.group(1))
    return sorted(names)


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_header_symbols_exported(header):
    lib = _lib.core() if HEADERS[header] == "core" else _lib.hip()
    other = _lib.hip() if HEADERS[header] == "core" else _lib.core()
    names = declared(header)
    assert names, header
    split = SPLIT.get(header, set())
    missing = [n for n in names if not hasattr(other if n in split else lib, n)]
    assert not missing, missing


def _fields(text, struct):
    body = re.search(r"struct %s\s*\{(.*?)\};" % struct, text, re.S).group(1)
    return re.findall(r"\(\*(\w+)\)", body)


STRUCTS = {
    "_generic_N_Vector_Ops": ("sundials_nvector.h", 56),
    "_generic_SUNMatrix_Ops": ("sundials_matrix.h", 11),
    "_generic_SUNLinearSolver_Ops": ("sundials_linearsolver.h", 16),
    "_generic_SUNNonlinearSolver_Ops": ("sundials_nonlinearsolver.h", 14),
}


@pytest.mark.parametrize("struct", sorted(STRUCTS))
def test_abi_ops_tables(struct):
    """Ops tables have the reference's slot count and order
    (sundials_nvector.h:101-195, sundials_matrix.h:89-105,
    sundials_linearsolver.h:117-140, sundials_nonlinearsolver.h:103-121)."""
    ours = _fields(open(os.path.join(INC, "sundials_abi.h")).read(), struct)
    header, n = STRUCTS[struct]
    assert len(ours) == n
    ref = os.path.join("/root/reference/include/sundials", header)
    if not os.path.exists(ref):
        pytest.skip("reference headers not present on this host")
    assert ours == _fields(open(ref).read(), struct)
