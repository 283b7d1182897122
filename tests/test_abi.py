"""The C-ABI libraries load on a CPU-only host and export every function the
public headers declare (no compute calls are made here).  The restated
generic-object structs match the reference's declarations: every ops slot's
name, return type and parameter types in order, every field of the object
structs, and -- compiled with the host C compiler -- every offset and size,
against the layout the reference's declared field types give on LP64 with
its default scalar types (sunrealtype double, sunindextype 64-bit)."""
import ctypes as C
import os
import re
import shutil
import subprocess

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
         "learned_predictor.h": {"SUNAMD_TrainingLogWrite", "SUNAMD_TrainingLogInfoRead",
                                 "SUNAMD_TrainingLogRead"},
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


REF_INC = "/root/reference/include/sundials"
OBJECTS = {"_generic_N_Vector": "sundials_nvector.h", "_generic_SUNMatrix": "sundials_matrix.h",
           "_generic_SUNLinearSolver": "sundials_linearsolver.h",
           "_generic_SUNNonlinearSolver": "sundials_nonlinearsolver.h"}


def _strip(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _slots(text, struct):
    """(return type, name, parameter types) of every function-pointer slot."""
    body = re.search(r"struct %s\s*\{(.*?)\};" % struct, _strip(text), re.S).group(1)
    out = []
    for decl in body.split(";"):
        d = " ".join(decl.split())
        if not d:
            continue
        m = re.match(r"^(.*?)\(\s*\*\s*(\w+)\s*\)\s*\((.*)\)$", d)
        assert m, d
        args = tuple(" ".join(a.replace("*", " * ").split()) for a in m.group(3).split(","))
        out.append((" ".join(m.group(1).replace("*", " * ").split()), m.group(2), args))
    return out


def _members(text, struct):
    """(type, name) of the members of a plain struct."""
    body = re.search(r"struct %s\s*\{(.*?)\};" % struct, _strip(text), re.S).group(1)
    out = []
    for decl in body.split(";"):
        d = " ".join(decl.replace("*", " * ").split())
        if d:
            t, n = d.rsplit(" ", 1)
            out.append((t, n))
    return out


def _ref(header):
    path = os.path.join(REF_INC, header)
    if not os.path.exists(path):
        pytest.skip("reference headers not present on this host")
    return open(path).read()


@pytest.mark.parametrize("struct", sorted(STRUCTS))
def test_abi_ops_signatures_match_reference(struct):
    ours = _slots(open(os.path.join(INC, "sundials_abi.h")).read(), struct)
    ref = _slots(_ref(STRUCTS[struct][0]), struct)
    assert ours == ref


@pytest.mark.parametrize("struct", sorted(OBJECTS))
def test_abi_object_members_match_reference(struct):
    ours = _members(open(os.path.join(INC, "sundials_abi.h")).read(), struct)
    assert ours == _members(_ref(OBJECTS[struct]), struct)


def _ref_default_index_bits():
    path = "/root/reference/cmake/SundialsBuildOptionsPre.cmake"
    if not os.path.exists(path):
        pytest.skip("reference build files not present on this host")
    # the non-xSDK branch: if(USE_XSDK_DEFAULTS) ... else() <default> endif()
    m = re.search(r'else\(\)\s*sundials_option\(\s*SUNDIALS_INDEX_SIZE\s+STRING\s+"[^"]*"\s+'
                  r'"(\d+)"', open(path).read())
    assert m
    return int(m.group(1))


def test_abi_compiled_layout():
    """Offsets and sizes of this library's structs as the C compiler lays them
    out equal those the reference's member types give: every ops slot and
    object member a pointer (8 bytes on LP64, in declaration order), and the
    scalar types of the reference's default build."""
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    assert _ref_default_index_bits() == 64
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "sunamd/sundials_abi.h"',
             "int main(void) {",
             'printf("scalar %zu %zu %zu\\n", sizeof(sunrealtype), sizeof(sunindextype), '
             "sizeof(sunbooleantype));"]
    expect = {"scalar": [8, _ref_default_index_bits() // 8, 4]}
    for struct, (header, n) in STRUCTS.items():
        names = [x[1] for x in _slots(_ref(header), struct)]
        offs = " ".join("%zu" for _ in names)
        args = ", ".join("offsetof(struct %s, %s)" % (struct, f) for f in names)
        lines.append('printf("%s %%zu %s\\n", sizeof(struct %s), %s);' % (struct, offs, struct,
                                                                         args))
        expect[struct] = [8 * n] + [8 * i for i in range(n)]
    for struct, header in OBJECTS.items():
        mem = _members(_ref(header), struct)
        assert all("*" in t or t in ("N_Vector_Ops", "SUNMatrix_Ops", "SUNLinearSolver_Ops",
                                     "SUNNonlinearSolver_Ops", "SUNContext") for t, _ in mem)
        offs = " ".join("%zu" for _ in mem)
        args = ", ".join("offsetof(struct %s, %s)" % (struct, f) for _, f in mem)
        lines.append('printf("%s %%zu %s\\n", sizeof(struct %s), %s);' % (struct, offs, struct,
                                                                         args))
        expect[struct] = [8 * len(mem)] + [8 * i for i in range(len(mem))]
    lines.append("return 0; }")
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run([cc, "-std=c11", "-I" + os.path.join(ROOT, "include"), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = {ln.split()[0]: [int(x) for x in ln.split()[1:]] for ln in out.splitlines()}
    assert got == expect


def test_massaction_upload_rejects_oversized_network():
    """SUNAMD_MassActionToDevice refuses, before any device call, a network
    whose per-cell Jacobian work area does not fit the workgroup's 64 KiB of
    LDS (m = 64 with 2000 reactions) and m outside 1..64."""
    core, hip = _lib.core(), _lib.hip()
    for m, R in ((64, 2000), (65, 10)):
        p = core.SUNAMD_MassActionCreate(m, R, 1, 5, 0.2)
        assert p
        assert not hip.SUNAMD_MassActionToDevice(p, None)
        core.SUNAMD_MassActionFree(p)
