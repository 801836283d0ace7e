"""The C++ a reference maintainer is told to add (the
SweeperHIP plug-in class, the device-loop solve(), the PETSc-GMRES operator,
the defect call, the priority keys) compiles against the reference's own
headers (src/SweeperAbstract.hh:47-70, src/KrylovSolver.hh:50-135,
src/Global.hh:109-133) and this library's header.  Syntax and types only
(g++ -fsyntax-only); skipped where the reference tree is absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"

pytestmark = pytest.mark.skipif(
    not os.path.isdir(REF_SRC) or shutil.which("g++") is None,
    reason="needs the reference's src/ headers and g++")


def cpp_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, re.S)


def block(first_line_prefix):
    for b in cpp_blocks():
        if b.lstrip().startswith(first_line_prefix):
            return b
    raise AssertionError(f"INTEGRATION.md has no cpp block starting "
                         f"{first_line_prefix!r}")


def syntax_check(tmp_path, name, source):
    path = tmp_path / name
    path.write_text(source)
    res = subprocess.run(
        ["g++", "-std=c++11", "-fsyntax-only", "-Wall", f"-I{tmp_path}",
         f"-I{REF_SRC}", f"-I{os.path.join(ROOT, 'include')}", str(path)],
        capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


PRELUDE = """#include "SweeperHIP.hh"
#include "Global.hh"
#include "PsiData.hh"
#include "KrylovSolver.hh"
#include "Problem.hh"
#include "Assert.hh"
#include <vector>
static void check(int rc) { Insist(rc == T2_OK, t2_last_error()); }
"""


@pytest.fixture
def header(tmp_path):
    (tmp_path / "SweeperHIP.hh").write_text(block("// SweeperHIP.hh"))
    return tmp_path


def test_plugin_class_compiles(header):
    """SweeperHIP.hh and SweeperHIP.cc as written (reference host loops)."""
    syntax_check(header, "SweeperHIP.cc", block("// SweeperHIP.cc"))


def test_device_loop_solve_compiles(header):
    """The solve() that runs the outer iteration on the device."""
    syntax_check(header, "solve_device.cc",
                 PRELUDE + block("void SweeperHIP::solve()"))


def test_petsc_operator_and_defect_and_priorities_compile(header):
    """The lhsOperator replacement and its call sequence in
    SourceIteration::krylov, the defect apply and the priority keys, each
    placed in a function that supplies the names the text assumes."""
    krylov = block("// replaces lhsOperator")
    head, _, body = krylov.partition("// in SourceIteration::krylov")
    body = body.split("\n", 1)[1]          # the rest of the comment line
    defect = block("std::vector<double> d(c_psi.size());")
    prio = block("check(t2_set_priority(")
    src = (PRELUDE + head +
           "void krylov_snippet(t2_sweeper *engine, UINT vecSize,\n"
           "                    const PsiData &source, PsiData &psi)\n{\n"
           + body + "}\n"
           "void defect_snippet(t2_sweeper *c_engine, PsiData &c_psi,\n"
           "                    std::vector<double> &phi)\n{\n"
           + defect + "}\n"
           "void priority_snippet(t2_sweeper *c_engine)\n{\n"
           + prio + "}\n")
    syntax_check(header, "snippets.cc", src)
