"""The C ABI library loads and exports every entry point include/pica_amd.h declares (no
compute call: this runs without a GPU), and the Python layer has no CPU fallback."""
import os
import re
import subprocess
import sys

import pica_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pica_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pk_[a-z_0-9]+)\s*\(", text)))


def test_header_and_python_symbol_lists_agree():
    assert sorted(P.ABI_SYMBOLS) == _declared()


def test_library_exports_every_declared_symbol():
    lib = P.lib()
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", P.LIBPATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (pk_[a-z_0-9]+)\b", out))
    assert set(_declared()) <= exported


def test_no_fallback_without_the_native_library(tmp_path):
    code = ("import pica_amd as P; P.LIBPATH = %r\n"
            "try:\n    P.lib()\nexcept RuntimeError as e:\n    print('raised', e)\n" % str(tmp_path / "missing.so"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT)
    assert "raised" in r.stdout


def test_cxx_surface_compiles():
    """include/pica_amd/pica.hpp (pica's names over the C ABI) compiles on its own, with every
    member and free function of the FP64 and FP32 forms instantiated"""
    src = """#include "pica_amd/pica.hpp"
using namespace pica_amd;
template class pica_amd::YeeGrid<double>;
template class pica_amd::YeeGrid<float>;
template class pica_amd::EnsembleSupercells<double>;
template class pica_amd::EnsembleSupercells<float>;
template <class R> static void all()
{
    void (*push)(EnsembleSupercells<R>&, YeeGrid<R>&, double, FormFactor, Depositor) = &pushParticles<R>;
    void (*one)(EnsembleSupercells<R>&, YeeGrid<R>&, double, FormFactor, Depositor, int) = &step<R>;
    void (*many)(std::vector<EnsembleSupercells<R>*>&, YeeGrid<R>&, double, FormFactor, Depositor) = &step<R>;
    Diagnostics (*diag)(EnsembleSupercells<R>&, YeeGrid<R>&, FormFactor) = &diagnostics<R>;
    void (YeeSolver::*ue)(YeeGrid<R>&, R) = &YeeSolver::updateE<R>;
    void (YeeSolver::*ub)(YeeGrid<R>&, R) = &YeeSolver::updateB<R>;
    (void)push; (void)one; (void)many; (void)diag; (void)ue; (void)ub;
}
int main() { all<double>(); all<float>(); return 0; }
"""
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-x", "c++", "-"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_pica_compat_surface_compiles():
    """include/pica_amd/pica_compat.hpp (pica's own namespace and signatures) compiles the pica
    caller tests/cxx/pica_loop.cpp with only its includes and particle loop switched"""
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-DUSE_PICA_AMD", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cxx", "pica_loop.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
