"""CPU checks of the thrust_func.h drop-ins: every overload the header declares
is exported by libsegclip with C++ linkage, a device call that fails throws
(as Thrust does) instead of returning silently, and the host stable sort keeps
ties in input order (tests/facade/thrust_check.cpp; no GPU call)."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "segment-tree-based-clipping_amd")
SRC = os.path.join(ROOT, "tests", "facade", "thrust_check.cpp")
BIN = os.path.join(ROOT, "tests", "facade", "thrust_check")


def _declared():
    src = open(os.path.join(ROOT, "include", "thrust_func.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.findall(r"void\s+(\w+)\(([^)]*)\)", src)


def test_overloads_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", os.path.join(PKG, "libsegclip.so")], capture_output=True,
                         text=True, check=True).stdout
    decl = _declared()
    assert len(decl) == 6
    for name, args in decl:
        types = ", ".join(re.sub(r"\s*\b\w+$", "", a.strip()).replace(" *", "*") for a in args.split(","))
        assert "%s(%s)" % (name, types) in out, (name, types)


def test_errors_throw_and_host_sort_is_stable():
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < os.path.getmtime(SRC):
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-o", BIN, SRC, "-L" + PKG, "-lsegclip",
                               "-Wl,-rpath," + PKG])
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
