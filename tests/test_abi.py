"""The C-ABI library loads and exports every function include/*.h declares
(CPU only: no compute call is made).  Builds the library if it is absent."""

import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_DECL = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?[\s\*])([A-Za-z_]\w*)\s*\(", re.M)


def declared_functions(header):
    """Function names declared at file scope in a C header (comments stripped)."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
    names = []
    for m in _DECL.finditer(text):
        name = m.group(1)
        if name not in ("if", "while", "for", "return", "sizeof"):
            names.append(name)
    return names


def _headers():
    return sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def _libraries():
    from jaxmaterials_amd import build

    libs = {}
    for path in build.all_library_paths():
        if not os.path.exists(path):
            build.build()
            break
    for path in build.all_library_paths():
        libs[os.path.basename(path)] = ctypes.CDLL(path)
    return libs


def test_header_parser():
    names = declared_functions(os.path.join(ROOT, "include", "lippmann_schwinger.h"))
    assert "lippmann_schwinger_solve" in names
    assert "ls_solver_create" in names
    assert len(names) == len(set(names))


def test_every_declared_symbol_is_exported():
    libs = _libraries()
    assert libs, "no library built"
    missing = []
    for h in _headers():
        for name in declared_functions(h):
            if not any(hasattr(lib, name) for lib in libs.values()):
                missing.append((os.path.basename(h), name))
    assert not missing, f"declared but not exported: {missing}"


def test_loader_signatures_cover_header():
    """The ctypes table of the Python layer binds exactly what the headers declare."""
    from jaxmaterials_amd import _lib

    declared = set()
    for h in _headers():
        declared.update(declared_functions(h))
    bound = set(_lib.SIGNATURES)
    assert declared == bound, (declared - bound, bound - declared)


def test_drop_in_library_name_and_symbol():
    """The reference loads `liblippmannschwinger.so` and binds
    `lippmann_schwinger_solve` (src/jaxmaterials/solver/lippmann_schwinger.py:286-291)."""
    from jaxmaterials_amd import build

    assert os.path.basename(build.LIB_PATH) == "liblippmannschwinger.so"
    lib = _libraries()["liblippmannschwinger.so"]
    assert hasattr(lib, "lippmann_schwinger_solve")


def test_product_path_has_no_oracle_import():
    """The shipped package never imports the CPU oracle."""
    pkg = os.path.join(ROOT, "jaxmaterials_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", src, re.M), f
