"""The C-ABI boundary (CPU only, no compute calls): the library loads, exports
every entry point declared in include/sedona_gpu.h, and the ctypes mirrors
have the C compiler's layout."""
import ctypes
import os
import re
import subprocess

import pytest

from pubsed_amd import abi, load_library, LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sedona_gpu.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sgpu_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(abi.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "build the library first (python -m pubsed_amd.build)"
    lib = load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    for name in declared_functions():
        assert name in exported, name
        assert getattr(lib, name) is not None


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sizes.c"
    names = sorted(abi.PY_STRUCTS)
    body = "\n".join('printf("%s %%zu\\n", sizeof(%s));' % (n, n) for n in names)
    src.write_text('#include <stdio.h>\n#include "%s"\nint main(void){%s return 0;}\n' % (HEADER, body))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for n in names:
        assert int(got[n]) == ctypes.sizeof(abi.PY_STRUCTS[n]), n


def test_field_offsets_match_c(tmp_path):
    """Offsets of every field of the larger descriptors."""
    checks = []
    for cname, cls in abi.PY_STRUCTS.items():
        for f, _ in cls._fields_:
            cf = "del" if f == "del_" else f
            checks.append((cname, cf, getattr(cls, f).offset))
    src = tmp_path / "offs.c"
    body = "\n".join('printf("%%zu\\n", offsetof(%s, %s));' % (c, f) for c, f, _ in checks)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "%s"\nint main(void){%s return 0;}\n'
                   % (HEADER, body))
    exe = tmp_path / "offs"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    for (c, f, o), v in zip(checks, vals):
        assert o == v, (c, f, o, v)


def test_oracle_is_not_imported_by_the_package():
    """The product path never imports, loads or links the CPU oracle."""
    pkg = os.path.join(ROOT, "pubsed_amd")
    pat = re.compile(r"(^\s*(import|from)\s+oracle\b)|liboracle|oracle_py", re.M)
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                assert not pat.search(open(os.path.join(dp, f)).read()), f
