"""The C-ABI library loads and exports every symbol include/mbx.h declares; the ctypes
mirror of mbx_params matches the compiled layout (no GPU calls)."""
import ctypes
import os
import re

import pytest

from mandelbulber2_amd import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(mbx_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_declared_symbols():
    lib = runtime.load()
    names = declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(runtime.EXPORTS)


def test_params_layout_matches_library(oracle_mod):
    lib = runtime.load()
    assert lib.mbx_version() == 9
    assert lib.mbx_sizeof_params() == ctypes.sizeof(abi.MbxParams)
    assert oracle_mod.lib().mbxo_sizeof_params() == ctypes.sizeof(abi.MbxParams)


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "mandelbulber2_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in text.replace("oracle.", "").lower() or fn == "__init__.py", fn


def test_formula_structs_match_c_layout(tmp_path):
    """mbx_formula_info, mbx_extended_aux, mbx_outputs and mbx_params: ctypes sizes and field
    offsets equal the C compiler's (a probe built from include/mbx.h with gcc)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    fields = {"mbx_formula_info": abi.MbxFormulaInfo, "mbx_extended_aux": abi.MbxExtendedAux,
              "mbx_outputs": abi.MbxOutputs, "mbx_params": abi.MbxParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mbx.h"', "int main(void) {"]
    for cname, cls in fields.items():
        lines.append('printf("%%s sizeof %%zu\\n", "%s", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%%s %%s %%zu\\n", "%s", "%s", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        parts = ln.split()
        if len(parts) == 3:
            got[(parts[0], parts[1])] = int(parts[2])
    for cname, cls in fields.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)
