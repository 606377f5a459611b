"""CPU checks of the C-ABI library: it loads, exports every function that
include/*.h declares, and the Python binding declares the same set.  No
compute call is made (there may be no GPU here)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "lela_amd", "liblela_amd.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mlz_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert "mlz_zp_gemm" in names and "mlz_zp_trsm" in names and "mlz_ctx_create" in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="liblela_amd.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from lela_amd import _lib
    assert declared_functions() == set(_lib.SIGNATURES)


def test_no_gpu_context_fails_loudly():
    """Without a usable GPU the context constructor raises; nothing falls back."""
    import lela_amd as la
    if not os.path.exists(LIB):
        with pytest.raises(RuntimeError):
            la.Context(la.Modular(101))
        return
    try:
        ctx = la.Context(la.Modular(101))
    except la.LELAError:
        return                      # no GPU: a clear error, no fallback
    ctx.close()
    pytest.skip("GPU present")


def test_task_struct_layout_matches_header(tmp_path):
    """struct mlz_task as gcc lays it out == the ctypes Task of the binding."""
    import shutil
    import subprocess
    from lela_amd._lib import Task
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    fields = [f for f, _ in Task._fields_]
    src = tmp_path / "layout.c"
    body = "\n".join(f'printf("{f} %zu\\n", offsetof(mlz_task, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "mlz.h"\nint main(void){\n'
                   + body + '\nprintf("size %zu\\n", sizeof(mlz_task));\nreturn 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    for f in fields:
        assert int(out[f]) == getattr(Task, f).offset, f
    assert int(out["size"]) == ctypes.sizeof(Task)
