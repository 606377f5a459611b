"""The row-echelon-form tool (lela_amd/util/row_echelon_form.cpp): argument
handling and its messages, which LELA's util/row-echelon-form.C and
util/support.h define.  These paths end before any GPU work, so they run on
the CPU; the computing paths are in test_cli_gpu.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "lela_amd", "bin", "row-echelon-form")
SMS = os.path.join(ROOT, "tests", "golden", "data", "sms.matrix")

pytestmark = pytest.mark.skipif(not os.path.exists(TOOL), reason="row-echelon-form not built")


def run(*args):
    r = subprocess.run([TOOL, *args], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("args,message", [
    (["-m", "bogus", SMS, "out.m"], "Invalid method (use 'standard', 'afast', or 'f4')"),
    (["-k", "gf2", "-m", "m4ri", SMS, "out.m"], "Invalid method (use 'standard', 'afast', or 'f4')"),
    (["-k", "zz", SMS, "out.m"], "Invalid ring (use 'gf2' or 'modular')"),
    (["-t", "hybrid", SMS, "out.m"], "Invalid matrix-type (use 'dense' or 'sparse')"),
    (["-k", "gf2", "-t", "list", SMS, "out.m"], "Invalid matrix-type (use 'dense', 'sparse', or 'hybrid')"),
    ([SMS, "out.txt"], "Could not guess output file-format"),
    (["-o", "png", SMS, "out.txt"], "Invalid output-file-format"),
    (["-i", "xml", SMS, "out.m"], "Invalid input-file-format"),
    (["-p", "18446744073709551557", SMS, "out.m"], "no GPU path"),
])
def test_argument_errors(tmp_path, args, message):
    args = [str(tmp_path / a) if a.startswith("out") else a for a in args]
    rc, out = run(*args)
    assert rc != 0 and message in out


def test_input_errors(tmp_path):
    rc, out = run(str(tmp_path / "missing.matrix"), str(tmp_path / "o.m"))
    assert rc != 0 and "Could not open input-file" in out
    f = tmp_path / "x.matrix"
    f.write_text("%%MatrixMarket matrix coordinate integer general\n2 2 1\n1 1 1\n")
    rc, out = run(str(f), str(tmp_path / "o.m"))
    assert rc != 0 and "Unable to determine format of input-file" in out


def test_help():
    rc, out = run("-h")
    assert rc != 0 and "Usage:" in out and "-m M" in out
    rc, out = run()
    assert rc != 0 and "Usage:" in out
