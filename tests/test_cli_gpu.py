"""The row-echelon-form tool on the GPU: matrix files in, echelon forms out,
compared with the oracle's forms of the same matrices (the dispatch of
EchelonForm::echelonize per matrix type and method, lela/solutions/echelon-form.h:61-154)."""
import os
import subprocess

import numpy as np
import pytest

import lela_amd as la
import oracle_lib as orc
from lela_amd import matrix_io as mio
from lela_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "lela_amd", "bin", "row-echelon-form")
SMS = os.path.join(ROOT, "tests", "golden", "data", "sms.matrix")


def tool(*args):
    print("row-echelon-form", *args, flush=True)
    r = subprocess.run([TOOL, *map(str, args)], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout + r.stderr


def read_back(path, F, fmt, rows=0, cols=0):
    return mio.MatrixReader(F).read_rows(path, fmt, rows, cols)


def test_sms_fixture_reduced_and_not(tmp_path):
    p = 65521
    F = la.Modular(p, "uint16")
    A = F.to_canonical(mio.MatrixReader(F).read(SMS, mio.FORMAT_DUMAS))
    rc0, R, rank, piv, _ = orc.zp_echelonize(p, A, reduced=True)
    _, U, urank, _, _ = orc.zp_elim(p, A, compute_L=False)
    _, Us, srank, _, _, _ = orc.zp_elim_sparse(p, A, 0)
    for method in ("standard", "afast"):
        out = tmp_path / f"r_{method}.txt"
        rc, msg = tool("-r", "-m", method, "-i", "dumas", "-o", "dumas", SMS, out)
        assert rc == 0, msg
        S = read_back(out, F, mio.FORMAT_DUMAS)
        assert S.shape == (11, 11) and np.array_equal(S.dense(), R.astype(np.uint64))
        out = tmp_path / f"u_{method}.txt"
        rc, msg = tool("-m", method, "-o", "turner", SMS, out)
        assert rc == 0, msg
        assert np.array_equal(read_back(out, F, mio.FORMAT_TURNER, 11, 11).dense(), U.astype(np.uint64))
    out = tmp_path / "s.txt"
    rc, msg = tool("-m", "standard", "-t", "sparse", "-o", "pretty", SMS, out)
    assert rc == 0, msg
    assert np.array_equal(read_back(out, F, mio.FORMAT_PRETTY).dense(), Us.astype(np.uint64))
    # sparse rows with afast: LELA's EchelonForm rejects the method
    rc, msg = tool("-m", "afast", "-t", "sparse", "-o", "pretty", SMS, out)
    assert rc != 0 and "Invalid method for choice of matrix" in msg


def test_f4_on_f4_shaped_file(tmp_path):
    p = 65521
    F = la.Modular(p, "uint32")
    csr, shape = synth.fl_matrix(11, p, na=150, nc=60, nb=120, band=40, nnz_a=8, density=0.05)
    X = synth.csr_to_dense(csr, shape).astype(np.uint32)
    src = tmp_path / "x.sms"
    mio.MatrixWriter(F).write(src, X, mio.FORMAT_DUMAS)
    rc0, R, rank, _, _ = orc.fl_echelonize(p, X)
    assert rc0 == 0
    for t in ("dense", "sparse"):
        out = tmp_path / f"r_{t}.m"
        rc, msg = tool("-r", "-t", t, src, out)       # defaults: f4, output format from ".m"
        assert rc == 0, msg
        assert np.array_equal(read_back(out, F, mio.FORMAT_MATLAB).dense(), R.astype(np.uint64))
    # the dense f4 method is always reduced (echelon-form.h:139-144)
    out = tmp_path / "d.m"
    rc, msg = tool(src, out)
    assert rc == 0, msg
    assert np.array_equal(read_back(out, F, mio.FORMAT_MATLAB).dense(), R.astype(np.uint64))
    # a matrix whose heads decrease is not F4-shaped
    bad = tmp_path / "bad.sms"
    mio.MatrixWriter(F).write(bad, np.array([[0, 1], [1, 0]], dtype=np.uint32), mio.FORMAT_DUMAS)
    rc, msg = tool("-r", bad, tmp_path / "o.m")
    assert rc != 0 and "Faugere-Lachartre" in msg


def test_floating_point_ring_writes_symmetric(tmp_path):
    p = 101
    Fd = la.Modular(p, "double")
    A = synth.zp_matrix(3, 30, 40, p).astype(np.uint32)
    src = tmp_path / "a.sms"
    mio.MatrixWriter(la.Modular(p, "uint32")).write(src, A, mio.FORMAT_DUMAS)
    out = tmp_path / "o.txt"
    rc, msg = tool("-f", "-p", p, "-r", "-m", "standard", "-o", "pretty", src, out)
    assert rc == 0, msg
    text = out.read_text()
    assert "-" in text                                  # symmetric representatives
    _, R, _, _, _ = orc.zp_echelonize(p, A, reduced=True)
    assert np.array_equal(read_back(out, Fd, mio.FORMAT_PRETTY).dense(), R.astype(np.uint64))


def test_gf2(tmp_path):
    rng = np.random.default_rng(5)
    m, n = 90, 150
    B = (rng.random((m, n)) < 0.05).astype(np.uint32)
    B[m - 1, n - 1] = 1          # Turner files carry no shape: the last row and column are read from entries
    src = tmp_path / "g.txt"
    mio.MatrixWriter(la.Modular(2, "uint16")).write(src, B, mio.FORMAT_TURNER)
    W = np.zeros((m, (n + 63) // 64), dtype=np.uint64)
    for j in range(n):
        W[:, j // 64] |= B[:, j].astype(np.uint64) << np.uint64(j % 64)
    Rg, rank, _ = orc.gf2_echelonize(m, n, W)
    out = tmp_path / "r.txt"
    rc, msg = tool("-k", "gf2", "-r", "-m", "standard", "-i", "turner", "-o", "dumas", src, out)
    assert rc == 0, msg
    got = read_back(out, la.GF2(), mio.FORMAT_DUMAS)
    bits = np.unpackbits(Rg.view(np.uint8), axis=1, bitorder="little")[:, :n]
    assert np.array_equal(got.dense(), bits.astype(np.uint64))
    # non-reduced forms: dense rows take the dense rule, sparse and hybrid rows the sparse rule
    _, Ud, _, _, _ = orc.zp_elim(2, B, compute_L=False)
    for t, want in (("dense", Ud), ("sparse", orc.zp_elim_sparse(2, B, 0, False)[1]),
                    ("hybrid", orc.zp_elim_sparse(2, B, 0, True)[1])):
        out = tmp_path / f"u_{t}.txt"
        rc, msg = tool("-k", "gf2", "-t", t, "-m", "standard", "-i", "turner", "-o", "dumas", src, out)
        assert rc == 0, msg
        assert np.array_equal(read_back(out, la.GF2(), mio.FORMAT_DUMAS).dense(), want.astype(np.uint64)), t
