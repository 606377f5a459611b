"""The LELA-side binding (include/lela_amd/hip-module.h) type-checks inside a
LELA translation unit.

A translation unit that instantiates every hook -- gemm and trsm for
Modular<uint16/uint32/double>, GF(2) gemm and trsm, EchelonForm (reduced, non-reduced,
Faugere-Lachartre method, rank) for the three Z/p storages and GF(2),
FaugereLachartre on sparse and dense rows, and Elimination (echelonize with
and without L, with the sparse pivot strategy, pluq, the reduced form, move_L) on dense and sparse rows, GF(2) EchelonForm on dense, sparse and hybrid rows --
goes through
`g++ -std=gnu++98 -fsyntax-only` against LELA's own headers.  This is a
syntax-only check: no object code is produced, nothing is linked or run, and
no result of LELA is taken from it.  LELA's headers include the
configure-generated lela/lela-config.h, which the source tree does not ship;
the test writes the handful of type macros it needs (configure.in:30-130 of
the LELA tree) into its own temporary directory.  Skipped where the LELA
source tree or g++ is absent (the GPU hosts have neither)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LELA = "/root/reference"
GMP = "/opt/conda/include"

CONFIG = """#define __LELA_INT8 char
#define __LELA_INT16 short
#define __LELA_INT32 int
#define __LELA_INT64 long
#define __LELA_UINT128 __uint128_t
#define __LELA_HAVE_LITTLE_ENDIAN 1
#define __LELA_SIZEOF_INT 4
#define __LELA_SIZEOF_LONG 8
"""

TU = r"""
#include "lela_amd/hip-module.h"
#include "lela/matrix/sparse.h"
#include "lela/vector/hybrid.h"
using namespace LELA;

template <class E>
void zp_hooks (const Modular<E> &F)
{
	Context<Modular<E>, HIPModule<Modular<E> > > ctx (F);
	DenseMatrix<E> A (8, 8), B (8, 8), C (8, 8);
	BLAS3::gemm (ctx, F.one (), A, B, F.zero (), C);
	typename DenseMatrix<E>::SubmatrixType As (A, 1, 1, 4, 4), Bs (B, 0, 0, 4, 4), Cs (C, 2, 2, 4, 4);
	BLAS3::gemm (ctx, F.one (), As, Bs, F.one (), Cs);
	BLAS3::trsm (ctx, F.one (), A, B, LowerTriangular, true);
	BLAS3::trsm (ctx, F.one (), A, B, UpperTriangular, false);
	typedef EchelonForm<Modular<E>, HIPModule<Modular<E> > > EF_t;
	EF_t EF (ctx);
	EF.echelonize (A, true);
	EF.echelonize (A, false);
	EF.echelonize (A, true, EF_t::METHOD_FAUGERE_LACHARTRE);
	EF.echelonize (A, false, EF_t::METHOD_STANDARD_GJ);
	size_t r = EF.rank (A);
	SparseMatrix<E> X (8, 8), R (8, 8);
	EF.echelonize (X, true);
	EF.echelonize (X, false, EF_t::METHOD_FAUGERE_LACHARTRE);
	r += EF.rank (X);
	FaugereLachartre<Modular<E>, HIPModule<Modular<E> > > FL (ctx);
	size_t rank;
	E det;
	FL.echelonize (R, X, rank, det, true);
	FL.echelonize (R, X, rank, det, false);
	DenseMatrix<E> Xd (8, 8), Rd (8, 8);
	FL.echelonize (Rd, Xd, rank, det, true);
	FL.det_wanted (false);
	FL.echelonize (Rd, Xd, rank, det);
	typedef Elimination<Modular<E>, HIPModule<Modular<E> > > Elim_t;
	Elim_t elim (ctx);
	typename Elim_t::Permutation P, Q;
	elim.echelonize (A, P, rank, det, true);
	elim.echelonize (A, P, rank, det, false);
	elim.pluq (A, P, Q, rank, det);
	elim.echelonize (X, P, rank, det, false);
	elim.echelonize (X, P, rank, det, true);
	elim.echelonize (X, P, rank, det, SparsePartialPivotStrategy<Modular<E>, HIPModule<Modular<E> > > (ctx), true);
	elim.pluq (X, P, Q, rank, det);
	EF.echelonize (X, false);
	DenseMatrix<E> L (8, 8);
	elim.echelonize_reduced (A, L, P, rank, det, true);
	elim.move_L (L, A);
	(void) r;
}

void gf2_hooks ()
{
	GF2 F;
	Context<GF2, HIPModule<GF2> > ctx (F);
	DenseMatrix<bool> A (64, 64), B (64, 64), C (64, 64);
	BLAS3::gemm (ctx, true, A, B, false, C);
	BLAS3::trsm (ctx, true, A, B, LowerTriangular, true);
	BLAS3::trsm (ctx, true, A, B, UpperTriangular, false);
	typedef EchelonForm<GF2, HIPModule<GF2> > EF_t;
	EF_t EF (ctx);
	EF.echelonize (A, true);
	EF.echelonize (A, false);
	EF.echelonize (A, true, EF_t::METHOD_FAUGERE_LACHARTRE);
	size_t r = EF.rank (A);
	FaugereLachartre<GF2, HIPModule<GF2> > FL (ctx);
	size_t rank;
	bool det;
	FL.echelonize (C, A, rank, det, true);
	typedef SparseMatrix<bool, HybridVector<DefaultEndianness<uint64>, uint16, uint64> > Hyb;
	Hyb X (64, 64), R (64, 64);
	FL.echelonize (R, X, rank, det, true);
	FL.echelonize (R, X, rank, det, false);
	SparseMatrix<bool> S (64, 64), T (64, 64);
	FL.echelonize (T, S, rank, det, true);
	EF.echelonize (X, true, EF_t::METHOD_FAUGERE_LACHARTRE);
	EF.echelonize (X, true);
	EF.echelonize (X, false);
	EF.echelonize (S, true);
	EF.echelonize (S, false);
	(void) r;
}

int main ()
{
	zp_hooks (Modular<uint16> (65521));
	zp_hooks (Modular<uint32> (65521));
	zp_hooks (Modular<double> (67108859));
	gf2_hooks ();
	return 0;
}
"""


@pytest.mark.skipif(not os.path.isdir(os.path.join(LELA, "lela")), reason="LELA source tree absent")
@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
@pytest.mark.skipif(not os.path.exists(os.path.join(GMP, "gmpxx.h")), reason="gmpxx.h absent")
def test_hip_module_header_type_checks_in_a_lela_unit(tmp_path):
    (tmp_path / "lela").mkdir()
    (tmp_path / "lela" / "lela-config.h").write_text(CONFIG)
    tu = tmp_path / "hooks.C"
    tu.write_text(TU)
    res = subprocess.run(["g++", "-std=gnu++98", "-fsyntax-only", "-Wall", "-Wextra",
                          "-I", str(tmp_path), "-I", LELA, "-I", GMP, "-I", os.path.join(ROOT, "include"),
                          str(tu)], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr[-4000:]
    own = [l for l in res.stderr.splitlines() if re.search(r"hip-module\.h:\d+:\d+: (warning|error)", l)]
    assert not own, "\n".join(own)
