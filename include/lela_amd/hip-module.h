/* include/lela_amd/hip-module.h -- LELA-side binding of the MI355X backend.
 *
 * A module for LELA's Context<Ring, Modules> (lela/blas/context.h:32-85):
 *
 *     Context<Modular<uint16>, HIPModule<Modular<uint16> > > ctx (F);
 *     BLAS3::gemm (ctx, a, A, B, b, C);          // runs mlz_zp_gemm
 *
 * It follows the module protocol of lela/blas/README:41-91: a struct built
 * from the ring with a nested Tag naming its parent, and _gemm/_trsm
 * specialisations for that Tag that dispatch on Matrix::StorageType and
 * forward anything they do not handle to Tag::Parent (the precedents are
 * lela/blas/level3-modular.h:29-68 and lela/blas/level3-m4ri.h:89-128).
 * EchelonForm and FaugereLachartre are specialised as whole classes, like
 * the GF2 + M4RI specialisation in lela/solutions/echelon-form-gf2.h:37-182.
 *
 * What runs where:
 *   BLAS3::gemm, BLAS3::trsm on three dense Z/p matrices   -> GPU
 *   BLAS3::gemm, BLAS3::trsm on little-endian dense GF(2) matrices -> GPU
 *   EchelonForm::echelonize (A, reduced) of a dense matrix   -> GPU (mlz_zp_rref / mlz_gf2_rref for
 *       the reduced form, mlz_zp_echelonize for the non-reduced one over Z/p)
 *   EchelonForm::echelonize (A, *, METHOD_FAUGERE_LACHARTRE) -> GPU FL (dense A, as LELA's dense
 *       overload, always the reduced form: lela/solutions/echelon-form.h:133-139)
 *   FaugereLachartre::echelonize (R, X, rank, det, true)     -> GPU FL (Z/p: any row type; GF(2):
 *       hybrid rows of little-endian words and dense 0-1 matrices)
 *   Elimination::echelonize (A, P, rank, det, compute_L) and
 *   Elimination::pluq (A, P, Q, rank, det) over Z/p with the
 *       default pivot strategy: dense rows                  -> GPU (mlz_zp_echelonize)
 *                               sparse rows                 -> GPU (mlz_zp_echelonize_sparse)
 *   EchelonForm::echelonize (A, reduced) of sparse rows      -> GPU (the reduced form, or Elimination's
 *       non-reduced form with the sparse pivot rule; GF(2): sparse and hybrid 0-1 rows over Z/2)
 *   other explicit pivot strategies and every other storage -> LELA's own CPU code
 *
 * Determinant of FaugereLachartre: LELA multiplies det by the heads of the
 * scanned rows of X and by the heads of its (non-reduced) echelon form of
 * D - C A^-1 B (lela/algorithms/faugere-lachartre.tcc:119, :831-833); the GPU
 * call returns that product.  A caller that does not use det can say so
 * (det_wanted (false)): the GPU then skips D's non-reduced form and det is 1.
 *
 * The header is compiled inside a LELA build (it includes LELA's headers,
 * C++98) and links against lela_amd/liblela_amd.so.  Return codes of the C
 * ABI are turned back into LELA's exceptions.
 */
#ifndef LELA_AMD_HIP_MODULE_H
#define LELA_AMD_HIP_MODULE_H

#include <algorithm>
#include <iostream>      // LELA's bit-iterator.h uses the stream operators of bool without including it
#include <map>
#include <vector>

#include "lela/algorithms/elimination.h"
#include "lela/algorithms/faugere-lachartre.h"   // WrongMatrixForm (faugere-lachartre.tcc:32-45)
#include "lela/blas/context.h"
#include "lela/blas/level3-ll.h"
#include "lela/blas/level3.h"
#include "lela/matrix/dense.h"
#include "lela/ring/gf2.h"
#include "lela/ring/modular.h"
#include "lela/solutions/echelon-form.h"
#include "lela/util/error.h"
#include "lela/vector/traits.h"

#include "../mlz.h"

namespace LELA
{

/** One HIP context (stream + device workspace) shared by a LELA Context and
 * its copies: LELA's algorithm classes copy the Context they are given
 * (echelon-form.h:33, faugere-lachartre.h:41-42), and those copies run on
 * the same thread, so they share one stream. */
class HIPDevice
{
	struct Shared { mlz_ctx *ctx; long refs; };
	Shared *_s;

	void release ()
	{
		if (_s && --_s->refs == 0) {
			if (_s->ctx)
				mlz_ctx_destroy (_s->ctx);
			delete _s;
		}
		_s = 0;
	}

public:
	HIPDevice () : _s (new Shared)
	{
		_s->ctx = 0;
		_s->refs = 1;
		if (mlz_ctx_create (&_s->ctx, 0, 0) != MLZ_OK) {
			release ();
			throw LELAError ("HIPModule: no usable MI355X device");
		}
	}

	HIPDevice (const HIPDevice &d) : _s (d._s) { ++_s->refs; }

	HIPDevice &operator = (const HIPDevice &d)
	{
		if (d._s != _s) {
			release ();
			_s = d._s;
			++_s->refs;
		}
		return *this;
	}

	~HIPDevice () { release (); }

	mlz_ctx *ctx () const { return _s->ctx; }

	/** Maps a return code to LELA's exceptions (idx: trsm diagonal index). */
	void check (int rc, size_t idx = 0) const
	{
		if (rc == MLZ_OK)
			return;
		if (rc == MLZ_E_NOT_INVERTIBLE)
			throw DiagonalEntryNotInvertible (idx);
		throw LELAError (mlz_last_error (_s->ctx));
	}
};

template <class Element> struct HIPElement;
template <> struct HIPElement<uint16> { static mlz_elem value () { return MLZ_U16; } };
template <> struct HIPElement<uint32> { static mlz_elem value () { return MLZ_U32; } };
template <> struct HIPElement<double> { static mlz_elem value () { return MLZ_F64_SYM; } };

template <class Ring> struct HIPModule;

/** Z/p: everything not handled on the GPU falls back to ZpModule. */
template <class Element>
struct HIPModule<Modular<Element> > : public ZpModule<Element>
{
	struct Tag { typedef typename ZpModule<Element>::Tag Parent; };

	HIPDevice dev;

	HIPModule (const Modular<Element> &R) : ZpModule<Element> (R) {}
};

/** GF(2): everything not handled on the GPU falls back to the generic module. */
template <>
struct HIPModule<GF2> : public GenericModule<GF2>
{
	struct Tag { typedef GenericModule<GF2>::Tag Parent; };

	HIPDevice dev;

	HIPModule (const GF2 &R) : GenericModule<GF2> (R) {}
};

/** Canonical [0, p) value of an element as the C ABI takes scalars. */
template <class Element>
inline uint64_t hip_scalar (const Modular<Element> &F, const Element &x)
{
	if (x < 0)                                  // Modular<double> symmetric range
		return (uint64_t) (x + F._modulus);
	return (uint64_t) x;
}

/** First word and word stride (uint64 words between rows) of a dense 0-1 matrix. */
template <class Matrix>
inline uint64 *hip_words (Matrix &A)
	{ return &*A.rowBegin ()->word_begin (); }

template <class Matrix>
inline const uint64 *hip_words (const Matrix &A)
	{ return &*A.rowBegin ()->word_begin (); }

template <class Matrix>
inline size_t hip_row_words (const Matrix &A)
{
	if (A.rowdim () < 2)
		return (A.coldim () + 63) / 64;
	typename Matrix::ConstRowIterator r = A.rowBegin ();
	const uint64 *w0 = &*r->word_begin ();
	++r;
	return (size_t) (&*r->word_begin () - w0);
}

namespace BLAS3
{

template <class Element>
class _gemm<Modular<Element>, typename HIPModule<Modular<Element> >::Tag>
{
	typedef typename HIPModule<Modular<Element> >::Tag Tag;

	template <class Modules, class Matrix1, class Matrix2, class Matrix3>
	static Matrix3 &gemm_impl (const Modular<Element> &F, Modules &M, const Element &a, const Matrix1 &A,
				   const Matrix2 &B, const Element &b, Matrix3 &C,
				   MatrixStorageTypes::Generic, MatrixStorageTypes::Generic, MatrixStorageTypes::Generic)
		{ return _gemm<Modular<Element>, typename Tag::Parent>::op (F, M, a, A, B, b, C); }

	template <class Modules, class Matrix1, class Matrix2, class Matrix3>
	static Matrix3 &gemm_impl (const Modular<Element> &F, Modules &M, const Element &a, const Matrix1 &A,
				   const Matrix2 &B, const Element &b, Matrix3 &C,
				   MatrixStorageTypes::Dense, MatrixStorageTypes::Dense, MatrixStorageTypes::Dense)
	{
		lela_check (A.coldim () == B.rowdim ());
		lela_check (A.rowdim () == C.rowdim ());
		lela_check (B.coldim () == C.coldim ());
		if (C.rowdim () == 0 || C.coldim () == 0)
			return C;
		M.dev.check (mlz_zp_gemm (M.dev.ctx (), HIPElement<Element>::value (), (uint64_t) F._modulus,
					  C.rowdim (), C.coldim (), A.coldim (), hip_scalar (F, a),
					  A.coldim () ? &A[0][0] : 0, A.disp (), A.coldim () ? &B[0][0] : 0, B.disp (),
					  hip_scalar (F, b), &C[0][0], C.disp (), MLZ_ALGO_AUTO));
		return C;
	}

public:
	template <class Modules, class Matrix1, class Matrix2, class Matrix3>
	static Matrix3 &op (const Modular<Element> &F, Modules &M, const Element &a, const Matrix1 &A,
			    const Matrix2 &B, const Element &b, Matrix3 &C)
		{ return gemm_impl (F, M, a, A, B, b, C, typename Matrix1::StorageType (),
				    typename Matrix2::StorageType (), typename Matrix3::StorageType ()); }
};

template <class Element>
class _trsm<Modular<Element>, typename HIPModule<Modular<Element> >::Tag>
{
	typedef typename HIPModule<Modular<Element> >::Tag Tag;

	template <class Modules, class Matrix1, class Matrix2>
	static Matrix2 &trsm_impl (const Modular<Element> &F, Modules &M, const Element &a, const Matrix1 &A,
				   Matrix2 &B, TriangularMatrixType type, bool diagIsOne,
				   MatrixStorageTypes::Generic, MatrixStorageTypes::Generic)
		{ return _trsm<Modular<Element>, typename Tag::Parent>::op (F, M, a, A, B, type, diagIsOne); }

	template <class Modules, class Matrix1, class Matrix2>
	static Matrix2 &trsm_impl (const Modular<Element> &F, Modules &M, const Element &a, const Matrix1 &A,
				   Matrix2 &B, TriangularMatrixType type, bool diagIsOne,
				   MatrixStorageTypes::Dense, MatrixStorageTypes::Dense)
	{
		lela_check (A.coldim () == B.rowdim ());
		lela_check (A.rowdim () == B.rowdim ());
		if (A.rowdim () == 0 || B.coldim () == 0)
			return B;
		size_t bad = 0;
		const int rc = mlz_zp_trsm (M.dev.ctx (), HIPElement<Element>::value (), (uint64_t) F._modulus,
					    type == UpperTriangular, diagIsOne, A.rowdim (), B.coldim (),
					    hip_scalar (F, a), &A[0][0], A.disp (), &B[0][0], B.disp (), &bad,
					    MLZ_ALGO_AUTO);
		M.dev.check (rc, bad);
		return B;
	}

public:
	template <class Modules, class Matrix1, class Matrix2>
	static Matrix2 &op (const Modular<Element> &F, Modules &M, const Element &a, const Matrix1 &A, Matrix2 &B,
			    TriangularMatrixType type, bool diagIsOne)
		{ return trsm_impl (F, M, a, A, B, type, diagIsOne, typename Matrix1::StorageType (),
				    typename Matrix2::StorageType ()); }
};

template <>
class _gemm<GF2, HIPModule<GF2>::Tag>
{
	typedef HIPModule<GF2>::Tag Tag;

	template <class Modules, class Matrix1, class Matrix2, class Matrix3>
	static Matrix3 &gpu (const GF2 &, Modules &M, bool a, const Matrix1 &A, const Matrix2 &B, bool b, Matrix3 &C)
	{
		lela_check (A.coldim () == B.rowdim ());
		lela_check (A.rowdim () == C.rowdim ());
		lela_check (B.coldim () == C.coldim ());
		if (C.rowdim () == 0 || C.coldim () == 0)
			return C;
		M.dev.check (mlz_gf2_gemm (M.dev.ctx (), C.rowdim (), C.coldim (), A.coldim (), a,
					   A.coldim () && A.rowdim () ? hip_words (A) : 0, hip_row_words (A),
					   A.coldim () && B.coldim () ? hip_words (B) : 0, hip_row_words (B),
					   b, hip_words (C), hip_row_words (C)));
		return C;
	}

public:
	/* Anything but three little-endian dense 0-1 matrices: the parent module. */
	template <class Modules, class Matrix1, class Matrix2, class Matrix3>
	static Matrix3 &op (const GF2 &F, Modules &M, bool a, const Matrix1 &A, const Matrix2 &B, bool b, Matrix3 &C)
		{ return _gemm<GF2, Tag::Parent>::op (F, M, a, A, B, b, C); }

	template <class Modules, class It, class CIt>
	static Dense01Matrix<It, CIt, LittleEndian<uint64> > &
	op (const GF2 &F, Modules &M, bool a, const Dense01Matrix<It, CIt, LittleEndian<uint64> > &A,
	    const Dense01Matrix<It, CIt, LittleEndian<uint64> > &B, bool b, Dense01Matrix<It, CIt, LittleEndian<uint64> > &C)
		{ return gpu (F, M, a, A, B, b, C); }

#ifndef __LELA_HAVE_M4RI
	/* DenseMatrix<bool> is a little-endian Dense01Matrix without M4RI
	 * (dense-zero-one.h:369-394, bit-iterator.h:300-312); the exact-match
	 * overload keeps it from the generic template above. */
	template <class Modules>
	static DenseMatrix<bool> &op (const GF2 &F, Modules &M, bool a, const DenseMatrix<bool> &A,
				      const DenseMatrix<bool> &B, bool b, DenseMatrix<bool> &C)
		{ return gpu (F, M, a, A, B, b, C); }
#endif
};

template <>
class _trsm<GF2, HIPModule<GF2>::Tag>
{
	typedef HIPModule<GF2>::Tag Tag;

	template <class Modules, class Matrix1, class Matrix2>
	static Matrix2 &gpu (const GF2 &, Modules &M, bool a, const Matrix1 &A, Matrix2 &B, TriangularMatrixType type,
			     bool diagIsOne)
	{
		lela_check (A.coldim () == B.rowdim ());
		lela_check (A.rowdim () == B.rowdim ());
		if (A.rowdim () == 0 || B.coldim () == 0)
			return B;
		size_t bad = 0;
		const int rc = mlz_gf2_trsm (M.dev.ctx (), type == UpperTriangular, diagIsOne, A.rowdim (), B.coldim (), a,
					     hip_words (A), hip_row_words (A), hip_words (B), hip_row_words (B), &bad);
		M.dev.check (rc, bad);
		return B;
	}

public:
	/* Anything but two little-endian dense 0-1 matrices: the parent module. */
	template <class Modules, class Matrix1, class Matrix2>
	static Matrix2 &op (const GF2 &F, Modules &M, bool a, const Matrix1 &A, Matrix2 &B, TriangularMatrixType type,
			    bool diagIsOne)
		{ return _trsm<GF2, Tag::Parent>::op (F, M, a, A, B, type, diagIsOne); }

	template <class Modules, class It, class CIt>
	static Dense01Matrix<It, CIt, LittleEndian<uint64> > &
	op (const GF2 &F, Modules &M, bool a, const Dense01Matrix<It, CIt, LittleEndian<uint64> > &A,
	    Dense01Matrix<It, CIt, LittleEndian<uint64> > &B, TriangularMatrixType type, bool diagIsOne)
		{ return gpu (F, M, a, A, B, type, diagIsOne); }

#ifndef __LELA_HAVE_M4RI
	template <class Modules>
	static DenseMatrix<bool> &op (const GF2 &F, Modules &M, bool a, const DenseMatrix<bool> &A, DenseMatrix<bool> &B,
				      TriangularMatrixType type, bool diagIsOne)
		{ return gpu (F, M, a, A, B, type, diagIsOne); }
#endif
};

} // namespace BLAS3

/** Appends the nonzero entries of a row to (col, val): sparse rows as they
 * are, dense rows scanned for nonzeros. */
template <class Element, class Row>
inline void hip_append_row (const Modular<Element> &, const Row &row, std::vector<uint32_t> &col,
			    std::vector<Element> &val, VectorRepresentationTypes::Sparse)
{
	for (typename Row::const_iterator j = row.begin (); j != row.end (); ++j) {
		col.push_back ((uint32_t) j->first);
		val.push_back (j->second);
	}
}

template <class Element, class Row>
inline void hip_append_row (const Modular<Element> &F, const Row &row, std::vector<uint32_t> &col,
			    std::vector<Element> &val, VectorRepresentationTypes::Dense)
{
	uint32_t idx = 0;
	for (typename Row::const_iterator j = row.begin (); j != row.end (); ++j, ++idx)
		if (!F.isZero (*j)) {
			col.push_back (idx);
			val.push_back (*j);
		}
}

/** Elimination over Z/p on the HIP module (lela/algorithms/elimination.h:34-238).
 * echelonize (with and without compute_L) and pluq run on the GPU with the
 * pivot strategy LELA takes from the row type (DefaultPivotStrategy,
 * pivot-strategy.h:138-160): dense matrices with DensePivotStrategy
 * (mlz_zp_echelonize), matrices with sparse rows with
 * SparsePartialPivotStrategy (mlz_zp_echelonize_sparse, also when that
 * strategy is passed explicitly); both reproduce the transpositions.  Other
 * explicit strategies, echelonize_reduced, move_L and other row types run
 * LELA's Elimination on a ZpModule context. */
template <class Element>
class Elimination<Modular<Element>, HIPModule<Modular<Element> > >
{
public:
	typedef std::pair<uint32, uint32> Transposition;
	typedef std::vector<Transposition> Permutation;

private:
	Context<Modular<Element>, HIPModule<Modular<Element> > > &ctx;
	Context<Modular<Element>, ZpModule<Element> > _cpu_ctx;
	Elimination<Modular<Element>, ZpModule<Element> > _cpu;

	size_t gpu (DenseMatrix<Element> &A, int mode, Permutation &P, Permutation *Q, Element &det) const
	{
		const size_t r = std::min (A.rowdim (), A.coldim ());
		std::vector<uint32_t> pr (r + 1), pc (r + 1);
		size_t rank = 0;
		uint64_t d = 1;
		P.clear ();
		if (Q)
			Q->clear ();
		if (A.rowdim () && A.coldim ()) {
			const int rc = mlz_zp_echelonize (ctx.M.dev.ctx (), HIPElement<Element>::value (),
							  (uint64_t) ctx.F._modulus, A.rowdim (), A.coldim (), &A[0][0],
							  A.disp (), mode, &rank, &d, &pr[0], &pc[0], MLZ_ALGO_AUTO);
			if (rc == MLZ_E_NOT_INVERTIBLE)
				throw LELAError ("Could not invert pivot-element in the ring");   // elimination.tcc:103-104
			ctx.M.dev.check (rc);
		}
		for (size_t s = 0; s < rank; ++s) {
			if (pr[s] != s)
				P.push_back (Transposition ((uint32) s, (uint32) pr[s]));
			if (Q && pc[s] != s)
				Q->push_back (Transposition ((uint32) s, (uint32) pc[s]));
		}
		ctx.F.init (det, (long) d);
		return rank;
	}

	/** Sparse rows: unpacked into a dense array, reduced on the GPU with the
	 * sparse pivot rule, packed back without zeros (LELA's sparse axpy drops
	 * them, level1-generic.tcc:222-254). */
	template <class Matrix>
	size_t gpu_sparse (Matrix &A, int mode, Permutation &P, Permutation *Q, Element &det) const
	{
		const size_t m = A.rowdim (), n = A.coldim (), r = std::min (m, n);
		std::vector<Element> D (m * n, ctx.F.zero ());
		size_t i = 0;
		for (typename Matrix::RowIterator row = A.rowBegin (); row != A.rowEnd (); ++row, ++i)
			for (typename Matrix::Row::const_iterator j = row->begin (); j != row->end (); ++j)
				D[i * n + j->first] = j->second;
		std::vector<uint32_t> pr (r + 1), pc (r + 1);
		size_t rank = 0;
		uint64_t d = 1;
		P.clear ();
		if (Q)
			Q->clear ();
		if (m && n) {
			const int rc = mlz_zp_echelonize_sparse (ctx.M.dev.ctx (), HIPElement<Element>::value (),
								 (uint64_t) ctx.F._modulus, m, n, &D[0], n, mode, 0,
								 &rank, &d, &pr[0], &pc[0]);
			if (rc == MLZ_E_NOT_INVERTIBLE)
				throw LELAError ("Could not invert pivot-element in the ring");   // elimination.tcc:103-104
			ctx.M.dev.check (rc);
		}
		i = 0;
		for (typename Matrix::RowIterator row = A.rowBegin (); row != A.rowEnd (); ++row, ++i) {
			row->clear ();
			for (size_t j = 0; j < n; ++j)
				if (!ctx.F.isZero (D[i * n + j]))
					row->push_back (typename Matrix::Row::value_type (j, D[i * n + j]));
		}
		for (size_t s = 0; s < rank; ++s) {
			if (pr[s] != s)
				P.push_back (Transposition ((uint32) s, (uint32) pr[s]));
			if (Q && pc[s] != s)
				Q->push_back (Transposition ((uint32) s, (uint32) pc[s]));
		}
		ctx.F.init (det, (long) d);
		return rank;
	}

	template <class Matrix>
	Matrix &echelonize_rows (Matrix &A, Permutation &P, size_t &rank, Element &det, bool compute_L,
				 VectorRepresentationTypes::Sparse) const
	{
		rank = gpu_sparse (A, compute_L ? MLZ_ECH_TRANSFORM : MLZ_ECH_U, P, 0, det);
		return A;
	}

	template <class Matrix, class Rep>
	Matrix &echelonize_rows (Matrix &A, Permutation &P, size_t &rank, Element &det, bool compute_L, Rep) const
		{ return _cpu.echelonize (A, P, rank, det, compute_L); }

	template <class Matrix>
	Matrix &pluq_rows (Matrix &A, Permutation &P, Permutation &Q, size_t &rank, Element &det,
			   VectorRepresentationTypes::Sparse) const
	{
		rank = gpu_sparse (A, MLZ_ECH_PLUQ, P, &Q, det);
		std::reverse (P.begin (), P.end ());
		std::reverse (Q.begin (), Q.end ());
		return A;
	}

	template <class Matrix, class Rep>
	Matrix &pluq_rows (Matrix &A, Permutation &P, Permutation &Q, size_t &rank, Element &det, Rep) const
		{ return _cpu.pluq (A, P, Q, rank, det); }

public:
	Elimination (Context<Modular<Element>, HIPModule<Modular<Element> > > &_ctx)
		: ctx (_ctx), _cpu_ctx (_ctx.F), _cpu (_cpu_ctx) {}

	/** Dense: the GPU (elimination.tcc:47-139). */
	DenseMatrix<Element> &echelonize (DenseMatrix<Element> &A, Permutation &P, size_t &rank, Element &det,
					  bool compute_L = true) const
	{
		rank = gpu (A, compute_L ? MLZ_ECH_TRANSFORM : MLZ_ECH_U, P, 0, det);
		return A;
	}

	/** Sparse rows: the GPU with the sparse pivot rule; other rows: the CPU. */
	template <class Matrix>
	Matrix &echelonize (Matrix &A, Permutation &P, size_t &rank, Element &det, bool compute_L = true) const
	{
		return echelonize_rows (A, P, rank, det, compute_L,
					typename VectorTraits<Modular<Element>, typename Matrix::Row>::RepresentationType ());
	}

	template <class Matrix, class Modules>
	Matrix &echelonize (Matrix &A, Permutation &P, size_t &rank, Element &det,
			    SparsePartialPivotStrategy<Modular<Element>, Modules>, bool compute_L) const
	{
		return echelonize_rows (A, P, rank, det, compute_L,
					typename VectorTraits<Modular<Element>, typename Matrix::Row>::RepresentationType ());
	}

	template <class Matrix, class PivotStrategy>
	Matrix &echelonize (Matrix &A, Permutation &P, size_t &rank, Element &det, PivotStrategy PS, bool compute_L) const
		{ return _cpu.echelonize (A, P, rank, det, PS, compute_L); }

	template <class Matrix1, class Matrix2>
	Matrix1 &echelonize_reduced (Matrix1 &A, Matrix2 &L, Permutation &P, size_t &rank, Element &det,
				     bool compute_L = false) const
		{ return _cpu.echelonize_reduced (A, L, P, rank, det, compute_L); }

	template <class Matrix1, class Matrix2, class PivotStrategy>
	Matrix1 &echelonize_reduced (Matrix1 &A, Matrix2 &L, Permutation &P, size_t &rank, Element &det,
				     PivotStrategy PS, bool compute_L) const
		{ return _cpu.echelonize_reduced (A, L, P, rank, det, PS, compute_L); }

	/** Dense: the GPU (elimination.tcc:266-370; P and Q in LELA's reversed order, :357-358). */
	DenseMatrix<Element> &pluq (DenseMatrix<Element> &A, Permutation &P, Permutation &Q, size_t &rank,
				    Element &det) const
	{
		rank = gpu (A, MLZ_ECH_PLUQ, P, &Q, det);
		std::reverse (P.begin (), P.end ());
		std::reverse (Q.begin (), Q.end ());
		return A;
	}

	/** Sparse rows: the GPU with the sparse pivot rule; other rows: the CPU. */
	template <class Matrix>
	Matrix &pluq (Matrix &A, Permutation &P, Permutation &Q, size_t &rank, Element &det) const
	{
		return pluq_rows (A, P, Q, rank, det,
				  typename VectorTraits<Modular<Element>, typename Matrix::Row>::RepresentationType ());
	}

	template <class Matrix, class PivotStrategy>
	Matrix &pluq (Matrix &A, Permutation &P, Permutation &Q, size_t &rank, Element &det, PivotStrategy PS) const
		{ return _cpu.pluq (A, P, Q, rank, det, PS); }

	template <class Matrix1, class Matrix2>
	void move_L (Matrix1 &L, Matrix2 &A) const { _cpu.move_L (L, A); }
};

/** FaugereLachartre over Z/p on the HIP module
 * (lela/algorithms/faugere-lachartre.h:40-81): the F4 matrix X goes to the
 * device as CSR (any row type), R comes back dense and is copied into R.
 * The non-reduced and only_D forms run LELA's CPU FaugereLachartre. */
template <class Element>
class FaugereLachartre<Modular<Element>, HIPModule<Modular<Element> > >
{
	Context<Modular<Element>, HIPModule<Modular<Element> > > &ctx;
	bool _det_wanted;

	template <class Matrix>
	void cpu (Matrix &R, const Matrix &X, size_t &rank, Element &det, bool reduced, bool only_D)
	{
		Context<Modular<Element>, ZpModule<Element> > cctx (ctx.F);
		FaugereLachartre<Modular<Element>, ZpModule<Element> > FL (cctx);
		FL.echelonize (R, X, rank, det, reduced, only_D);
	}

public:
	FaugereLachartre (Context<Modular<Element>, HIPModule<Modular<Element> > > &_ctx)
		: ctx (_ctx), _det_wanted (true) {}

	/** Whether the caller uses det.  If not, the GPU skips the non-reduced
	 * echelon form of D that det needs, and det is set to 1. */
	void det_wanted (bool wanted) { _det_wanted = wanted; }

	template <class Matrix>
	void echelonize (Matrix &R, const Matrix &X, size_t &rank, Element &det, bool reduced = true, bool only_D = false)
	{
		if (!reduced || only_D) {
			cpu (R, X, rank, det, reduced, only_D);
			return;
		}

		std::vector<uint64_t> row_ptr (1, 0);
		std::vector<uint32_t> col;
		std::vector<Element> val;
		for (typename Matrix::ConstRowIterator i = X.rowBegin (); i != X.rowEnd (); ++i) {
			hip_append_row (ctx.F, *i, col, val,
					typename VectorTraits<Modular<Element>, typename Matrix::ConstRow>::RepresentationType ());
			row_ptr.push_back (col.size ());
		}

		DenseMatrix<Element> Rd (X.rowdim (), X.coldim ());
		std::vector<uint32_t> pivots (std::min (X.rowdim (), X.coldim ()) + 1);
		uint64_t det_c = 1;
		size_t bad_row = 0, r = 0;
		const int rc = mlz_zp_fl_echelonize (ctx.M.dev.ctx (), HIPElement<Element>::value (),
						     (uint64_t) ctx.F._modulus, X.rowdim (), X.coldim (), &row_ptr[0],
						     col.empty () ? 0 : &col[0], val.empty () ? 0 : &val[0],
						     X.rowdim () && X.coldim () ? &Rd[0][0] : 0, Rd.disp (), &r,
						     &pivots[0], _det_wanted ? &det_c : 0, &bad_row, MLZ_ALGO_AUTO);
		if (rc == MLZ_E_WRONG_FORM)
			throw WrongMatrixForm (bad_row);
		if (rc == MLZ_E_NOT_INVERTIBLE) {        // LELA leaves R as it was (:790-799)
			rank = r;
			ctx.F.init (det, (long) det_c);
			return;
		}
		ctx.M.dev.check (rc);
		rank = r;
		ctx.F.init (det, (long) det_c);
		BLAS3::copy (ctx, Rd, R);
	}
};

/** EchelonForm over Z/p on the HIP module (lela/solutions/echelon-form.h:31-164).
 * Dense matrices, sparse rows and the Faugere-Lachartre method run on the
 * GPU; other storage runs LELA's own Elimination on a ZpModule context.  The dense
 * non-reduced form is Elimination's with transposition pivoting, which both
 * methods give (echelon-form.h:113-131, tests/test-gauss-jordan.C:92-99). */
template <class Element>
class EchelonForm<Modular<Element>, HIPModule<Modular<Element> > >
{
	Context<Modular<Element>, HIPModule<Modular<Element> > > &_ctx;
	Context<Modular<Element>, ZpModule<Element> > _cpu;
	std::map<const void *, size_t> _rank_table;

	template <class Matrix>
	size_t cpu_echelonize (Matrix &A, bool reduced)
	{
		Elimination<Modular<Element>, ZpModule<Element> > elim (_cpu);
		typename Elimination<Modular<Element>, ZpModule<Element> >::Permutation P;
		size_t rank = 0;
		Element d;
		if (reduced) {
			DenseMatrix<Element> L;
			elim.echelonize_reduced (A, L, P, rank, d, false);
		} else
			elim.echelonize (A, P, rank, d, false);
		return rank;
	}

	/** Sparse rows, standard method (:86-91): the non-reduced form is
	 * Elimination's with the sparse pivot rule, on the GPU; the reduced form
	 * is unique, so the dense reduced kernel gives it. */
	template <class Matrix>
	size_t std_echelonize (Matrix &A, bool reduced, VectorRepresentationTypes::Sparse)
	{
		size_t rank = 0;
		Element d;
		if (!reduced) {
			Elimination<Modular<Element>, HIPModule<Modular<Element> > > elim (_ctx);
			typename Elimination<Modular<Element>, HIPModule<Modular<Element> > >::Permutation P;
			elim.echelonize (A, P, rank, d, false);
			return rank;
		}
		const size_t m = A.rowdim (), n = A.coldim ();
		if (!m || !n)
			return 0;
		std::vector<Element> D (m * n, _ctx.F.zero ());
		size_t i = 0;
		for (typename Matrix::RowIterator row = A.rowBegin (); row != A.rowEnd (); ++row, ++i)
			for (typename Matrix::Row::const_iterator j = row->begin (); j != row->end (); ++j)
				D[i * n + j->first] = j->second;
		std::vector<uint32_t> pivots (std::min (m, n) + 1);
		_ctx.M.dev.check (mlz_zp_rref (_ctx.M.dev.ctx (), HIPElement<Element>::value (), (uint64_t) _ctx.F._modulus,
					       m, n, &D[0], n, &rank, &pivots[0], MLZ_ALGO_AUTO));
		i = 0;
		for (typename Matrix::RowIterator row = A.rowBegin (); row != A.rowEnd (); ++row, ++i) {
			row->clear ();
			for (size_t j = 0; j < n; ++j)
				if (!_ctx.F.isZero (D[i * n + j]))
					row->push_back (typename Matrix::Row::value_type (j, D[i * n + j]));
		}
		return rank;
	}

	template <class Matrix, class Rep>
	size_t std_echelonize (Matrix &A, bool reduced, Rep)
		{ return cpu_echelonize (A, reduced); }

	template <class Matrix>
	size_t fl_echelonize (Matrix &A, bool reduced)
	{
		FaugereLachartre<Modular<Element>, HIPModule<Modular<Element> > > FL (_ctx);
		FL.det_wanted (false);                  // EchelonForm keeps only the rank
		size_t rank = 0;
		Element d;
		FL.echelonize (A, A, rank, d, reduced);
		return rank;
	}

public:
	enum Method { METHOD_UNKNOWN, METHOD_STANDARD_GJ, METHOD_ASYMPTOTICALLY_FAST_GJ, METHOD_FAUGERE_LACHARTRE };

	EchelonForm (Context<Modular<Element>, HIPModule<Modular<Element> > > &ctx)
		: _ctx (ctx), _cpu (ctx.F) {}

	/** Any matrix: the standard method (sparse rows on the GPU, other rows
	 * LELA's CPU Elimination), or Faugere-Lachartre with the caller's
	 * `reduced` (:86-94). */
	template <class Matrix>
	Matrix &echelonize (Matrix &A, bool reduced = false, Method method = METHOD_STANDARD_GJ)
	{
		size_t rank;
		if (method == METHOD_STANDARD_GJ)
			rank = std_echelonize (A, reduced,
					       typename VectorTraits<Modular<Element>, typename Matrix::Row>::RepresentationType ());
		else if (method == METHOD_FAUGERE_LACHARTRE)
			rank = fl_echelonize (A, reduced);
		else
			throw LELAError ("Invalid method for choice of matrix");
		_rank_table[&A] = rank;
		return A;
	}

	/** Dense matrices (:103-154). */
	DenseMatrix<Element> &echelonize (DenseMatrix<Element> &A, bool reduced = false,
					  Method method = METHOD_ASYMPTOTICALLY_FAST_GJ)
	{
		size_t rank = 0;
		switch (method) {
		case METHOD_STANDARD_GJ:
		case METHOD_ASYMPTOTICALLY_FAST_GJ:
			if (!reduced) {
				// Elimination::echelonize (:78); GaussJordan + move_L gives the
				// same form (:128-131, tests/test-gauss-jordan.C:92-99)
				uint64_t d = 1;
				if (A.rowdim () && A.coldim ())
					_ctx.M.dev.check (mlz_zp_echelonize (_ctx.M.dev.ctx (), HIPElement<Element>::value (),
									     (uint64_t) _ctx.F._modulus, A.rowdim (),
									     A.coldim (), &A[0][0], A.disp (), MLZ_ECH_U,
									     &rank, &d, 0, 0, MLZ_ALGO_AUTO));
				break;
			}
			if (A.rowdim () && A.coldim ()) {
				std::vector<uint32_t> pivots (std::min (A.rowdim (), A.coldim ()) + 1);
				_ctx.M.dev.check (mlz_zp_rref (_ctx.M.dev.ctx (), HIPElement<Element>::value (),
							       (uint64_t) _ctx.F._modulus, A.rowdim (), A.coldim (),
							       &A[0][0], A.disp (), &rank, &pivots[0], MLZ_ALGO_AUTO));
			}
			break;

		case METHOD_FAUGERE_LACHARTRE:
			// LELA's dense overload runs the reduced FL whatever `reduced` says (:133-139)
			rank = fl_echelonize (A, true);
			break;

		default:
			throw LELAError ("Invalid method for choice of matrix");
		}
		_rank_table[&A] = rank;
		return A;
	}

	/** Rank recorded for A by echelonize; 0 for a matrix never passed to it. */
	template <class Matrix>
	size_t rank (const Matrix &A) const
	{
		std::map<const void *, size_t>::const_iterator i = _rank_table.find (&A);
		return i == _rank_table.end () ? 0 : i->second;
	}
};

/** The nonzero words of a row of a 0-1 matrix, with their word indices, for
 * the hybrid-row ABI: rows that are hybrid vectors with little-endian 64-bit
 * words (lela/vector/hybrid.h:30-33) go as they are; false for any other row
 * (LELA's CPU code then runs). */
template <class Row>
inline bool hip_hybrid_words (const Row &row, std::vector<uint32_t> &blk, std::vector<uint64_t> &words,
			      LittleEndian<uint64>)
{
	for (typename Row::const_iterator j = row.begin (); j != row.end (); ++j) {
		blk.push_back ((uint32_t) j->first);
		words.push_back ((uint64_t) j->second);
	}
	return true;
}

template <class Row, class Endianness>
inline bool hip_hybrid_words (const Row &, std::vector<uint32_t> &, std::vector<uint64_t> &, Endianness)
	{ return false; }

template <class Row>
inline bool hip_hybrid_row (const Row &row, std::vector<uint32_t> &blk, std::vector<uint64_t> &words,
			    VectorRepresentationTypes::Hybrid01)
	{ return hip_hybrid_words (row, blk, words, typename Row::Endianness ()); }

template <class Row>
inline bool hip_hybrid_row (const Row &, std::vector<uint32_t> &, std::vector<uint64_t> &,
			    VectorRepresentationTypes::Generic)
	{ return false; }

/** FaugereLachartre over GF(2) on the HIP module
 * (lela/algorithms/faugere-lachartre.h:40-81): X with hybrid rows of
 * little-endian words (the reference test's SparseMatrix<bool, HybridVector<..>>,
 * tests/test-faugere-lachartre.C:278-285) or a dense 0-1 matrix goes to the
 * GPU (mlz_gf2_fl_echelonize) and R comes back dense; other rows, the
 * non-reduced form and only_D run LELA's CPU FaugereLachartre.  Every head is
 * 1 over GF(2), so det is true. */
template <>
class FaugereLachartre<GF2, HIPModule<GF2> >
{
	Context<GF2, HIPModule<GF2> > &ctx;

	template <class Matrix>
	void cpu (Matrix &R, const Matrix &X, size_t &rank, bool &det, bool reduced, bool only_D)
	{
		Context<GF2, GenericModule<GF2> > cctx (ctx.F);
		FaugereLachartre<GF2, GenericModule<GF2> > FL (cctx);
		FL.echelonize (R, X, rank, det, reduced, only_D);
	}

	/* true when the GPU ran: rank, det and Rd (m x n) filled */
	bool gpu (const std::vector<uint64_t> &row_ptr, const std::vector<uint32_t> &blk,
		  const std::vector<uint64_t> &words, size_t m, size_t n, DenseMatrix<bool> &Rd, size_t &rank, bool &det)
	{
		size_t bad_row = 0, r = 0;
		const int rc = mlz_gf2_fl_echelonize (ctx.M.dev.ctx (), m, n, &row_ptr[0], blk.empty () ? 0 : &blk[0],
						      words.empty () ? 0 : &words[0], m && n ? hip_words (Rd) : 0,
						      hip_row_words (Rd), &r, 0, &bad_row);
		if (rc == MLZ_E_WRONG_FORM)
			throw WrongMatrixForm (bad_row);
		ctx.M.dev.check (rc);
		rank = r;
		det = true;
		return true;
	}

public:
	FaugereLachartre (Context<GF2, HIPModule<GF2> > &_ctx) : ctx (_ctx) {}

	template <class Matrix>
	void echelonize (Matrix &R, const Matrix &X, size_t &rank, bool &det, bool reduced = true, bool only_D = false)
	{
		std::vector<uint64_t> row_ptr (1, 0);
		std::vector<uint32_t> blk;
		std::vector<uint64_t> words;
		bool ok = reduced && !only_D;
		for (typename Matrix::ConstRowIterator i = X.rowBegin (); ok && i != X.rowEnd (); ++i) {
			ok = hip_hybrid_row (*i, blk, words,
					     typename VectorTraits<GF2, typename Matrix::ConstRow>::RepresentationType ());
			row_ptr.push_back (blk.size ());
		}
#ifndef __LELA_HAVE_M4RI
		if (ok) {
			DenseMatrix<bool> Rd (X.rowdim (), X.coldim ());
			gpu (row_ptr, blk, words, X.rowdim (), X.coldim (), Rd, rank, det);
			BLAS3::copy (ctx, Rd, R);
			return;
		}
#endif
		cpu (R, X, rank, det, reduced, only_D);
	}

#ifndef __LELA_HAVE_M4RI
	/** Dense 0-1 matrices (little-endian words without M4RI): the nonzero words as hybrid rows. */
	void echelonize (DenseMatrix<bool> &R, const DenseMatrix<bool> &X, size_t &rank, bool &det, bool reduced = true,
			 bool only_D = false)
	{
		if (!reduced || only_D) {
			cpu (R, X, rank, det, reduced, only_D);
			return;
		}
		const size_t m = X.rowdim (), n = X.coldim (), nw = (n + 63) / 64, ld = m && n ? hip_row_words (X) : 0;
		std::vector<uint64_t> row_ptr (1, 0);
		std::vector<uint32_t> blk;
		std::vector<uint64_t> words;
		const uint64 *W = m && n ? hip_words (X) : 0;
		for (size_t i = 0; i < m; ++i) {
			for (size_t w = 0; w < nw; ++w)
				if (W[i * ld + w]) {
					blk.push_back ((uint32_t) w);
					words.push_back (W[i * ld + w]);
				}
			row_ptr.push_back (blk.size ());
		}
		DenseMatrix<bool> Rd (m, n);
		gpu (row_ptr, blk, words, m, n, Rd, rank, det);
		BLAS3::copy (ctx, Rd, R);
	}
#endif
};

/** EchelonForm over GF(2) on the HIP module.  On the GPU: dense 0-1
 * matrices (reduced: mlz_gf2_rref; non-reduced: Elimination's form over Z/2),
 * sparse and hybrid 0-1 rows with the standard method (reduced: mlz_gf2_rref;
 * non-reduced: the sparse pivot rule over Z/2) and the Faugere-Lachartre
 * method on hybrid and dense rows; other rows run LELA's Elimination /
 * FaugereLachartre on a GenericModule context.  The reference's
 * own GF(2) specialisation (echelon-form-gf2.h:37-182) exists only with M4RI;
 * without it EchelonForm<GF2> uses the generic class (echelon-form.h:31-164),
 * whose dense overload runs GaussJordan (reduced: the same RREF). */
template <>
class EchelonForm<GF2, HIPModule<GF2> >
{
	Context<GF2, HIPModule<GF2> > &_ctx;
	Context<GF2, GenericModule<GF2> > _cpu;
	std::map<const void *, size_t> _rank_table;

	template <class Matrix>
	size_t cpu_echelonize (Matrix &A, bool reduced)
	{
		Elimination<GF2, GenericModule<GF2> > elim (_cpu);
		Elimination<GF2, GenericModule<GF2> >::Permutation P;
		size_t rank = 0;
		bool d;
		if (reduced) {
			DenseMatrix<bool> L;
			elim.echelonize_reduced (A, L, P, rank, d, false);
		} else
			elim.echelonize (A, P, rank, d, false);
		return rank;
	}

	/** Rows of 0-1 matrices as one 16-bit Z/2 entry per column (the GPU's
	 * non-reduced forms over GF(2) run over Z/2: the same eliminations) and
	 * back; false for rows the ABI does not take (LELA's CPU code then runs). */
	template <class Row>
	static bool unpack_row (const Row &row, uint16_t *D, size_t, VectorRepresentationTypes::Sparse01)
	{
		for (typename Row::const_iterator j = row.begin (); j != row.end (); ++j)
			D[*j] = 1;
		return true;
	}

	template <class Row>
	static bool unpack_row (const Row &row, uint16_t *D, size_t n, VectorRepresentationTypes::Hybrid01)
	{
		std::vector<uint32_t> blk;
		std::vector<uint64_t> words;
		if (!hip_hybrid_row (row, blk, words, VectorRepresentationTypes::Hybrid01 ()))
			return false;
		for (size_t e = 0; e < blk.size (); ++e)
			for (size_t b = 0; b < 64 && 64 * (size_t) blk[e] + b < n; ++b)
				if ((words[e] >> b) & 1)
					D[64 * (size_t) blk[e] + b] = 1;
		return true;
	}

	template <class Row>
	static void pack_row (Row &row, const uint16_t *D, size_t n, VectorRepresentationTypes::Sparse01)
	{
		row.clear ();
		for (size_t j = 0; j < n; ++j)
			if (D[j])
				row.push_back (j);
	}

	template <class Row>
	static void pack_row (Row &row, const uint16_t *D, size_t n, VectorRepresentationTypes::Hybrid01)
	{
		row.clear ();
		for (size_t w = 0; 64 * w < n; ++w) {
			uint64_t x = 0;
			for (size_t b = 0; b < 64 && 64 * w + b < n; ++b)
				x |= (uint64_t) (D[64 * w + b] & 1) << b;
			if (x)
				row.push_back (typename Row::value_type (w, x));
		}
	}

	/** Sparse and hybrid 0-1 rows, standard method (echelon-form.h:86-91):
	 * the reduced form is unique (mlz_gf2_rref on the bits); the non-reduced
	 * form is Elimination's with SparsePartialPivotStrategy, which counts
	 * entries for sparse rows and words for hybrid rows
	 * (pivot-strategy.tcc:101-188). */
	template <class Matrix, class Rep>
	size_t std_sparse (Matrix &A, bool reduced, Rep rep)
	{
		const size_t m = A.rowdim (), n = A.coldim ();
		if (!m || !n)
			return 0;
		std::vector<uint16_t> D (m * n, 0);
		size_t i = 0;
		for (typename Matrix::RowIterator row = A.rowBegin (); row != A.rowEnd (); ++row, ++i)
			if (!unpack_row (*row, &D[i * n], n, rep))
				return cpu_echelonize (A, reduced);
		size_t rank = 0;
		if (reduced) {
			const size_t nw = (n + 63) / 64;
			std::vector<uint64_t> W (m * nw, 0);
			for (size_t r = 0; r < m; ++r)
				for (size_t j = 0; j < n; ++j)
					W[r * nw + j / 64] |= (uint64_t) D[r * n + j] << (j % 64);
			std::vector<uint32_t> pivots (std::min (m, n) + 1);
			_ctx.M.dev.check (mlz_gf2_rref (_ctx.M.dev.ctx (), m, n, &W[0], nw, &rank, &pivots[0]));
			for (size_t r = 0; r < m; ++r)
				for (size_t j = 0; j < n; ++j)
					D[r * n + j] = (uint16_t) ((W[r * nw + j / 64] >> (j % 64)) & 1);
		} else {
			uint64_t d = 1;
			_ctx.M.dev.check (mlz_zp_echelonize_sparse (_ctx.M.dev.ctx (), MLZ_U16, 2, m, n, &D[0], n, MLZ_ECH_U,
								    hip_counts_words (rep), &rank, &d, 0, 0));
		}
		i = 0;
		for (typename Matrix::RowIterator row = A.rowBegin (); row != A.rowEnd (); ++row, ++i)
			pack_row (*row, &D[i * n], n, rep);
		return rank;
	}

	static int hip_counts_words (VectorRepresentationTypes::Sparse01) { return 0; }
	static int hip_counts_words (VectorRepresentationTypes::Hybrid01) { return 1; }

	template <class Matrix>
	size_t std_echelonize (Matrix &A, bool reduced, VectorRepresentationTypes::Sparse01 r)
		{ return std_sparse (A, reduced, r); }

	template <class Matrix>
	size_t std_echelonize (Matrix &A, bool reduced, VectorRepresentationTypes::Hybrid01 r)
		{ return std_sparse (A, reduced, r); }

	template <class Matrix, class Rep>
	size_t std_echelonize (Matrix &A, bool reduced, Rep)
		{ return cpu_echelonize (A, reduced); }

	template <class Matrix>
	size_t cpu_fl (Matrix &A, bool reduced)
	{
		FaugereLachartre<GF2, HIPModule<GF2> > FL (_ctx);   // the GPU where the rows allow it
		size_t rank = 0;
		bool d;
		FL.echelonize (A, A, rank, d, reduced);
		return rank;
	}

public:
	enum Method { METHOD_UNKNOWN, METHOD_STANDARD_GJ, METHOD_ASYMPTOTICALLY_FAST_GJ, METHOD_M4RI,
		      METHOD_FAUGERE_LACHARTRE };

	EchelonForm (Context<GF2, HIPModule<GF2> > &ctx) : _ctx (ctx), _cpu (ctx.F) {}

	template <class Matrix>
	Matrix &echelonize (Matrix &A, bool reduced = false, Method method = METHOD_STANDARD_GJ)
	{
		size_t rank;
		if (method == METHOD_STANDARD_GJ)
			rank = std_echelonize (A, reduced, typename VectorTraits<GF2, typename Matrix::Row>::RepresentationType ());
		else if (method == METHOD_FAUGERE_LACHARTRE)
			rank = cpu_fl (A, reduced);
		else
			throw LELAError ("Invalid method for choice of matrix");
		_rank_table[&A] = rank;
		return A;
	}

#ifndef __LELA_HAVE_M4RI
	/** Dense 0-1 matrices: the reduced form on the GPU. */
	DenseMatrix<bool> &echelonize (DenseMatrix<bool> &A, bool reduced = false,
				       Method method = METHOD_ASYMPTOTICALLY_FAST_GJ)
	{
		size_t rank = 0;
		if (method == METHOD_FAUGERE_LACHARTRE)
			rank = cpu_fl (A, true);
		else if (method != METHOD_STANDARD_GJ && method != METHOD_ASYMPTOTICALLY_FAST_GJ &&
			 method != METHOD_M4RI)
			throw LELAError ("Invalid method for choice of matrix");
		else if (!reduced) {
			// Elimination::echelonize with DensePivotStrategy (echelon-form.h:113-131) over Z/2
			const size_t m = A.rowdim (), n = A.coldim (), nw = hip_row_words (A);
			if (m && n) {
				const uint64_t *W = hip_words (A);
				std::vector<uint16_t> D (m * n);
				for (size_t r = 0; r < m; ++r)
					for (size_t j = 0; j < n; ++j)
						D[r * n + j] = (uint16_t) ((W[r * nw + j / 64] >> (j % 64)) & 1);
				uint64_t d = 1;
				_ctx.M.dev.check (mlz_zp_echelonize (_ctx.M.dev.ctx (), MLZ_U16, 2, m, n, &D[0], n, MLZ_ECH_U,
								     &rank, &d, 0, 0, MLZ_ALGO_AUTO));
				uint64_t *Wo = hip_words (A);
				for (size_t r = 0; r < m; ++r)
					for (size_t j = 0; j < n; ++j) {
						const uint64_t bit = 1ull << (j % 64);
						Wo[r * nw + j / 64] = D[r * n + j] ? (Wo[r * nw + j / 64] | bit)
										   : (Wo[r * nw + j / 64] & ~bit);
					}
			}
		} else if (A.rowdim () && A.coldim ()) {
			std::vector<uint32_t> pivots (std::min (A.rowdim (), A.coldim ()) + 1);
			_ctx.M.dev.check (mlz_gf2_rref (_ctx.M.dev.ctx (), A.rowdim (), A.coldim (), hip_words (A),
							hip_row_words (A), &rank, &pivots[0]));
		}
		_rank_table[&A] = rank;
		return A;
	}
#endif

	template <class Matrix>
	size_t rank (const Matrix &A) const
	{
		std::map<const void *, size_t>::const_iterator i = _rank_table.find (&A);
		return i == _rank_table.end () ? 0 : i->second;
	}
};

} // namespace LELA

#endif // LELA_AMD_HIP_MODULE_H
