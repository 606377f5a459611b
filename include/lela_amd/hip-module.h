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
 * EchelonForm is specialised as a whole class, like the GF2 + M4RI
 * specialisation in lela/solutions/echelon-form-gf2.h:37-182.
 *
 * This header is compiled inside a LELA build (it includes LELA's headers)
 * and links against lela_amd/liblela_amd.so.  Return codes of the C ABI are
 * turned back into LELA's exceptions.
 */
#ifndef LELA_AMD_HIP_MODULE_H
#define LELA_AMD_HIP_MODULE_H

#include <map>
#include <vector>

#include "lela/algorithms/faugere-lachartre.h"   // WrongMatrixForm (faugere-lachartre.tcc:32-45)
#include "lela/blas/context.h"
#include "lela/blas/level3-ll.h"
#include "lela/matrix/dense.h"
#include "lela/ring/gf2.h"
#include "lela/ring/modular.h"
#include "lela/solutions/echelon-form.h"
#include "lela/util/error.h"

#include "../mlz.h"

namespace LELA
{

/** One HIP context (stream + device workspace) per LELA Context. */
class HIPDevice
{
	mlz_ctx *_ctx;

	HIPDevice (const HIPDevice &);
	HIPDevice &operator = (const HIPDevice &);

public:
	HIPDevice () : _ctx (0)
	{
		if (mlz_ctx_create (&_ctx, 0, 0) != MLZ_OK)
			throw LELAError ("HIPModule: no usable MI355X device");
	}

	~HIPDevice () { if (_ctx) mlz_ctx_destroy (_ctx); }

	mlz_ctx *ctx () const { return _ctx; }

	/** Maps a return code to LELA's exceptions (idx: trsm diagonal index). */
	void check (int rc, size_t idx = 0) const
	{
		if (rc == MLZ_OK)
			return;
		if (rc == MLZ_E_NOT_INVERTIBLE)
			throw DiagonalEntryNotInvertible (idx);
		throw LELAError (mlz_last_error (_ctx));
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

/** Word stride (uint64 words between rows) of a Dense01Matrix. */
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

template <>
class _gemm<GF2, HIPModule<GF2>::Tag>
{
	typedef HIPModule<GF2>::Tag Tag;

public:
	/* Anything but three little-endian Dense01Matrix: the parent module. */
	template <class Modules, class Matrix1, class Matrix2, class Matrix3>
	static Matrix3 &op (const GF2 &F, Modules &M, bool a, const Matrix1 &A, const Matrix2 &B, bool b, Matrix3 &C)
		{ return _gemm<GF2, Tag::Parent>::op (F, M, a, A, B, b, C); }

	template <class Modules, class It, class CIt>
	static Dense01Matrix<It, CIt, LittleEndian<uint64> > &
	op (const GF2 &F, Modules &M, bool a, const Dense01Matrix<It, CIt, LittleEndian<uint64> > &A,
	    const Dense01Matrix<It, CIt, LittleEndian<uint64> > &B, bool b, Dense01Matrix<It, CIt, LittleEndian<uint64> > &C)
	{
		lela_check (A.coldim () == B.rowdim ());
		lela_check (A.rowdim () == C.rowdim ());
		lela_check (B.coldim () == C.coldim ());
		if (C.rowdim () == 0 || C.coldim () == 0)
			return C;
		M.dev.check (mlz_gf2_gemm (M.dev.ctx (), C.rowdim (), C.coldim (), A.coldim (), a,
					   A.coldim () ? &*A.rowBegin ()->word_begin () : 0, hip_row_words (A),
					   A.coldim () ? &*B.rowBegin ()->word_begin () : 0, hip_row_words (B),
					   b, &*C.rowBegin ()->word_begin (), hip_row_words (C)));
		return C;
	}
};

} // namespace BLAS3

/** FaugereLachartre over Z/p on the HIP module
 * (lela/algorithms/faugere-lachartre.h:40-81): the F4 matrix X goes to the
 * device as CSR, R comes back dense and is written into R row by row. */
template <class Element>
class FaugereLachartre<Modular<Element>, HIPModule<Modular<Element> > >
{
	Context<Modular<Element>, HIPModule<Modular<Element> > > &ctx;

public:
	FaugereLachartre (Context<Modular<Element>, HIPModule<Modular<Element> > > &_ctx) : ctx (_ctx) {}

	/* Only the reduced form (reduced = true, only_D = false) is offered by
	 * the device path; the other forms throw. */
	template <class Matrix>
	void echelonize (Matrix &R, const Matrix &X, size_t &rank, Element &det, bool reduced = true, bool only_D = false)
	{
		if (!reduced || only_D)
			throw LELAError ("HIPModule: FaugereLachartre runs the reduced form only");

		std::vector<uint64_t> row_ptr (1, 0);
		std::vector<uint32_t> col;
		std::vector<Element> val;
		for (typename Matrix::ConstRowIterator i = X.rowBegin (); i != X.rowEnd (); ++i) {
			for (typename Matrix::ConstRow::const_iterator j = i->begin (); j != i->end (); ++j) {
				col.push_back ((uint32_t) j->first);
				val.push_back (j->second);
			}
			row_ptr.push_back (col.size ());
		}

		DenseMatrix<Element> Rd (X.rowdim (), X.coldim ());
		uint64_t det_c = 1;
		size_t bad_row = 0;
		const int rc = mlz_zp_fl_echelonize (ctx.M.dev.ctx (), HIPElement<Element>::value (),
						     (uint64_t) ctx.F._modulus, X.rowdim (), X.coldim (), &row_ptr[0],
						     col.empty () ? 0 : &col[0], val.empty () ? 0 : &val[0],
						     &Rd[0][0], Rd.disp (), &rank, 0, &det_c, &bad_row, MLZ_ALGO_AUTO);
		ctx.F.init (det, (long) det_c);
		if (rc == MLZ_E_WRONG_FORM)
			throw WrongMatrixForm (bad_row);
		if (rc == MLZ_E_NOT_INVERTIBLE)
			return;                         // LELA leaves R as it was (:790-799)
		ctx.M.dev.check (rc);
		BLAS3::copy (ctx, Rd, R);
	}
};

/** EchelonForm over Z/p on the HIP module: the reduced form of a dense matrix
 * runs on the GPU (mlz_zp_rref); the non-reduced forms, whose layout depends
 * on LELA's pivot order, run LELA's own Elimination on a ZpModule context. */
template <class Element>
class EchelonForm<Modular<Element>, HIPModule<Modular<Element> > >
{
	Context<Modular<Element>, HIPModule<Modular<Element> > > &_ctx;
	Context<Modular<Element>, ZpModule<Element> > _cpu;
	EchelonForm<Modular<Element>, ZpModule<Element> > _fallback;
	std::map<const void *, size_t> _rank_table;

public:
	enum Method { METHOD_UNKNOWN, METHOD_STANDARD_GJ, METHOD_ASYMPTOTICALLY_FAST_GJ, METHOD_FAUGERE_LACHARTRE };

	EchelonForm (Context<Modular<Element>, HIPModule<Modular<Element> > > &ctx)
		: _ctx (ctx), _cpu (ctx.F), _fallback (_cpu) {}

	DenseMatrix<Element> &echelonize (DenseMatrix<Element> &A, bool reduced = false,
					  Method method = METHOD_ASYMPTOTICALLY_FAST_GJ)
	{
		if (!reduced || method == METHOD_FAUGERE_LACHARTRE) {
			_fallback.echelonize (A, reduced,
					      (typename EchelonForm<Modular<Element>, ZpModule<Element> >::Method) method);
			_rank_table[&A] = _fallback.rank (A);
			return A;
		}
		size_t rank = 0;
		std::vector<uint32_t> pivots (std::min (A.rowdim (), A.coldim ()) + 1);
		if (A.rowdim () && A.coldim ())
			_ctx.M.dev.check (mlz_zp_rref (_ctx.M.dev.ctx (), HIPElement<Element>::value (),
						       (uint64_t) _ctx.F._modulus, A.rowdim (), A.coldim (), &A[0][0],
						       A.disp (), &rank, &pivots[0], MLZ_ALGO_AUTO));
		_rank_table[&A] = rank;
		return A;
	}

	template <class Matrix>
	size_t rank (const Matrix &A) const { return _rank_table.find (&A)->second; }
};

} // namespace LELA

#endif // LELA_AMD_HIP_MODULE_H
