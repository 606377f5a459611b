// row-echelon-form: reads a matrix file, computes its row-echelon form on the
// GPU and writes it -- the command-line tool of LELA's util/row-echelon-form.C
// (options, defaults and messages as there, :201-312), built on the C ABI
// (include/mlz.h, include/mlz_io.h) only.
//
// Dispatch follows EchelonForm<Ring>::echelonize as LELA builds it without
// M4RI (lela/solutions/echelon-form.h:61-154):
//   dense  standard / afast: reduced -> mlz_zp_rref / mlz_gf2_rref, else
//                            Elimination's U (mlz_zp_echelonize, U mode)
//   dense  f4:               FaugereLachartre, always reduced (:139-144)
//   sparse standard:         reduced -> the reduced form; else Elimination with
//                            SparsePartialPivotStrategy (mlz_zp_echelonize_sparse)
//   sparse f4:               FaugereLachartre (reduced form only here)
//   sparse afast:            "Invalid method for choice of matrix"
// GF(2) non-reduced forms run over Z/2 on one 16-bit entry per bit.

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mlz.h"
#include "mlz_io.h"

namespace {

enum Method { METHOD_UNKNOWN, METHOD_STANDARD_GJ, METHOD_ASYMPTOTICALLY_FAST_GJ, METHOD_FAUGERE_LACHARTRE };
enum MatrixType { MATRIX_UNKNOWN, MATRIX_DENSE, MATRIX_SPARSE, MATRIX_HYBRID };

struct Options {
	bool reduced = false;
	const char *ring = "guess";
	std::string modulus = "65521";
	bool floating = false;
	const char *method = "f4";
	const char *in_format = "guess";
	const char *out_format = "guess";
	const char *type = "dense";
	int device = 0;
	const char *input = nullptr;
	const char *output = nullptr;
};

void help(const char *prog)
{
	printf("Usage: %s [options] <input-filename> <output-filename>\n\n"
	       "Where [options] are the following:\n"
	       "  -r        Compute the reduced row-echelon form (default OFF)\n"
	       "  -k K      Ring over which to compute ('guess', 'gf2', 'modular') (default guess)\n"
	       "  -p P      Modulus of ring, when ring is 'modular' (default 65521)\n"
	       "  -f        Compute using floating point, when ring is 'modular' (default OFF)\n"
	       "  -m M      Method to be used ('standard', 'afast', or 'f4') (default f4)\n"
	       "  -i I      Input file format ('guess', 'dumas', 'turner', 'maple', 'matlab', 'sage', 'pretty')\n"
	       "  -o O      Output file format ('guess', 'dumas', 'turner', 'maple', 'matlab', 'sage', 'pretty')\n"
	       "  -t T      Type to use for matrix ('dense', 'sparse', 'hybrid') (default dense)\n"
	       "  -d D      HIP device (default 0)\n"
	       "  -h or -?  Display this message\n",
	       prog);
}

int format_tag(const char *s)
{
	static const struct {
		const char *name;
		int tag;
	} t[] = {{"guess", MLZ_FORMAT_DETECT}, {"dumas", MLZ_FORMAT_DUMAS},   {"turner", MLZ_FORMAT_TURNER},
		 {"maple", MLZ_FORMAT_MAPLE},  {"matlab", MLZ_FORMAT_MATLAB}, {"sage", MLZ_FORMAT_SAGE},
		 {"pretty", MLZ_FORMAT_PRETTY}};
	for (auto &e : t)
		if (!strcmp(s, e.name))
			return e.tag;
	return MLZ_FORMAT_UNKNOWN;
}

// support.h's guess_format_tag: by the output file's extension.
int guess_format_tag(const char *filename)
{
	const char *ext = strrchr(filename, '.');
	if (!ext || ext[1] == '\0')
		return MLZ_FORMAT_UNKNOWN;
	++ext;
	if (!strcmp(ext, "m"))
		return MLZ_FORMAT_MATLAB;
	if (!strcmp(ext, "sage"))
		return MLZ_FORMAT_SAGE;
	return MLZ_FORMAT_UNKNOWN;
}

Method get_method(const char *s)
{
	if (!strcmp(s, "standard"))
		return METHOD_STANDARD_GJ;
	if (!strcmp(s, "afast"))
		return METHOD_ASYMPTOTICALLY_FAST_GJ;
	if (!strcmp(s, "f4"))
		return METHOD_FAUGERE_LACHARTRE;
	return METHOD_UNKNOWN;
}

MatrixType get_matrix_type(const char *s, bool gf2)
{
	if (!strcmp(s, "dense"))
		return MATRIX_DENSE;
	if (!strcmp(s, "sparse"))
		return MATRIX_SPARSE;
	if (gf2 && !strcmp(s, "hybrid"))
		return MATRIX_HYBRID;
	return MATRIX_UNKNOWN;
}

// support.C's parseArguments: single-letter switches, values in the next argument.
bool parse(int argc, char **argv, Options &o)
{
	for (int i = 1; i < argc; ++i) {
		const char *a = argv[i];
		if (a[0] == '-' && a[1] != '\0') {
			const char c = a[1];
			auto value = [&]() -> const char * {
				if (i + 1 >= argc) {
					fprintf(stderr, "ERROR: Missing value for %s\n", a);
					exit(1);
				}
				return argv[++i];
			};
			switch (c) {
			case 'h':
			case '?':
				help(argv[0]);
				exit(1);
			case 'v':
				break;
			case 'r':
				o.reduced = true;
				break;
			case 'f':
				o.floating = true;
				break;
			case 'k':
				o.ring = value();
				break;
			case 'p':
				o.modulus = value();
				break;
			case 'm':
				o.method = value();
				break;
			case 'i':
				o.in_format = value();
				break;
			case 'o':
				o.out_format = value();
				break;
			case 't':
				o.type = value();
				break;
			case 'd':
				o.device = atoi(value());
				break;
			default:
				fprintf(stderr, "ERROR: Bad argument %s\n", a);
				return true;
			}
		} else if (!o.input) {
			o.input = a;
		} else if (!o.output) {
			o.output = a;
		}
	}
	return true;
}

struct Ring {
	bool gf2;
	int elem;            // mlz_elem for Z/p
	uint64_t p;
	size_t esize;        // bytes per stored element
};

// The matrix read from the input, as host rows and as a dense array.
struct Input {
	size_t m = 0, n = 0;
	std::vector<uint64_t> row_ptr, val;
	std::vector<uint32_t> col;
};

int read_input(const Options &o, int fmt, const Ring &R, Input &in)
{
	mlz_file_matrix *h = nullptr;
	const int rc = mlz_matrix_read(o.input, fmt, R.gf2 ? MLZ_IO_GF2 : R.elem, R.p, 0, 0, &h);
	if (rc == MLZ_E_FILE) {
		fprintf(stderr, "Could not open input-file\n");
		return -1;
	}
	if (rc == MLZ_E_FORMAT) {
		fprintf(stderr, "Unable to determine format of input-file\n");
		return -1;
	}
	if (rc) {
		fprintf(stderr, "Error reading the input-file: %s\n", mlz_io_last_error());
		return -1;
	}
	size_t nnz = 0;
	mlz_matrix_shape(h, &in.m, &in.n, &nnz);
	in.row_ptr.resize(in.m + 1);
	in.col.resize(nnz);
	in.val.resize(nnz);
	mlz_matrix_get_csr(h, in.row_ptr.data(), in.col.data(), in.val.data());
	mlz_matrix_release(h);
	return 0;
}

// Dense array of the input in the ring's storage (Z/p) or one uint16 per bit (GF(2) over Z/2).
std::vector<unsigned char> densify(const Input &in, const Ring &R, size_t es)
{
	std::vector<unsigned char> A(in.m * in.n * es, 0);
	for (size_t i = 0; i < in.m; ++i)
		for (uint64_t e = in.row_ptr[i]; e < in.row_ptr[i + 1]; ++e) {
			const size_t at = i * in.n + in.col[e];
			const uint64_t v = in.val[e];
			if (es == 2)
				((uint16_t *) A.data())[at] = (uint16_t) v;
			else if (es == 4)
				((uint32_t *) A.data())[at] = (uint32_t) v;
			else
				((double *) A.data())[at] = (double) (2 * v > R.p ? (long long) v - (long long) R.p : (long long) v);
		}
	return A;
}

std::vector<uint64_t> pack_bits(const Input &in)
{
	const size_t nw = (in.n + 63) / 64;
	std::vector<uint64_t> W(in.m * std::max<size_t>(nw, 1), 0);
	for (size_t i = 0; i < in.m; ++i)
		for (uint64_t e = in.row_ptr[i]; e < in.row_ptr[i + 1]; ++e)
			W[i * std::max<size_t>(nw, 1) + in.col[e] / 64] |= 1ull << (in.col[e] % 64);
	return W;
}

int report(mlz_ctx *ctx, int rc, size_t bad)
{
	if (rc == MLZ_OK)
		return 0;
	if (rc == MLZ_E_WRONG_FORM)
		fprintf(stderr, "Bad input-matrix for Faugere-Lachartre at row-index %zu\n", bad);
	else if (rc == MLZ_E_NOT_INVERTIBLE)
		fprintf(stderr, "Could not invert pivot-element in the ring\n");
	else
		fprintf(stderr, "Error: %s\n", mlz_last_error(ctx));
	return -1;
}

int run(const Options &o, const Ring &R, Method method, MatrixType type, int in_fmt, int out_fmt)
{
	Input in;
	if (read_input(o, in_fmt, R, in))
		return -1;
	if (type != MATRIX_DENSE && method == METHOD_ASYMPTOTICALLY_FAST_GJ) {
		fprintf(stderr, "Invalid method for choice of matrix\n");
		return -1;
	}
	mlz_ctx *ctx = nullptr;
	if (mlz_ctx_create(&ctx, o.device, nullptr) != MLZ_OK) {
		fprintf(stderr, "Could not create a GPU context on device %d\n", o.device);
		return -1;
	}
	const bool fl = method == METHOD_FAUGERE_LACHARTRE;
	// the dense FL specialisation ignores `reduced` (echelon-form.h:139-144)
	const bool reduced = o.reduced || (fl && type == MATRIX_DENSE);
	if (fl && !reduced) {
		fprintf(stderr, "The non-reduced Faugere-Lachartre form is not available on the GPU\n");
		mlz_ctx_destroy(ctx);
		return -1;
	}
	int rc = MLZ_OK;
	size_t rank = 0, bad = 0;
	const int elem_out = R.gf2 ? MLZ_IO_GF2 : R.elem;
	const size_t m = in.m, n = in.n;
	if (R.gf2 && (reduced || fl)) {
		std::vector<uint64_t> W = pack_bits(in);
		const size_t nw = std::max<size_t>((n + 63) / 64, 1);
		if (fl) {
			// hybrid rows: the nonzero words of each row
			std::vector<uint64_t> rp(m + 1, 0), words;
			std::vector<uint32_t> blk;
			for (size_t i = 0; i < m; ++i) {
				for (size_t w = 0; w < nw; ++w)
					if (W[i * nw + w]) {
						blk.push_back((uint32_t) w);
						words.push_back(W[i * nw + w]);
					}
				rp[i + 1] = blk.size();
			}
			std::fill(W.begin(), W.end(), 0);
			rc = mlz_gf2_fl_echelonize(ctx, m, n, rp.data(), blk.data(), words.data(), W.data(), nw, &rank,
						   nullptr, &bad);
		} else {
			rc = mlz_gf2_rref(ctx, m, n, W.data(), nw, &rank, nullptr);
		}
		if (report(ctx, rc, bad) == 0)
			rc = mlz_matrix_write_dense(o.output, out_fmt, elem_out, 2, m, n, W.data(), nw);
		else
			rc = -1;
	} else if (fl) {
		std::vector<unsigned char> R0(m * n * R.esize, 0);
		std::vector<unsigned char> vals(in.val.size() * R.esize);
		for (size_t e = 0; e < in.val.size(); ++e) {
			const uint64_t v = in.val[e];
			if (R.esize == 2)
				((uint16_t *) vals.data())[e] = (uint16_t) v;
			else if (R.esize == 4)
				((uint32_t *) vals.data())[e] = (uint32_t) v;
			else
				((double *) vals.data())[e] = (double) (2 * v > R.p ? (long long) v - (long long) R.p : (long long) v);
		}
		rc = mlz_zp_fl_echelonize(ctx, (mlz_elem) R.elem, R.p, m, n, in.row_ptr.data(), in.col.data(), vals.data(),
					  R0.data(), n, &rank, nullptr, nullptr, &bad, MLZ_ALGO_AUTO);
		if (report(ctx, rc, bad) == 0)
			rc = mlz_matrix_write_dense(o.output, out_fmt, elem_out, R.p, m, n, R0.data(), n);
		else
			rc = -1;
	} else {
		// Z/p, or GF(2) as Z/2 with one 16-bit entry per bit
		const int elem = R.gf2 ? MLZ_U16 : R.elem;
		const uint64_t p = R.gf2 ? 2 : R.p;
		const size_t es = R.gf2 ? 2 : R.esize;
		std::vector<unsigned char> A = densify(in, R, es);
		uint64_t det = 0;
		if (reduced)
			rc = mlz_zp_rref(ctx, (mlz_elem) elem, p, m, n, A.data(), n, &rank, nullptr, MLZ_ALGO_AUTO);
		else if (type == MATRIX_DENSE)
			rc = mlz_zp_echelonize(ctx, (mlz_elem) elem, p, m, n, A.data(), n, MLZ_ECH_U, &rank, &det, nullptr,
					       nullptr, MLZ_ALGO_AUTO);
		else
			rc = mlz_zp_echelonize_sparse(ctx, (mlz_elem) elem, p, m, n, A.data(), n, MLZ_ECH_U,
						      type == MATRIX_HYBRID, &rank, &det, nullptr, nullptr);
		if (report(ctx, rc, bad) == 0) {
			if (R.gf2) {
				Input out;
				out.m = m;
				out.n = n;
				out.row_ptr.assign(m + 1, 0);
				for (size_t i = 0; i < m; ++i) {
					for (size_t j = 0; j < n; ++j)
						if (((uint16_t *) A.data())[i * n + j]) {
							out.col.push_back((uint32_t) j);
							out.val.push_back(1);
						}
					out.row_ptr[i + 1] = out.col.size();
				}
				std::vector<uint64_t> W = pack_bits(out);
				rc = mlz_matrix_write_dense(o.output, out_fmt, MLZ_IO_GF2, 2, m, n, W.data(),
							    std::max<size_t>((n + 63) / 64, 1));
			} else {
				rc = mlz_matrix_write_dense(o.output, out_fmt, elem, p, m, n, A.data(), n);
			}
		} else {
			rc = -1;
		}
	}
	mlz_ctx_destroy(ctx);
	if (rc > 0) {
		fprintf(stderr, "Could not write the output-file: %s\n", mlz_io_last_error());
		return -1;
	}
	return rc;
}

}  // namespace

int main(int argc, char **argv)
{
	Options o;
	parse(argc, argv, o);
	if (!o.input || !o.output) {
		help(argv[0]);
		return -1;
	}
	const int in_fmt = format_tag(o.in_format);
	int out_fmt = format_tag(o.out_format);
	if (in_fmt == MLZ_FORMAT_UNKNOWN) {
		fprintf(stderr, "Invalid input-file-format (use 'guess', 'dumas', 'turner', 'maple', 'matlab', 'sage', "
				"or 'pretty')\n");
		return -1;
	}
	if (out_fmt == MLZ_FORMAT_UNKNOWN) {
		fprintf(stderr, "Invalid output-file-format (use 'guess', 'dumas', 'turner', 'maple', 'matlab', 'sage', "
				"or 'pretty')\n");
		return -1;
	}
	if (out_fmt == MLZ_FORMAT_DETECT) {
		out_fmt = guess_format_tag(o.output);
		if (out_fmt == MLZ_FORMAT_UNKNOWN) {
			fprintf(stderr, "Could not guess output file-format\n");
			return -1;
		}
	}
	bool gf2;
	if (!strcmp(o.ring, "gf2"))
		gf2 = true;
	else if (!strcmp(o.ring, "modular") || !strcmp(o.ring, "guess"))
		gf2 = false;
	else {
		fprintf(stderr, "Invalid ring (use 'gf2' or 'modular')\n");
		return -1;
	}
	const Method method = get_method(o.method);
	if (method == METHOD_UNKNOWN) {
		fprintf(stderr, "Invalid method (use 'standard', 'afast', or 'f4')\n");
		return -1;
	}
	const MatrixType type = get_matrix_type(o.type, gf2);
	if (type == MATRIX_UNKNOWN) {
		fprintf(stderr, gf2 ? "Invalid matrix-type (use 'dense', 'sparse', or 'hybrid')\n"
				    : "Invalid matrix-type (use 'dense' or 'sparse')\n");
		return -1;
	}
	Ring R{gf2, 0, 2, 0};
	if (!gf2) {
		char *end = nullptr;
		errno = 0;
		const unsigned long long p = strtoull(o.modulus.c_str(), &end, 10);
		if (errno || !end || *end || p < 2) {
			fprintf(stderr, "Invalid modulus %s\n", o.modulus.c_str());
			return -1;
		}
		R.p = p;
		// the ring LELA picks (row-echelon-form.C:290-305): Modular<float> below
		// 2^7 and Modular<double> below 2^21 with -f (both symmetric, written
		// alike), else Modular<uint8/uint16> below 2^16 and Modular<uint32>
		// below 2^32; larger moduli take Modular<integer>, which has no GPU path.
		if (o.floating && p < (1ull << 21)) {
			R.elem = MLZ_F64_SYM;
			R.esize = 8;
		} else if (!o.floating && p < (1ull << 16)) {
			R.elem = MLZ_U16;
			R.esize = 2;
		} else if (!o.floating && p < (1ull << 32)) {
			R.elem = MLZ_U32;
			R.esize = 4;
		} else {
			fprintf(stderr, "Modulus %s needs Modular<integer>, which has no GPU path\n", o.modulus.c_str());
			return -1;
		}
	}
	return run(o, R, method, type, in_fmt, out_fmt);
}
