// Matrix files: LELA's text formats read into and written from host rows.
// Host code only (no GPU), compiled into liblela_amd.so beside the kernels.
//
// Formats and their text follow lela/matrix/io.tcc (MatrixReader :33-413,
// MatrixWriter :443-651); the element syntax follows Ring::read / write
// (lela/ring/modular.h:399-417, lela/ring/gf2.h:105-121).

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mlz.h"
#include "mlz_io.h"

namespace {

thread_local std::string g_err = "no error";

int fail(int rc, const std::string &msg)
{
	g_err = msg;
	return rc;
}

// The ring of the entries: Modular<T>(p) in one of three storages, or GF2.
struct Ring {
	int elem;
	uint64_t p;

	bool gf2() const { return elem == MLZ_IO_GF2; }

	// Ring::read's reduction of a signed 64-bit value: Modular<T> reduces mod
	// p into the canonical class (ModularTraits<T>::reduce, modular.h:62-63,
	// 123-124, 167-168, 236-237); GF2 maps any nonzero value to 1 (gf2.h:110-111).
	uint64_t reduce(int64_t v) const
	{
		if (gf2())
			return v != 0;
		const int64_t pi = (int64_t) p;
		if (p > (uint64_t) INT64_MAX) {
			uint64_t u = v < 0 ? (uint64_t) 0 - (uint64_t) (-(v + 1)) - 1 : (uint64_t) v;
			return u % p;
		}
		int64_t r = v % pi;
		if (r < 0)
			r += pi;
		return (uint64_t) r;
	}

	// Ring::elementWidth (modular.h:75-80, 183-184, 307-308; gf2.h:120-121).
	unsigned width() const
	{
		if (gf2())
			return 1;
		if (elem == MLZ_F64_SYM)
			return (unsigned) ceil(log((double) p / 2) / M_LN10) + 1;
		return (unsigned) ceil(log((double) p) / M_LN10);
	}

	// The printed value of a canonical residue: symmetric for Modular<double>
	// (ModularTraits<double>::write prints (long long) x, modular.h:305-306).
	long long printed(uint64_t v) const
	{
		if (elem == MLZ_F64_SYM && 2 * v > p)
			return (long long) v - (long long) p;
		return (long long) v;
	}
};

bool valid_ring(int elem, uint64_t p, std::string &why)
{
	if (elem == MLZ_IO_GF2)
		return true;
	if (elem != MLZ_U16 && elem != MLZ_U32 && elem != MLZ_F64_SYM) {
		why = "elem must be MLZ_U16, MLZ_U32, MLZ_F64_SYM or MLZ_IO_GF2";
		return false;
	}
	const uint64_t lim = elem == MLZ_U16 ? (1ull << 16) : elem == MLZ_U32 ? (1ull << 32) : (1ull << 26);
	if (p < 2 || p >= lim) {
		why = "modulus out of range for the storage";
		return false;
	}
	return true;
}

// ---------------------------------------------------------------------------
// Input: a cursor over the file's bytes with the std::istream operations the
// readers use (whitespace skipping, numeric extraction, peek).

struct Cursor {
	const std::string &s;
	size_t at = 0;

	explicit Cursor(const std::string &text) : s(text) {}

	bool eof() const { return at >= s.size(); }
	int peek() const { return eof() ? -1 : (unsigned char) s[at]; }
	void skip_ws()
	{
		while (!eof() && isspace((unsigned char) s[at]))
			++at;
	}
	// `is >> c` for a char: skips whitespace, -1 at the end.
	int next_char()
	{
		skip_ws();
		return eof() ? -1 : (unsigned char) s[at++];
	}
	size_t line() const
	{
		size_t l = 1;
		for (size_t i = 0; i < at && i < s.size(); ++i)
			l += s[i] == '\n';
		return l;
	}

	// `is >> size_t`: optional sign, decimal digits; a '-' wraps modulo 2^64
	// as strtoull does ("-1" is SIZE_MAX, the readers' end marker).
	bool read_index(uint64_t &v)
	{
		skip_ws();
		size_t i = at;
		bool neg = false;
		if (i < s.size() && (s[i] == '-' || s[i] == '+'))
			neg = s[i++] == '-';
		if (i >= s.size() || !isdigit((unsigned char) s[i]))
			return false;
		uint64_t x = 0;
		for (; i < s.size() && isdigit((unsigned char) s[i]); ++i) {
			const uint64_t d = (uint64_t) (s[i] - '0');
			if (x > (UINT64_MAX - d) / 10)
				return false;
			x = x * 10 + d;
		}
		at = i;
		v = neg ? (uint64_t) 0 - x : x;
		return true;
	}

	// `is >> int` as GF2::read does (gf2.h:110-111): a value outside int sets
	// the stream's failbit (the extracted value, INT_MAX or INT_MIN, is
	// nonzero); *failed reports it.
	bool read_int(int64_t &v, bool &failed)
	{
		const size_t start = at;
		if (!read_integer(v))
			return false;
		// re-read the digits to see whether the value fits in int
		size_t i = start;
		while (i < s.size() && isspace((unsigned char) s[i]))
			++i;
		const bool neg = s[i] == '-';
		if (s[i] == '-' || s[i] == '+')
			++i;
		uint64_t x = 0;
		failed = false;
		for (; i < at; ++i) {
			x = x * 10 + (uint64_t) (s[i] - '0');
			if (x > (uint64_t) INT32_MAX + (neg ? 1 : 0)) {
				failed = true;
				break;
			}
		}
		if (failed)
			v = neg ? INT32_MIN : INT32_MAX;
		return true;
	}

	// `is >> integer` (GMP, decimal) followed by mpz_get_si: an integer of any
	// length; the result is the low 63 bits of |x| with x's sign, as
	// mpz_get_si returns them (exact whenever x fits in a signed 64-bit word).
	bool read_integer(int64_t &v)
	{
		skip_ws();
		size_t i = at;
		bool neg = false;
		if (i < s.size() && (s[i] == '-' || s[i] == '+'))
			neg = s[i++] == '-';
		if (i >= s.size() || !isdigit((unsigned char) s[i]))
			return false;
		uint64_t low = 0;             // |x| mod 2^64
		for (; i < s.size() && isdigit((unsigned char) s[i]); ++i)
			low = low * 10 + (uint64_t) (s[i] - '0');
		at = i;
		if (!neg)
			v = (int64_t) (low & (uint64_t) INT64_MAX);
		else
			v = low == 0 ? 0 : -1 - (int64_t) ((low - 1) & (uint64_t) INT64_MAX);
		return true;
	}
};

// Ring::read (modular.h:405-413, gf2.h:110-111): the canonical residue of
// the next entry; *failed is set when the stream would fail after it (GF(2)
// entries outside int), which ends LELA's read loops.
bool read_entry(Cursor &c, const Ring &R, uint64_t &v, bool &failed)
{
	int64_t x;
	failed = false;
	if (R.gf2() ? !c.read_int(x, failed) : !c.read_integer(x))
		return false;
	v = R.reduce(x);
	return true;
}

// Host entries being read, with setEntry semantics: nonzero values only, a
// repeated (i, j) keeps the last one (resolved when the rows are assembled).
struct Rows {
	size_t m = 0, n = 0;
	bool fixed = false;           // dimensions given (Dumas header or caller)
	struct Entry {
		uint32_t i, j;
		uint64_t x;
	};
	std::vector<Entry> ent;       // in input order

	void size_to(size_t r, size_t c)
	{
		m = r;
		n = c;
		fixed = true;
	}

	// Extends an inferred shape to cover (i, j); false outside a fixed shape.
	bool touch(uint64_t i, uint64_t j)
	{
		if (fixed)
			return i < m && j < n;
		if (i >= UINT32_MAX || j >= UINT32_MAX)
			return false;
		if (i + 1 > m)
			m = i + 1;
		if (j + 1 > n)
			n = j + 1;
		return true;
	}

	// A.setEntry (i, j, x) for a nonzero x.
	bool set(uint64_t i, uint64_t j, uint64_t x)
	{
		if (!touch(i, j))
			return false;
		ent.push_back({(uint32_t) i, (uint32_t) j, x});
		return true;
	}
};

}  // namespace

struct mlz_file_matrix {
	size_t m, n;
	std::vector<uint64_t> row_ptr;
	std::vector<uint32_t> col;
	std::vector<uint64_t> val;
};

namespace {

int slurp(const char *path, std::string &out)
{
	if (!path)
		return fail(MLZ_E_ARG, "null path");
	FILE *f = fopen(path, "rb");
	if (!f)
		return fail(MLZ_E_FILE, std::string("cannot open ") + path + ": " + strerror(errno));
	char buf[1 << 16];
	size_t got;
	out.clear();
	while ((got = fread(buf, 1, sizeof buf, f)) > 0)
		out.append(buf, got);
	const bool bad = ferror(f);
	fclose(f);
	if (bad)
		return fail(MLZ_E_FILE, std::string("read error on ") + path);
	return MLZ_OK;
}

// Classification of the first non-blank line by the shapes the writers emit.
int detect(const std::string &text)
{
	size_t b = 0;
	while (b < text.size() && isspace((unsigned char) text[b]))
		++b;
	size_t e = text.find('\n', b);
	std::string line = text.substr(b, e == std::string::npos ? std::string::npos : e - b);
	while (!line.empty() && isspace((unsigned char) line.back()))
		line.pop_back();
	if (line.empty())
		return MLZ_FORMAT_UNKNOWN;

	// whitespace-separated tokens of the line
	std::vector<std::string> tok;
	for (size_t i = 0; i < line.size();) {
		while (i < line.size() && isspace((unsigned char) line[i]))
			++i;
		size_t j = i;
		while (j < line.size() && !isspace((unsigned char) line[j]))
			++j;
		if (j > i)
			tok.push_back(line.substr(i, j - i));
		i = j;
	}
	auto is_uint = [](const std::string &t) {
		if (t.empty())
			return false;
		for (char c : t)
			if (!isdigit((unsigned char) c))
				return false;
		return true;
	};
	auto is_int = [&](const std::string &t) {
		return is_uint(t) || (t.size() > 1 && (t[0] == '-' || t[0] == '+') && is_uint(t.substr(1)));
	};
	if (tok.size() == 3 && is_uint(tok[0]) && is_uint(tok[1]) && tok[2] == "M")
		return MLZ_FORMAT_DUMAS;
	if (tok.size() == 3 && is_uint(tok[0]) && is_uint(tok[1]) && is_int(tok[2]))
		return MLZ_FORMAT_TURNER;
	if (line.compare(0, 7, "matrix(") == 0)
		return MLZ_FORMAT_SAGE;
	if (line[0] != '[')
		return MLZ_FORMAT_UNKNOWN;
	size_t k = 1;
	while (k < line.size() && isspace((unsigned char) line[k]))
		++k;
	if (k < line.size() && line[k] == '[')
		return MLZ_FORMAT_MAPLE;
	if (line == "[]" || line.find(',') != std::string::npos || line.find(';') != std::string::npos)
		return MLZ_FORMAT_MATLAB;
	return MLZ_FORMAT_PRETTY;
}

// MatrixReader::readTurner (io.tcc:191-229): lines "i j x", 0-based, until a
// row index of -1 or the end of the input.
int read_turner(const std::string &text, const Ring &R, Rows &A)
{
	size_t pos = 0;
	while (pos < text.size()) {
		size_t e = text.find('\n', pos);
		if (e == std::string::npos)
			e = text.size();
		const std::string line = text.substr(pos, e - pos);
		pos = e + 1;
		Cursor c(line);
		uint64_t i, j, v;
		bool failed;
		if (!c.read_index(i)) {
			c.skip_ws();
			if (c.eof())
				continue;          // blank line: LELA reads zeros and stores nothing
			return fail(MLZ_E_INPUT, "Turner: bad row index in line '" + line + "'");
		}
		if (i == UINT64_MAX)
			break;
		if (!c.read_index(j) || !read_entry(c, R, v, failed))
			return fail(MLZ_E_INPUT, "Turner: expected 'i j x' in line '" + line + "'");
		if (!(v ? A.set(i, j, v) : A.touch(i, j)))
			return fail(MLZ_E_INPUT, "Turner: entry (" + std::to_string(i) + ", " + std::to_string(j) +
					  ") outside the matrix");
		if (failed)
			break;                 // the stream has failed: LELA's loop ends
	}
	return MLZ_OK;
}

// MatrixReader::readDumas (io.tcc:231-269): "m n M", then 1-based "i j x"
// until a row index of 0 (or -1), whose "j x" is consumed too.
int read_dumas(const std::string &text, const Ring &R, Rows &A)
{
	Cursor c(text);
	uint64_t m, n;
	if (!c.read_index(m) || !c.read_index(n) || c.next_char() != 'M')
		return fail(MLZ_E_INPUT, "Dumas: the header must be 'm n M'");
	if (m > (1ull << 40) || n > UINT32_MAX)
		return fail(MLZ_E_INPUT, "Dumas: dimensions too large");
	A.size_to(m, n);
	uint64_t i, j, v;
	bool failed;
	while (c.read_index(i)) {
		if (i == 0 || i == UINT64_MAX) {
			c.read_index(j);
			read_entry(c, R, v, failed);
			break;
		}
		if (!c.read_index(j))
			return fail(MLZ_E_INPUT, "Dumas: missing column index near line " + std::to_string(c.line()));
		if (i > m || j > n || j == 0)
			return fail(MLZ_E_INPUT, "Dumas: entry (" + std::to_string(i) + ", " + std::to_string(j) +
					  ") outside the " + std::to_string(m) + " x " + std::to_string(n) + " matrix");
		if (!read_entry(c, R, v, failed))
			return fail(MLZ_E_INPUT, "Dumas: missing value near line " + std::to_string(c.line()));
		if (v)
			A.set(i - 1, j - 1, v);
		if (failed)
			break;                 // the stream has failed: LELA's loop ends
	}
	return MLZ_OK;
}

// MatrixReader::readMatlab (io.tcc:285-337): '[', entries separated by ','
// within a row and ';' between rows, ']'.  An entry must start with a digit
// or '-' (so "[]" is invalid input, as in LELA).
int read_matlab(const std::string &text, const Ring &R, Rows &A)
{
	Cursor c(text);
	if (c.next_char() != '[')
		return fail(MLZ_E_INPUT, "Matlab: expected '['");
	uint64_t i = 0, j = 0;
	for (;;) {
		const int ch = c.next_char();
		if (!(ch >= 0 && (isdigit(ch) || ch == '-')))
			return fail(MLZ_E_INPUT, "Matlab: expected an entry near line " + std::to_string(c.line()));
		--c.at;
		uint64_t v;
		bool failed;
		if (!read_entry(c, R, v, failed))
			return fail(MLZ_E_INPUT, "Matlab: bad entry near line " + std::to_string(c.line()));
		if (!(v ? A.set(i, j, v) : A.touch(i, j)))
			return fail(MLZ_E_INPUT, "Matlab: entry outside the matrix");
		if (failed)
			return fail(MLZ_E_INPUT, "Matlab: entry out of range for GF(2)'s int read");
		const int sep = c.next_char();
		if (sep == ';') {
			++i;
			j = 0;
		} else if (sep == ']') {
			break;
		} else if (sep == ',') {
			++j;
		} else {
			return fail(MLZ_E_INPUT, "Matlab: expected ',', ';' or ']' near line " + std::to_string(c.line()));
		}
	}
	if (!A.fixed && i + 1 > A.m)
		A.m = i + 1;
	return MLZ_OK;
}

// MatrixReader::readPretty (io.tcc:339-413): one '[' ... ']' per row,
// entries separated by whitespace, '.' for a zero entry.
int read_pretty(const std::string &text, const Ring &R, Rows &A)
{
	Cursor c(text);
	uint64_t i = 0;
	for (;;) {
		c.skip_ws();
		if (c.eof())
			break;
		if (c.next_char() != '[')
			return fail(MLZ_E_INPUT, "pretty: expected '[' at line " + std::to_string(c.line()));
		uint64_t j = 0;
		for (;;) {
			c.skip_ws();
			if (c.eof())
				break;
			const int ch = c.peek();
			if (ch == ']') {
				++c.at;
				break;
			}
			if (ch == '.') {
				++c.at;
				if (!A.touch(i, j))
					return fail(MLZ_E_INPUT, "pretty: entry outside the matrix");
				++j;
				continue;
			}
			uint64_t v;
			bool failed;
			if (!read_entry(c, R, v, failed) || failed)
				return fail(MLZ_E_INPUT, "pretty: bad entry at line " + std::to_string(c.line()));
			if (!(v ? A.set(i, j, v) : A.touch(i, j)))
				return fail(MLZ_E_INPUT, "pretty: entry outside the matrix");
			++j;
		}
		if (!A.fixed && j == 0 && i + 1 > A.m)
			A.m = i + 1;
		++i;
	}
	return MLZ_OK;
}

// ---------------------------------------------------------------------------
// Output: a buffered FILE* writer and a row source over either input form.

struct Out {
	FILE *f;
	std::string buf;
	bool bad = false;

	explicit Out(FILE *file) : f(file) { buf.reserve(1 << 20); }
	void flush()
	{
		if (!buf.empty() && fwrite(buf.data(), 1, buf.size(), f) != buf.size())
			bad = true;
		buf.clear();
	}
	Out &operator<<(const char *s)
	{
		buf += s;
		if (buf.size() >= (1 << 20))
			flush();
		return *this;
	}
	Out &operator<<(const std::string &s) { return *this << s.c_str(); }
	Out &num(long long v) { return *this << std::to_string(v); }
	Out &unum(unsigned long long v) { return *this << std::to_string(v); }
	// os.width (w) << v: right-aligned in w columns.
	Out &padded(long long v, unsigned w)
	{
		std::string t = std::to_string(v);
		if (t.size() < w)
			buf.append(w - t.size(), ' ');
		return *this << t;
	}
};

// Row i as canonical (column, value) pairs with nonzero values, increasing column.
struct Source {
	const Ring &R;
	size_t m, n;
	// CSR form
	const uint64_t *rp = nullptr;
	const uint32_t *ci = nullptr;
	const uint64_t *cv = nullptr;
	// dense form
	const void *A = nullptr;
	size_t lda = 0;

	Source(const Ring &ring, size_t rows, size_t cols) : R(ring), m(rows), n(cols) {}

	uint64_t dense_at(size_t i, size_t j) const
	{
		if (R.gf2())
			return (((const uint64_t *) A)[i * lda + j / 64] >> (j % 64)) & 1;
		switch (R.elem) {
		case MLZ_U16:
			return ((const uint16_t *) A)[i * lda + j];
		case MLZ_U32:
			return ((const uint32_t *) A)[i * lda + j];
		default: {
			const double d = ((const double *) A)[i * lda + j];
			const long long v = (long long) d;
			return v < 0 ? (uint64_t) (v + (long long) R.p) : (uint64_t) v;
		}
		}
	}

	void row(size_t i, std::vector<std::pair<uint32_t, uint64_t>> &out) const
	{
		out.clear();
		if (rp) {
			for (uint64_t e = rp[i]; e < rp[i + 1]; ++e) {
				const uint64_t v = R.gf2() ? cv[e] != 0 : cv[e] % R.p;
				if (v)
					out.emplace_back(ci[e], v);
			}
		} else {
			for (size_t j = 0; j < n; ++j) {
				const uint64_t v = dense_at(i, j);
				if (v)
					out.emplace_back((uint32_t) j, v);
			}
		}
	}
};

// writeTurner / writeOneBased / writeDumas (io.tcc:484-522).
void write_triples(Out &o, const Source &S, int format)
{
	const long long base = format == MLZ_FORMAT_TURNER ? 0 : 1;
	if (format == MLZ_FORMAT_DUMAS) {
		o.unum(S.m) << " ";
		o.unum(S.n) << " M\n";
	}
	std::vector<std::pair<uint32_t, uint64_t>> r;
	for (size_t i = 0; i < S.m; ++i) {
		S.row(i, r);
		for (auto &e : r) {
			o.num((long long) i + base) << " ";
			o.num((long long) e.first + base) << " ";
			o.num(S.R.printed(e.second)) << "\n";
		}
	}
	if (format == MLZ_FORMAT_TURNER)
		o << "-1\n";
	else if (format == MLZ_FORMAT_DUMAS)
		o << "0 0 0\n";
}

// Entries of row i with zeros in place, as getEntry gives them.
void dense_row(const Source &S, size_t i, std::vector<std::pair<uint32_t, uint64_t>> &tmp,
	       std::vector<uint64_t> &full)
{
	S.row(i, tmp);
	full.assign(S.n, 0);
	for (auto &e : tmp)
		if (e.first < S.n)
			full[e.first] = e.second;
}

// writeMaple (io.tcc:524-560), writeMatlab (:562-595), writeSage (:597-635).
void write_bracketed(Out &o, const Source &S, int format)
{
	if (S.m == 0) {
		o << (format == MLZ_FORMAT_SAGE ? "matrix([])\n" : "[]");
		return;
	}
	o << (format == MLZ_FORMAT_SAGE ? "matrix([ " : "[ ");
	std::vector<std::pair<uint32_t, uint64_t>> tmp;
	std::vector<uint64_t> full;
	for (size_t i = 0; i < S.m; ++i) {
		dense_row(S, i, tmp, full);
		if (format == MLZ_FORMAT_MAPLE)
			o << "[";
		else if (format == MLZ_FORMAT_SAGE)
			o << (i == 0 ? "[ " : "         [ ");
		for (size_t j = 0; j < S.n; ++j) {
			o.num(S.R.printed(full[j]));
			if (j + 1 < S.n)
				o << ", ";
		}
		const bool last = i + 1 == S.m;
		if (format == MLZ_FORMAT_MAPLE)
			o << (last ? "] ]\n" : "], ");
		else if (format == MLZ_FORMAT_MATLAB)
			o << (last ? "" : "; ");
		else
			o << (last ? " ] ])\n" : " ],\n");
	}
	if (format == MLZ_FORMAT_MATLAB)
		o << "]\n";
}

// writePretty (io.tcc:637-670).
void write_pretty(Out &o, const Source &S)
{
	if (S.m == 0 || S.n == 0) {
		o << "(empty ";
		o.unum(S.m) << "x";
		o.unum(S.n) << " matrix)\n";
		return;
	}
	const unsigned w = S.R.width();
	const std::string dot = std::string(w > 0 ? w - 1 : 0, ' ') + ". ";
	std::vector<std::pair<uint32_t, uint64_t>> tmp;
	std::vector<uint64_t> full;
	for (size_t i = 0; i < S.m; ++i) {
		dense_row(S, i, tmp, full);
		o << "  [ ";
		for (size_t j = 0; j < S.n; ++j) {
			if (full[j] == 0) {
				o << dot;
			} else {
				o.padded(S.R.printed(full[j]), w);
				o << " ";
			}
		}
		o << "]\n";
	}
}

int write_any(const char *path, int format, const Source &S)
{
	if (!path)
		return fail(MLZ_E_ARG, "null path");
	if (format < MLZ_FORMAT_TURNER || format > MLZ_FORMAT_PRETTY)
		return fail(MLZ_E_ARG, "bad output format");
	FILE *f = fopen(path, "wb");
	if (!f)
		return fail(MLZ_E_FILE, std::string("cannot open ") + path + ": " + strerror(errno));
	Out o(f);
	switch (format) {
	case MLZ_FORMAT_TURNER:
	case MLZ_FORMAT_ONE_BASED:
	case MLZ_FORMAT_DUMAS:
		write_triples(o, S, format);
		break;
	case MLZ_FORMAT_PRETTY:
		write_pretty(o, S);
		break;
	default:
		write_bracketed(o, S, format);
	}
	o.flush();
	const bool bad = o.bad || fclose(f) != 0;
	if (bad)
		return fail(MLZ_E_FILE, std::string("write error on ") + path);
	return MLZ_OK;
}

}  // namespace

extern "C" {

int mlz_matrix_detect_format(const char *path, int *format)
{
	if (!format)
		return fail(MLZ_E_ARG, "null format");
	std::string text;
	int rc = slurp(path, text);
	if (rc)
		return rc;
	*format = detect(text);
	return MLZ_OK;
}

int mlz_matrix_read(const char *path, int format, int elem, uint64_t p, size_t rows, size_t cols,
		    mlz_file_matrix **out)
{
	if (!out)
		return fail(MLZ_E_ARG, "null output");
	*out = nullptr;
	std::string why;
	if (!valid_ring(elem, p, why))
		return fail(MLZ_E_ARG, why);
	const Ring R{elem, elem == MLZ_IO_GF2 ? 2 : p};
	std::string text;
	int rc = slurp(path, text);
	if (rc)
		return rc;
	if (format == MLZ_FORMAT_DETECT)
		format = detect(text);
	Rows A;
	if ((rows || cols) && format != MLZ_FORMAT_DUMAS)
		A.size_to(rows, cols);
	switch (format) {
	case MLZ_FORMAT_TURNER:
		rc = read_turner(text, R, A);
		break;
	case MLZ_FORMAT_DUMAS:
		rc = read_dumas(text, R, A);
		break;
	case MLZ_FORMAT_MATLAB:
		rc = read_matlab(text, R, A);
		break;
	case MLZ_FORMAT_PRETTY:
		rc = read_pretty(text, R, A);
		break;
	case MLZ_FORMAT_ONE_BASED:
	case MLZ_FORMAT_MAPLE:
	case MLZ_FORMAT_SAGE:
		return fail(MLZ_E_UNSUPPORTED, "reading this format is not implemented (as in LELA)");
	case MLZ_FORMAT_UNKNOWN:
		return fail(MLZ_E_FORMAT, "unable to determine the format of the input file");
	default:
		return fail(MLZ_E_ARG, "bad input format");
	}
	if (rc)
		return rc;
	if (A.m >= UINT32_MAX || A.n > UINT32_MAX)
		return fail(MLZ_E_INPUT, "matrix dimensions too large");
	// rows in order, columns increasing; of a repeated (i, j) the last wins
	std::stable_sort(A.ent.begin(), A.ent.end(), [](const Rows::Entry &a, const Rows::Entry &b) {
		return a.i != b.i ? a.i < b.i : a.j < b.j;
	});
	auto *M = new mlz_file_matrix;
	M->m = A.m;
	M->n = A.n;
	M->row_ptr.assign(A.m + 1, 0);
	M->col.reserve(A.ent.size());
	M->val.reserve(A.ent.size());
	for (size_t e = 0; e < A.ent.size(); ++e) {
		if (e + 1 < A.ent.size() && A.ent[e + 1].i == A.ent[e].i && A.ent[e + 1].j == A.ent[e].j)
			continue;
		M->col.push_back(A.ent[e].j);
		M->val.push_back(A.ent[e].x);
		++M->row_ptr[A.ent[e].i + 1];
	}
	for (size_t i = 0; i < A.m; ++i)
		M->row_ptr[i + 1] += M->row_ptr[i];
	*out = M;
	return MLZ_OK;
}

int mlz_matrix_shape(const mlz_file_matrix *A, size_t *m, size_t *n, size_t *nnz)
{
	if (!A)
		return fail(MLZ_E_ARG, "null matrix");
	if (m)
		*m = A->m;
	if (n)
		*n = A->n;
	if (nnz)
		*nnz = A->col.size();
	return MLZ_OK;
}

int mlz_matrix_get_csr(const mlz_file_matrix *A, uint64_t *row_ptr, uint32_t *col, uint64_t *val)
{
	if (!A || !row_ptr || ((!col || !val) && !A->col.empty()))
		return fail(MLZ_E_ARG, "null argument");
	memcpy(row_ptr, A->row_ptr.data(), A->row_ptr.size() * sizeof(uint64_t));
	if (!A->col.empty()) {
		memcpy(col, A->col.data(), A->col.size() * sizeof(uint32_t));
		memcpy(val, A->val.data(), A->val.size() * sizeof(uint64_t));
	}
	return MLZ_OK;
}

int mlz_matrix_release(mlz_file_matrix *A)
{
	delete A;
	return MLZ_OK;
}

int mlz_matrix_write_csr(const char *path, int format, int elem, uint64_t p, size_t m, size_t n,
			 const uint64_t *row_ptr, const uint32_t *col, const uint64_t *val)
{
	std::string why;
	if (!valid_ring(elem, p, why))
		return fail(MLZ_E_ARG, why);
	if (!row_ptr || (row_ptr[m] > 0 && (!col || !val)))
		return fail(MLZ_E_ARG, "null CSR array");
	const Ring R{elem, elem == MLZ_IO_GF2 ? 2 : p};
	Source S(R, m, n);
	S.rp = row_ptr;
	S.ci = col;
	S.cv = val;
	return write_any(path, format, S);
}

int mlz_matrix_write_dense(const char *path, int format, int elem, uint64_t p, size_t m, size_t n,
			   const void *A, size_t lda)
{
	std::string why;
	if (!valid_ring(elem, p, why))
		return fail(MLZ_E_ARG, why);
	if (m && n && !A)
		return fail(MLZ_E_ARG, "null matrix");
	if (m > 1 && lda < (elem == MLZ_IO_GF2 ? (n + 63) / 64 : n))
		return fail(MLZ_E_ARG, "leading dimension smaller than the row");
	const Ring R{elem, elem == MLZ_IO_GF2 ? 2 : p};
	Source S(R, m, n);
	S.A = A;
	S.lda = lda;
	return write_any(path, format, S);
}

const char *mlz_io_last_error(void)
{
	return g_err.c_str();
}

}  // extern "C"
