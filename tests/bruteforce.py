"""Independent brute-force references (Python integers / numpy object arrays).

These do not follow LELA's loop structure: products are exact big-integer
matmuls reduced once, and row reduction picks the *last* nonzero candidate row
as pivot (LELA picks the first), which still yields the unique reduced
row-echelon form.  They pin the oracle on small inputs.
"""
import numpy as np


def matmul_mod(A, B, p):
    A = np.asarray(A).astype(object)
    B = np.asarray(B).astype(object)
    if A.shape[1] == 0:
        return np.zeros((A.shape[0], B.shape[1]), dtype=object)
    return (A.dot(B)) % p


def gemm_mod(p, a, A, B, b, C):
    return ((a * matmul_mod(A, B, p)) + b * np.asarray(C).astype(object)) % p


def rref_mod(A, p):
    """Reduced row-echelon form over Z/p (p prime), last-row pivoting."""
    M = [[int(x) % p for x in row] for row in np.asarray(A)]
    m = len(M)
    n = len(M[0]) if m else 0
    r = 0
    pivots = []
    for c in range(n):
        cand = [i for i in range(r, m) if M[i][c] % p]
        if not cand:
            continue
        i = cand[-1]
        M[r], M[i] = M[i], M[r]
        inv = pow(M[r][c], -1, p)
        M[r] = [(x * inv) % p for x in M[r]]
        for j in range(m):
            if j != r and M[j][c]:
                f = M[j][c]
                M[j] = [(x - f * y) % p for x, y in zip(M[j], M[r])]
        pivots.append(c)
        r += 1
        if r == m:
            break
    return np.array(M, dtype=object).reshape(m, n), r, pivots


def det_mod(A, p):
    """Determinant mod prime p by elimination over the field."""
    M = [[int(x) % p for x in row] for row in np.asarray(A)]
    n = len(M)
    d = 1
    for c in range(n):
        piv = next((i for i in range(c, n) if M[i][c]), None)
        if piv is None:
            return 0
        if piv != c:
            M[c], M[piv] = M[piv], M[c]
            d = -d
        d = d * M[c][c] % p
        inv = pow(M[c][c], -1, p)
        for i in range(c + 1, n):
            f = M[i][c] * inv % p
            if f:
                M[i] = [(x - f * y) % p for x, y in zip(M[i], M[c])]
    return d % p


def is_rref(R, p):
    """Structural check of the standard-layout RREF; returns the pivot list."""
    R = np.asarray(R).astype(object) % p
    m, n = R.shape
    pivots = []
    last = -1
    zero_seen = False
    for i in range(m):
        nz = [c for c in range(n) if R[i, c] != 0]
        if not nz:
            zero_seen = True
            continue
        assert not zero_seen, "nonzero row after a zero row"
        c = nz[0]
        assert c > last, "pivot columns not increasing"
        assert R[i, c] == 1, "pivot not 1"
        col = R[:, c]
        assert sum(1 for x in col if x != 0) == 1, "pivot column not a unit vector"
        pivots.append(c)
        last = c
    return pivots


def lela_pivot_sequence(A, p):
    """The pivots of LELA's transposition pivoting (first nonzero, columns left
    to right), characterised without eliminating A itself: at step i the pivot
    column is the first column c (after the previous pivot's) for which some
    row at a position >= i is, restricted to columns 0..c, outside the span of
    the pivot rows so far restricted the same way, and the pivot row is the
    first such position; its row then swaps with position i.  Ranks come from
    rref_mod (a different elimination order).  Returns (steps, order, cols):
    steps[i] = pivot position of step i, order = final row order, cols =
    pivot columns."""
    A = [[int(x) % p for x in row] for row in np.asarray(A)]
    m = len(A)
    n = len(A[0]) if m else 0
    order = list(range(m))
    steps, cols = [], []
    i, c = 0, 0

    def rank(rows):
        return rref_mod(np.array(rows, dtype=object).reshape(len(rows), -1), p)[1] if rows else 0

    while i < m and c < n:
        found = None
        for col in range(c, n):
            base = [A[order[t]][:col + 1] for t in range(i)]
            rb = rank(base)
            for pos in range(i, m):
                if rank(base + [A[order[pos]][:col + 1]]) > rb:
                    found = (col, pos)
                    break
            if found:
                break
        if not found:
            break
        col, pos = found
        steps.append(pos)
        cols.append(col)
        order[i], order[pos] = order[pos], order[i]
        i += 1
        c = col + 1
    return steps, order, cols


def echelon_given_order(A, order, cols, p):
    """The echelon form U fixed by the row order and pivot columns: U[t] is the
    unique vector of (PA)[t] + span(U[0..t-1]) that is zero on cols[0..t-1]
    (a closed form: subtract (PA)[t][J] (U_J)^-1 U with J the earlier pivot
    columns); rows past the rank are zero."""
    PA = np.array([[int(x) % p for x in np.asarray(A)[r]] for r in order], dtype=object).reshape(len(order), -1)
    m, n = PA.shape
    U = np.zeros((m, n), dtype=object)
    for t in range(len(cols)):
        row = PA[t].copy()
        if t:
            J = cols[:t]
            UJ = U[:t][:, J]
            inv = inverse_mod(UJ, p)
            coef = (PA[t, J].dot(inv)) % p
            row = (row - coef.dot(U[:t])) % p
        U[t] = row
    return U


def inverse_mod(M, p):
    """Inverse of a square matrix mod prime p (Gauss-Jordan on [M | I])."""
    M = [[int(x) % p for x in row] for row in np.asarray(M)]
    n = len(M)
    aug = [M[i] + [int(i == j) for j in range(n)] for i in range(n)]
    for c in range(n):
        piv = next(i for i in range(c, n) if aug[i][c])
        aug[c], aug[piv] = aug[piv], aug[c]
        inv = pow(aug[c][c], -1, p)
        aug[c] = [(x * inv) % p for x in aug[c]]
        for i in range(n):
            if i != c and aug[i][c]:
                f = aug[i][c]
                aug[i] = [(x - f * y) % p for x, y in zip(aug[i], aug[c])]
    return np.array([row[n:] for row in aug], dtype=object).reshape(n, n)


# ----------------------------------------------------------------- GF(2)

def gf2_unpack(W, ncols):
    W = np.asarray(W, dtype=np.uint64)
    rows, words = W.shape
    bits = np.zeros((rows, words * 64), dtype=np.uint8)
    for b in range(64):
        bits[:, b::64] = ((W >> np.uint64(b)) & np.uint64(1)).astype(np.uint8)
    return bits[:, :ncols]


def gf2_pack(bits):
    bits = np.asarray(bits, dtype=np.uint64)
    rows, ncols = bits.shape
    words = (ncols + 63) // 64
    padded = np.zeros((rows, words * 64), dtype=np.uint64)
    padded[:, :ncols] = bits
    W = np.zeros((rows, words), dtype=np.uint64)
    for b in range(64):
        W |= padded[:, b::64] << np.uint64(b)
    return W


def gf2_matmul_bits(A, B):
    return (np.asarray(A, dtype=np.int64) @ np.asarray(B, dtype=np.int64)) & 1


def gf2_rref_bits(A):
    M = np.array(A, dtype=np.uint8) & 1
    m, n = M.shape
    r = 0
    pivots = []
    for c in range(n):
        cand = np.nonzero(M[r:, c])[0]
        if cand.size == 0:
            continue
        i = r + cand[-1]
        if i != r:
            M[[r, i]] = M[[i, r]]
        rows = np.nonzero(M[:, c])[0]
        rows = rows[rows != r]
        M[rows] ^= M[r]
        pivots.append(c)
        r += 1
        if r == m:
            break
    return M, r, pivots


def fl_det(X, p):
    """FaugereLachartre's det output on an F4-shaped matrix from first
    principles: the heads of the rows scanned in order (a head that repeats
    marks a C|D row), then D' = D - C A^-1 B with A^-1 by Gauss-Jordan, and
    the heads of D' in the echelon form fixed by lela_pivot_sequence (each
    head is U[t][cols[t]] of echelon_given_order).  Returns (det, rank of D')."""
    X = np.asarray(X)
    m, n = X.shape
    d = 1 % p
    prow, pcol, last = [], [], -1
    for i in range(m):
        nz = np.nonzero(X[i] % p)[0]
        if nz.size == 0:
            break
        h = int(nz[0])
        assert h >= last
        if h > last:
            prow.append(i)
            pcol.append(h)
            last = h
        d = d * int(X[i, h]) % p
    rrows = [i for i in range(m) if i not in prow]
    rcols = [c for c in range(n) if c not in pcol]
    if not rrows or not rcols:
        return d, 0
    Xo = X.astype(object) % p
    D = Xo[np.ix_(rrows, rcols)]
    if prow:
        A = Xo[np.ix_(prow, pcol)]
        B = Xo[np.ix_(prow, rcols)]
        C = Xo[np.ix_(rrows, pcol)]
        D = (D - matmul_mod(C, matmul_mod(inverse_mod(A, p), B, p), p)) % p
    steps, order, cols = lela_pivot_sequence(D, p)
    U = echelon_given_order(D, order, cols, p)
    for t, c in enumerate(cols):
        d = d * int(U[t, c]) % p
    return d, len(cols)


def sparse_rule_replay(A, p, mode, words, perm_rows, perm_cols):
    """Checks a pivot sequence against the definition of LELA's
    SparsePartialPivotStrategy and replays the elimination it implies.

    At step i every row at a position >= i is described by (first nonzero at
    or after the search column, its count, its position), where the count is
    the number of nonzero entries from the search column on, or with `words`
    the number of nonzero 64-column words from the first nonzero's word on;
    the chosen row must carry the smallest such triple.  The elimination is
    then the textbook one (swap, clear the column below; mode 1 keeps the
    multipliers left of the diagonal, mode 2 also swaps the pivot column into
    place and stores a/x), over Python integers.  Returns the final matrix
    as an object array and det, or raises AssertionError."""
    M = [[int(x) % p for x in row] for row in np.asarray(A)]
    m = len(M)
    n = len(M[0]) if m else 0
    start = 0
    det = 1 % p
    for i, pr in enumerate(perm_rows):
        if mode == 2:
            start = i
        cands = []
        for r in range(i, m):
            nz = [c for c in range(start, n) if M[r][c] % p]
            if not nz:
                continue
            lead = nz[0]
            if words:
                cnt = len({c // 64 for c in range(lead, n) if M[r][c] % p})
            else:
                cnt = len(nz)
            cands.append((lead, cnt, r))
        assert cands, f"step {i}: no row left but a pivot was chosen"
        lead, _, best = min(cands)
        assert int(pr) == best, f"step {i}: chose row {pr}, the rule gives {best}"
        M[i], M[best] = M[best], M[i]
        if mode == 2:
            assert int(perm_cols[i]) == lead, f"step {i}: column {perm_cols[i]}, the rule gives {lead}"
            for r in range(m):
                M[r][i], M[r][lead] = M[r][lead], M[r][i]
        pc = i if mode == 2 else lead
        x = M[i][pc]
        det = det * x % p
        xinv = pow(x, -1, p)
        for j in range(i + 1, m):
            a = M[j][pc]
            if a == 0:
                continue
            c = (-a * xinv) % p
            c0 = pc if mode == 0 else 0 if mode == 1 else i + 1
            for col in range(c0, n):
                M[j][col] = (M[j][col] + c * M[i][col]) % p
            if mode == 1:
                M[j][i] = c
            elif mode == 2:
                M[j][i] = (-c) % p
        start = pc + 1
    # the sequence must be complete: no row left with a nonzero at or after start
    k = len(perm_rows)
    if mode == 2:
        start = k
    for r in range(k, m):
        assert not any(M[r][c] % p for c in range(start, n)), "elimination stopped early"
    return np.array(M, dtype=object).reshape(m, n), det
