"""Brute-force evaluation of BRMerge's result by its tree definition.

Independent of the merge loop: for every output column, the contributions
fl(a_p * b_pc) of the row's list positions p are combined by recursion over
the aligned power-of-two tree on positions [0, 2^ceil(log2 L)), a node with
one present child passing that child through.  Pure Python (IEEE binary64,
round to nearest, no FMA); for tiny inputs only.
"""


def _tree(contrib, lo, size):
    # contrib: sorted list of (pos, value) within [lo, lo + size)
    if not contrib:
        return None
    if size == 1:
        return contrib[0][1]
    half = size // 2
    left = [c for c in contrib if c[0] < lo + half]
    right = [c for c in contrib if c[0] >= lo + half]
    vl = _tree(left, lo, half)
    vr = _tree(right, lo + half, half)
    if vl is None:
        return vr
    if vr is None:
        return vl
    return vl + vr


def brute_spgemm(A, B):
    """A, B: (rpt, col, val[, ncols]) host arrays.  Returns (rpt, col, val) lists."""
    arpt, acol, aval = [list(map(int, A[0])), list(map(int, A[1])), list(map(float, A[2]))]
    brpt, bcol, bval = [list(map(int, B[0])), list(map(int, B[1])), list(map(float, B[2]))]
    M = len(arpt) - 1
    rpt, col, val = [0], [], []
    for i in range(M):
        L = arpt[i + 1] - arpt[i]
        per_col = {}
        for p in range(L):
            a = arpt[i] + p
            k = acol[a]
            for q in range(brpt[k], brpt[k + 1]):
                per_col.setdefault(bcol[q], []).append((p, aval[a] * bval[q]))
        size = 1
        while size < L:
            size *= 2
        for c in sorted(per_col):
            col.append(c)
            val.append(_tree(per_col[c], 0, size))
        rpt.append(len(col))
    return rpt, col, val
