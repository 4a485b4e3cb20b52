"""Writes tests/golden/kat_bracketing.json: known-answer SpGEMM cases whose
expected outputs were derived by hand from the merge loop of
BRMerge-SpGEMM/inc/spgemm.h:639-707 (pairs (2j, 2j+1) per level, the odd
list carried unchanged, equal columns combined as src1 + src2, empty B rows
still occupying their list position) and the expand step at :618-624 (each
product aval*bval rounded to double before any add).

The reference publishes no tests or golden vectors and cannot be compiled in
this image (its utils.h:3 includes TBB's scalable_allocator.h, which is
absent), so these hand-derived cases are what pins the oracle's bracketing.
Values are chosen so that the reference's pairwise tree, a left-to-right sum,
a right fold, dropping empty lists and fused multiply-add all give different
bits.  Expected values are literals, not computed by any tree formula.

Run:  python tests/golden/make_golden.py   (rewrites the JSON in place)
"""
import json
import os

X = 2.0 ** 53          # 9007199254740992; X + 1 rounds back to X (ties-to-even)
E = 1.0 + 2.0 ** -52   # E*E = 1 + 2^-51 + 2^-104 exactly; rounds to 1 + 2^-51


def h(v):
    return float(v).hex()


def b_single_col0(k_rows, empty=()):
    """B with k_rows rows; row r holds (col 0, 1.0) unless r is in `empty`."""
    rpt, col, val = [0], [], []
    for r in range(k_rows):
        if r not in empty:
            col.append(0)
            val.append(1.0)
        rpt.append(len(col))
    return rpt, col, val


def one_row_a(avals, cols):
    return [0, len(avals)], list(cols), list(avals)


CASES = []


def add(name, derivation, A, B, bncols, C):
    CASES.append({
        "name": name,
        "derivation": derivation,
        "A": {"rpt": A[0], "col": A[1], "val": [h(v) for v in A[2]]},
        "B": {"rpt": B[0], "col": B[1], "val": [h(v) for v in B[2]], "ncols": bncols},
        "C": {"rpt": C[0], "col": C[1], "val": [h(v) for v in C[2]]},
    })


# A: four lists on one column, values [X, 1, 1, -X].
# level0: (X+1) -> X ; (1+(-X)) -> -(X-1) exact.  level1: X + -(X-1) = 1.
# left-to-right would give ((X+1)+1)-X = 0; right fold gives 2.
add("four_lists_tree", "((X+1)+(1+(-X))) = X-(X-1) = 1",
    one_row_a([X, 1.0, 1.0, -X], [0, 1, 2, 3]), b_single_col0(4), 1,
    ([0, 1], [0], [1.0]))

# B: three lists [X, 1, 1]: level0 (X+1)->X, list 2 carried; level1 X+1 -> X.
# A right fold would give X+2.
add("odd_list_carried", "((X+1)+1) = X ; right fold X+(1+1) = X+2 differs",
    one_row_a([X, 1.0, 1.0], [0, 1, 2]), b_single_col0(3), 1,
    ([0, 1], [0], [X]))

# C: positions [X, empty, 1, 1] (A col 1 points at an empty B row).
# level0: (X + nothing) = X ; (1+1) = 2.  level1: X+2 (exact, even).
# If the empty list were dropped: [X,1,1] -> X.
add("empty_list_keeps_position", "(X) + (1+1) = X+2 ; dropping the empty list would give X",
    one_row_a([X, 5.0, 1.0, 1.0], [0, 1, 2, 3]), b_single_col0(4, empty=(1,)), 1,
    ([0, 1], [0], [X + 2.0]))

# D: five lists [X, 1, 1, -X, 0.5]: ((X+1)+(1-X)) + 0.5 = 1.5.
# left-to-right 0.5, right fold 2.
add("five_lists_two_carries", "(((X+1)+(1+(-X))) + 0.5) = 1.5",
    one_row_a([X, 1.0, 1.0, -X, 0.5], [0, 1, 2, 3, 4]), b_single_col0(5), 1,
    ([0, 1], [0], [1.5]))

# E: exact cancellation keeps a structural +0.0; columns merge in order.
# B row0 = {1: 1.5, 3: 2.0}, row1 = {1: -1.5, 2: 4.0}; A = [1, 1].
add("cancellation_kept",
    "col1: 1.5 + (-1.5) = +0.0 kept; col2: 4; col3: 2",
    ([0, 2], [0, 1], [1.0, 1.0]),
    ([0, 2, 4], [1, 3, 1, 2], [1.5, 2.0, -1.5, 4.0]), 4,
    ([0, 3], [1, 2, 3], [0.0, 4.0, 2.0]))

# G: products are rounded before the add (no FMA): E*E -> 1+2^-51, then
# + (-(1+2^-51)) = 0.  fma(E, E, -(1+2^-51)) would be 2^-104.
add("product_rounded_before_add", "fl(E*E) + (-(1+2^-51)) = 0 ; FMA gives 2^-104",
    ([0, 2], [0, 1], [E, -1.0]),
    ([0, 1, 2], [0, 0], [E, 1.0 + 2.0 ** -51]), 1,
    ([0, 1], [0], [0.0]))

# G2: the same two products in the other list order, so that contracting
# either operand's multiply into the add is caught.
add("product_rounded_before_add_swapped", "(-(1+2^-51)) + fl(E*E) = 0 ; FMA on the right product gives 2^-104",
    ([0, 2], [0, 1], [-1.0, E]),
    ([0, 1, 2], [0, 0], [1.0 + 2.0 ** -51, E]), 1,
    ([0, 1], [0], [0.0]))

# F: rectangular (A 4x3, B 3x5), an empty A row, and a row whose lists are
# all empty B rows.  B row0 = {0: 2, 4: 3}, row1 = {} (empty), row2 = {2: 5}.
# A row0 = [1@0, 2@2] -> {0: 2, 2: 10, 4: 3}
# A row1 = []          -> {}
# A row2 = [7@1]       -> {} (only an empty list)
# A row3 = [0.5@0, 0.25@1, 4@2] -> {0: 1, 2: 20, 4: 1.5}
add("rectangular_empty_rows",
    "row-by-row by hand; no column collides, so every value is one rounded product",
    ([0, 2, 2, 3, 6], [0, 2, 1, 0, 1, 2], [1.0, 2.0, 7.0, 0.5, 0.25, 4.0]),
    ([0, 2, 2, 3], [0, 4, 2], [2.0, 3.0, 5.0]), 5,
    ([0, 3, 3, 3, 6], [0, 2, 4, 0, 2, 4], [2.0, 10.0, 3.0, 1.0, 20.0, 1.5]))

# H: two A rows over the same four B rows r0..r3, each holding one entry in
# column 7 with values [X, 1, 1, -X].  Row0 of A = [1, 1, 1, 1] gives lists
# [X, 1, 1, -X] -> ((X+1)+(1-X)) = 1.  Row1 = [-1, 1, 1, -1] gives lists
# [-X, 1, 1, X]: level0 (-X+1) = -(X-1) exact ; (1+X) -> X ; level1 = 1.
add("two_rows_column7",
    "row0 ((X+1)+(1-X)) = 1 ; row1 ((-X+1)+(1+X)) = -(X-1)+X = 1",
    ([0, 4, 8], [0, 1, 2, 3, 0, 1, 2, 3], [1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0]),
    ([0, 1, 2, 3, 4], [7, 7, 7, 7], [X, 1.0, 1.0, -X]), 8,
    ([0, 1, 2], [7, 7], [1.0, 1.0]))


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_bracketing.json")
    with open(out, "w") as f:
        json.dump({"x": h(X), "cases": CASES}, f, indent=1)
    print("wrote", out, len(CASES), "cases")


if __name__ == "__main__":
    main()
