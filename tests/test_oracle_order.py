"""Pins of the order-dependent parts of the oracle's symbolic phase (CPU only).

The supernode row order, the supernodal child-list order, the amalgamation
decisions and the entry order of P A P^T are not fixed by the mathematics of
the factorization; they are fixed by the reference's traversal rules:

  P A P^T          each output column receives its entries in the order the
                   source is traversed: source columns ascending, each in
                   stored order (src/taucs_ccs_ops.c:231-253)
  new supernode    rows = {j}, then each child's update rows (children in
                   etree child order, ascending) not yet seen and > j, then
                   the rows of A's column j not yet seen, in stored order
                   (src/taucs_sn_llt.c:1455-1480)
  joined column    j joins its only child's supernode when A's column j is
                   marked by that child; j is swapped with the first update
                   row, the rest stays in place (:1413-1451)
  child list       the etree children's supernodes, each pushed at the front
                   of the list: the list is the etree order reversed
                   (:1483-1492)
  amalgamation     a parent merges with all its children when the rule of
                   :1716-1725 holds for the would-be merged supernode; merged
                   rows = the children's columns (in child-list order), the
                   parent's columns, then the children's and the parent's
                   update rows not yet seen; the new child list is the
                   grandchildren, each pushed at the front (:1606-1786)

Each pin below is derived by hand from those rules (the derivation is written
next to the expected values), or re-derives the decision from brute-force
column structures, so that a plausible mistake in the oracle -- sorted rows,
the child list in etree order, a shift instead of a swap, an over-eager or a
timid merge -- fails here.
"""
import numpy as np
import pytest

from oracle import oracle as O


def ccs(n, cols):
    """Lower CSC from per-column row lists in stored order (values: a
    diagonally dominant SPD matrix; only the pattern matters here)."""
    cp, ri, v = [0], [], []
    for j, rows in enumerate(cols):
        assert j in rows and all(i >= j for i in rows)
        for i in rows:
            ri.append(i)
            v.append(float(n + 1) if i == j else -1.0)
        cp.append(len(ri))
    return np.array(cp, np.int32), np.array(ri, np.int32), np.array(v)


def structure(sym):
    """{sn: (size, rows)} of the live slots and {sn: child list} (root too)."""
    live = {}
    for sn in range(sym["n_sn"]):
        s = int(sym["sn_size"][sn])
        if s:
            a = sym["struct_ptr"][sn]
            live[sn] = (s, [int(i) for i in sym["struct_idx"][a:a + sym["sn_up_size"][sn]]])
    kids = {}
    for sn in range(sym["n_sn"] + 1):
        c, lst = sym["first_child"][sn], []
        while c != -1:
            lst.append(int(c))
            c = sym["next_child"][c]
        if lst:
            kids[sn] = lst
    return live, kids


# ---- P A P^T entry order -------------------------------------------------

@pytest.mark.parametrize("seed", range(4))
def test_papt_columns_list_entries_in_source_traversal_order(seed):
    """Tag every stored entry of A with its position; in P A P^T the tags of
    each column must increase (source columns ascending, each in stored
    order).  A permutation that sorts rows, puts the diagonal first, or fills
    columns in any other order fails."""
    from tests.test_oracle import random_spd_pattern
    n, cp, ri, v = random_spd_pattern(40 + 10 * seed, 0.12, 100 + seed)
    inv = np.random.default_rng(seed).permutation(n).astype(np.int32)
    tags = np.arange(cp[n], dtype=np.float64)
    pcp, pri, tv = O.permute(n, cp, ri, tags, inv)
    for j in range(n):
        t = tv[pcp[j]:pcp[j + 1]]
        assert np.all(np.diff(t) > 0), f"column {j}: {t}"
    # and the entry with tag k lands in column min(inv[i], inv[j]), row max(..)
    col_of = np.repeat(np.arange(n), np.diff(cp))
    for j in range(n):
        for k, i in zip(tv[pcp[j]:pcp[j + 1]].astype(int), pri[pcp[j]:pcp[j + 1]]):
            a, b = inv[ri[k]], inv[col_of[k]]
            assert (j, i) == (min(a, b), max(a, b))
    # the rows within a column are generally NOT sorted: the order is a choice
    unsorted = sum(np.any(np.diff(pri[pcp[j]:pcp[j + 1]]) < 0) for j in range(n))
    assert unsorted > 0


# ---- hand-derived supernodal structures -----------------------------------

def comb20():
    """Twenty chains (a, b) = (2i, 2i+1) below a three-column root 40, 41, 42.
    A's column a = [42, 40, b, 41, a], column b and 40..42 the diagonal only.

    etree: parent(a) = b; struct(L_b) = {b, 40, 41, 42} so parent(b) = 40;
    parent(40) = 41, parent(41) = 42.  Children of 40: b = 1, 3, ..., 39.
    Fundamental supernodes, by the rules:
      a (leaf): rows [a] + A's column a in stored order = [a, 42, 40, b, 41]
      b joins (A's column b = [b], marked by a): b sits at position 3; it is
        swapped with position 1 (row 42): [a, b, 40, 42, 41]
        (a shift would give [a, b, 42, 40, 41]; sorting [a, b, 40, 41, 42])
      chain i becomes supernode i (the recursion visits b = 1, 3, ... in order)
      40 (20 children): rows [40] + chain 0's update rows > 40 not yet seen,
        in its order [40, 42, 41] -> 42, 41: [40, 42, 41]; supernode 20
      41 joins (A's column [41] marked by 40): 41 at position 2, swapped with
        position 1: [40, 41, 42]; 42 joins in place: size 3
      child list of 20: children b = 1, 3, ..., 39 pushed at the front in
        turn: [19, 18, ..., 0]; the root slot 21 has the child 20
    Amalgamation at 20: candidate columns 2*20 + 3 = 43 >= 16;
      nonzeros = 6 (own trapezoid) + 20 * 9 (each chain: (5-2)*2 + 3) = 186;
      candidate rows [38, 39, 36, 37, ..., 0, 1, 40, 41, 42], all 43 columns:
      the zero count of row r at position p is p+1 minus what the parts
      hold: chain i (at child-list position 19-i) holds 1 and 2 entries of
      its columns' rows and 2 of each of 40, 41, 42; 20 holds 1, 2, 3 of
      40, 41, 42.  Zeros = sum_k 4k (k = 0..19) = 760 >= 0.5 * 186
      (own size 3 < 50): refused.  The final structure is the fundamental one.
    """
    n = 43
    cols = [None] * n
    for i in range(20):
        a, b = 2 * i, 2 * i + 1
        cols[a] = [42, 40, b, 41, a]
        cols[b] = [b]
    for j in (40, 41, 42):
        cols[j] = [j]
    live = {i: (2, [2 * i, 2 * i + 1, 40, 42, 41]) for i in range(20)}
    live[20] = (3, [40, 41, 42])
    kids = {20: list(range(19, -1, -1)), 21: [20]}
    return n, cols, (21, live, kids), (21, live, kids)


def two_children():
    """n = 5; A's columns 0: [4, 2, 0], 1: [3, 2, 1], 2..4: the diagonal.

    etree: struct(L_0) = {0, 4, 2}, parent 2; struct(L_1) = {1, 3, 2},
    parent 2; struct(L_2) = {2, 4, 3}, parent 3; parent(3) = 4.
    Children of 2 in etree order: 0, 1.
    Fundamental:
      0 (leaf): [0, 4, 2]          supernode 0
      1 (leaf): [1, 3, 2]          supernode 1
      2 (two children): [2] + child 0's update rows > 2 not seen: 4;
        child 1's: 3; A's column 2 adds nothing: [2, 4, 3]; supernode 2,
        child list: 0, then 1 pushed at the front: [1, 0]
      3 joins 2: 3 at position 2 swapped with position 1: [2, 3, 4]
      4 joins: [2, 3, 4], size 3; the root slot 3 has the child 2
    Amalgamation at 2: candidate size 1 + 1 + 3 = 5 < 16: merged.  Rows:
      children's columns in child-list order (1, then 0), own columns
      [2, 3, 4], no new update rows: [1, 0, 2, 3, 4]; supernodes 0 and 1 are
      emptied; 2 has no grandchildren.
    """
    n = 5
    cols = [[4, 2, 0], [3, 2, 1], [2], [3], [4]]
    fund = (3, {0: (1, [0, 4, 2]), 1: (1, [1, 3, 2]), 2: (3, [2, 3, 4])}, {2: [1, 0], 3: [2]})
    final = (3, {2: (5, [1, 0, 2, 3, 4])}, {3: [2]})
    return n, cols, fund, final


def star5():
    """Leaves 0..4 under a center 5; A's column i = [5, i], column 5 = [5].
    Fundamental: leaf i -> supernode i, rows [i, 5]; the center gets [5]
    (its children's update rows hold only 5); child list [4, 3, 2, 1, 0].
    Amalgamation: candidate size 6 < 16, merged: the children's columns in
    child-list order, then 5: [4, 3, 2, 1, 0, 5]."""
    n = 6
    cols = [[5, i] for i in range(5)] + [[5]]
    fund = (6, {**{i: (1, [i, 5]) for i in range(5)}, 5: (1, [5])}, {5: [4, 3, 2, 1, 0], 6: [5]})
    final = (6, {5: (6, [4, 3, 2, 1, 0, 5])}, {6: [5]})
    return n, cols, fund, final


def path5():
    """Tridiagonal; A's column j = [j+1, j], column 4 = [4].  Columns 1..3
    are not nested in their child (each brings the new row j+1), so
    supernode j has rows [j, j+1] for j < 3, each the only child of the next;
    column 4 (A's column [4], marked by 3) joins supernode 3: [3, 4], size 2.
    Amalgamation from slot 3 merges bottom-up (every candidate < 16
    columns): [0, 1, 2, 3, 4] in one supernode, slot 3."""
    n = 5
    cols = [[j + 1, j] for j in range(4)] + [[4]]
    fund = (4, {0: (1, [0, 1]), 1: (1, [1, 2]), 2: (1, [2, 3]), 3: (2, [3, 4])},
            {1: [0], 2: [1], 3: [2], 4: [3]})
    final = (4, {3: (5, [0, 1, 2, 3, 4])}, {4: [3]})
    return n, cols, fund, final


def mesh2d_3():
    """mesh2d(3) in the natural order (x + 3y), the generator's entry order
    (y+1, x+1, diagonal).  Columns: 0 [3,1,0], 1 [4,2,1], 2 [5,2], 3 [6,4,3],
    4 [7,5,4], 5 [8,5], 6 [7,6], 7 [8,7], 8 [8].  The etree is the chain
    0 -> 1 -> ... -> 8.
      0 (leaf): [0, 3, 1]
      1: A's row 4 is not marked by 0: new; [1] + child's update rows > 1: 3;
         then A's 4, 2: [1, 3, 4, 2]
      2: new (row 5): [2] + 3, 4 (2 is not > 2) + A's 5: [2, 3, 4, 5]
      3: new (row 6): [3, 4, 5, 6];  4: new (row 7): [4, 5, 6, 7]
      5: new (row 8): [5, 6, 7, 8]
      6, 7, 8 are marked by their child: they join supernode 5 in place
         (each is already the first update row): size 4
    Amalgamation: every candidate is < 16 columns; bottom-up, each parent
    takes its child's columns first: one supernode [0, 1, ..., 8] in slot 5."""
    n = 9
    cols = [[3, 1, 0], [4, 2, 1], [5, 2], [6, 4, 3], [7, 5, 4], [8, 5], [7, 6], [8, 7], [8]]
    fund = (6, {0: (1, [0, 3, 1]), 1: (1, [1, 3, 4, 2]), 2: (1, [2, 3, 4, 5]),
                3: (1, [3, 4, 5, 6]), 4: (1, [4, 5, 6, 7]), 5: (4, [5, 6, 7, 8])},
            {1: [0], 2: [1], 3: [2], 4: [3], 5: [4], 6: [5]})
    final = (6, {5: (9, list(range(9)))}, {6: [5]})
    return n, cols, fund, final


EXAMPLES = {"comb20": comb20, "two_children": two_children, "star5": star5, "path5": path5,
            "mesh2d_3": mesh2d_3}


@pytest.mark.parametrize("name", sorted(EXAMPLES))
def test_hand_derived_supernodes(name):
    n, cols, fund, final = EXAMPLES[name]()
    cp, ri, v = ccs(n, cols)
    if name == "mesh2d_3":   # the columns above are the generator's own
        n2, cp2, ri2, _ = O.mesh2d(3)
        assert np.array_equal(cp, cp2) and np.array_equal(ri, ri2)
    for want, sym in ((fund, O.symbolic(n, cp, ri, amalgamate=False)), (final, O.symbolic(n, cp, ri))):
        n_sn, live, kids = want
        got_live, got_kids = structure(sym)
        assert sym["n_sn"] == n_sn
        assert got_live == live
        assert got_kids == kids


# ---- amalgamation decisions re-derived ------------------------------------

def _rule(new_size, own, zeros, nonzeros):
    return (new_size < 16 or (own < 50 and zeros < 0.5 * nonzeros)
            or (own < 250 and zeros < 0.25 * nonzeros) or (own < 500 and zeros < 0.10 * nonzeros)
            or zeros < 0.05 * nonzeros)


def _cases():
    from tests.test_oracle import random_spd_pattern
    out = []
    for seed in range(3):
        out.append(("random", random_spd_pattern(120 + 40 * seed, 0.03, 40 + seed), False))
    out.append(("mesh2d 24", O.mesh2d(24), True))
    out.append(("mesh3d 9", O.mesh3d(9, 9, 9), True))
    out.append(("mesh2d 50", O.mesh2d(50), True))
    return out


@pytest.mark.parametrize("case", range(6))
def test_every_merge_decision_rederived_from_brute_force(case):
    """For every supernode with children that amalgamation visits -- merged or
    refused alike -- the candidate's size, update size, nonzeros and zeros
    equal what the brute-force column structures of L give, and the decision
    equals the published rule evaluated on them (src/taucs_sn_llt.c:1716-1725).
    nonzeros = sum of the true column counts of the candidate's columns;
    zeros = the candidate's padded trapezoid minus those nonzeros."""
    from tests.test_oracle import brute_structure, dense
    name, (n, cp, ri, v), order = _cases()[case]
    inv = O.genmmd(n, cp, ri)[1] if order else np.arange(n, dtype=np.int32)
    pcp, pri, pv = O.permute(n, cp, ri, v, inv)
    S = brute_structure(dense(n, pcp, pri, pv))
    fund = O.symbolic(n, pcp, pri, amalgamate=False)
    sym = O.symbolic(n, pcp, pri, merge_log=True)
    log = sym["merge_log"]
    assert log, name
    fund_size = {int(sn): int(fund["sn_size"][sn]) for sn in range(fund["n_sn"])}
    merged = refused = 0
    for rec in log:
        cols = [int(c) for c in rec["cols"]]
        assert len(set(cols)) == len(cols) == rec["new_size"]
        rows = set().union(*[S[c] for c in cols])
        assert rows >= set(cols)
        nonzeros = sum(len(S[c]) for c in cols)
        s, m = rec["new_size"], len(rows)
        zeros = s * (s + 1) // 2 + (m - s) * s - nonzeros
        assert rec["new_up"] == m, (name, rec["sn"])
        assert rec["nonzeros"] == nonzeros and rec["zeros"] == zeros, (name, rec["sn"])
        assert rec["own_size"] == fund_size[rec["sn"]]
        assert rec["merged"] == _rule(s, rec["own_size"], zeros, nonzeros), (name, rec["sn"])
        merged += rec["merged"]
        refused += not rec["merged"]
    # the final partition is what the decisions say: every live supernode is
    # a fundamental one or the candidate of the last merge that formed it
    live, _ = structure(sym)
    last = {}
    for rec in log:
        if rec["merged"]:
            last[rec["sn"]] = sorted(int(c) for c in rec["cols"])
    fund_live, _ = structure(fund)
    for sn, (s, rows) in live.items():
        want = last.get(sn, sorted(fund_live[sn][1][:fund_live[sn][0]]))
        assert sorted(rows[:s]) == want
    if name.startswith("mesh2d 50"):
        assert merged > 0 and refused > 0   # both branches of the rule are exercised
