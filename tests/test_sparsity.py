"""Colourings, orderings and symbolic factorisations.

AMD, the elimination tree and the LDL / QR patterns are pinned bit-exactly
against the reference's vendored CSparse (compiled into oracle/_ref from
external_packages/CSparse, of which CasADi's SparsityInternal routines are
declared modified versions: casadi/core/sparsity_internal.cpp:35-448, 661-973).
Colourings are pinned by their defining properties and special cases.
"""
import numpy as np
import pytest

import casadi_amd as ca
import csparse_ref as CS
from casadi_amd import workloads as W
from casadi_amd.sx import Sparsity

needs_ref = pytest.mark.skipif(not CS.AVAILABLE, reason="oracle/_ref not built (reference tree absent)")


def random_sym(n, density, seed):
    rng = np.random.default_rng(seed)
    m = rng.random((n, n)) < density
    m = m | m.T | np.eye(n, dtype=bool)
    r, c = np.nonzero(m)
    return Sparsity.triplet(n, n, r, c)


def random_rect(m, n, density, seed):
    rng = np.random.default_rng(seed)
    a = rng.random((m, n)) < density
    a[np.arange(min(m, n)), np.arange(min(m, n))] = True
    r, c = np.nonzero(a)
    return Sparsity.triplet(m, n, r, c)


def patterns():
    kkt, _, _ = W.ocp_kkt_sparsity()
    yield "kkt274", kkt
    for k, (n, d) in enumerate([(20, 0.1), (60, 0.05), (150, 0.02), (40, 0.3)]):
        yield f"rand{n}", random_sym(n, d, 100 + k)


# -- colourings ----------------------------------------------------------------
def check_uni(A, D):
    """Columns sharing a colour never share a row (structurally orthogonal)."""
    mask = A.dense_mask()
    for c in range(D.ncol):
        cols = D.row[D.colind[c]:D.colind[c + 1]]
        assert np.all(mask[:, cols].sum(axis=1) <= 1)
    assert sorted(D.row.tolist()) == list(range(A.ncol))


def test_uni_coloring_properties():
    for name, sp in patterns():
        D = sp.uni_coloring()
        check_uni(sp, D)
    dense = Sparsity.dense(5, 7)
    assert dense.uni_coloring().ncol == 7
    assert Sparsity.diag(9).uni_coloring().ncol == 1


def check_star(A, D):
    n = A.ncol
    color = np.empty(n, np.int64)
    for c in range(D.ncol):
        color[D.row[D.colind[c]:D.colind[c + 1]]] = c
    adj = A.dense_mask() & ~np.eye(n, dtype=bool)
    # distance-1 colouring
    r, c = np.nonzero(adj)
    assert np.all(color[r] != color[c])
    # no path on four vertices is two-coloured (star colouring)
    for v in range(n):
        for w in np.nonzero(adj[v])[0]:
            for x in np.nonzero(adj[w])[0]:
                if x == v:
                    continue
                for y in np.nonzero(adj[x])[0]:
                    if y in (w, v):
                        continue
                    assert not (color[v] == color[x] and color[w] == color[y]), (v, w, x, y)


def test_star_coloring_properties():
    for name, sp in patterns():
        if sp.ncol > 160:
            continue
        check_star(sp, sp.star_coloring())
    assert Sparsity.diag(6).star_coloring().ncol == 1
    assert Sparsity.dense(6, 6).star_coloring().ncol == 6


def test_largest_first_is_degree_ordered():
    for name, sp in patterns():
        order = sp.largest_first()
        deg = np.diff(sp.colind)[order]
        assert np.all(np.diff(deg) <= 0)
        assert sorted(order.tolist()) == list(range(sp.ncol))


def test_jacobian_colouring_matches_fd_columns():
    """The compressed Jacobian recovers J exactly: entries from the colouring
    equal the dense Jacobian of the cart-pole step."""
    f = W.cartpole_step()
    plan = f.jac_plan()
    assert plan["forward"] and plan["ndir"] == 5


# -- orderings / symbolic factorisations vs the reference's CSparse ---------------
@needs_ref
def test_amd_matches_reference_csparse():
    for name, sp in patterns():
        assert np.array_equal(sp.amd(), CS.amd(sp)), name
    # the quadrotor OCP Hessian pattern
    w, lam, Lag = W.quadrotor_ocp_lagrangian(N=4)
    H, _ = ca.hessian(Lag, w)
    assert np.array_equal(H.sparsity.amd(), CS.amd(H.sparsity))


@needs_ref
def test_etree_matches_reference_csparse():
    for name, sp in patterns():
        assert np.array_equal(sp.etree(), CS.etree(sp)), name
    for k, (m, n) in enumerate([(30, 20), (50, 50)]):
        sp = random_rect(m, n, 0.1, 7 + k)
        assert np.array_equal(sp.etree(ata=True), CS.etree(sp, ata=True))


@needs_ref
def test_ldl_pattern_matches_cholesky_pattern():
    rng = np.random.default_rng(0)
    for name, sp in patterns():
        Lt, p = sp.ldl(amd=True)
        assert np.array_equal(p, sp.amd())
        perm_sp = Sparsity.triplet(sp.nrow, sp.ncol, np.argsort(p)[sp.row], np.argsort(p)[sp.get_col()])
        # SPD values on the permuted pattern: diagonally dominant
        dense = np.zeros(sp.shape)
        mask = perm_sp.dense_mask()
        dense[mask] = rng.uniform(-1, 1, mask.sum())
        dense = (dense + dense.T) / 2 + np.diag(np.abs(dense).sum(axis=1) + 1)
        vals = dense[perm_sp.row, perm_sp.get_col()]
        colind, row = CS.chol_pattern(perm_sp, vals)
        # strictly lower part of CSparse's L == transpose of CasADi's L^T
        Lc = Sparsity(sp.nrow, sp.ncol, colind, row)
        strict = [(r, c) for c in range(Lc.ncol) for r in Lc.row[Lc.colind[c]:Lc.colind[c + 1]] if r > c]
        mine = Lt.T
        mine_set = {(r, c) for c in range(mine.ncol) for r in mine.row[mine.colind[c]:mine.colind[c + 1]]}
        assert mine_set == set(strict), name


@needs_ref
def test_qr_patterns_match_reference_csparse():
    cases = [random_rect(40, 40, 0.08, 21), random_rect(60, 45, 0.06, 22)]
    kkt, _, _ = W.ocp_kkt_sparsity(N=5)
    cases.append(kkt)
    for sp in cases:
        V, R, prinv, pc = sp.qr_sparse(amd=False)
        vc, vr, rc, rr, pinv, m2 = CS.qr_patterns(sp)
        assert V.nrow == m2 and np.array_equal(prinv, pinv)
        assert np.array_equal(V.colind, vc) and np.array_equal(R.colind, rc)
        for c in range(sp.ncol):
            assert set(V.row[V.colind[c]:V.colind[c + 1]]) == set(vr[vc[c]:vc[c + 1]])
            assert set(R.row[R.colind[c]:R.colind[c + 1]]) == set(rr[rc[c]:rc[c + 1]])


# -- colourings traced by hand through the reference's steps -------------------
def _sym(n, edges):
    r = list(range(n)) + [i for i, j in edges] + [j for i, j in edges]
    c = list(range(n)) + [j for i, j in edges] + [i for i, j in edges]
    return Sparsity.triplet(n, n, r, c)


def _cols_of(D):
    return [D.row[D.colind[c]:D.colind[c + 1]].tolist() for c in range(D.ncol)]


def test_largest_first_ties_hand_traced():
    """largest_first (casadi/core/sparsity_internal.cpp:3511-3544): a bucket
    sort by degree filled from the last column down, then reversed, so that
    columns come by decreasing degree and, within one degree, by increasing
    index.  Arrow 4x4 with hub 1: degrees [2, 4, 2, 2] -> [1, 0, 2, 3].
    Path 0-1-2-3: degrees [2, 3, 3, 2] -> [1, 2, 0, 3]."""
    assert _sym(4, [(1, 0), (1, 2), (1, 3)]).largest_first().tolist() == [1, 0, 2, 3]
    assert _sym(4, [(0, 1), (1, 2), (2, 3)]).largest_first().tolist() == [1, 2, 0, 3]


def test_star_coloring_hand_traced():
    """star_coloring with ordering 1 (casadi/core/sparsity_internal.cpp:
    3398-3509): colour the pattern permuted by largest_first with GMP
    Alg. 4.1, then map the colours back.

    Arrow with hub 1: the permuted hub (vertex 0) takes colour 0, every leaf
    is forbidden colour 0 only and takes colour 1, so D = [[1], [0, 2, 3]].
    (Without the ordering the hub would be coloured second and get colour 1.)

    Path 0-1-2-3, permuted by p_inv = [2, 0, 1, 3] into the path 2-0-1-3:
    vertex 0 -> colour 0; 1 -> 1 (0 forbidden by its coloured neighbour);
    2 -> 1 (only 0 forbidden); 3: its neighbour 1 has colour 1, and 1's
    neighbour 0 has a neighbour (2) of colour 1, so colours 0 and 1 are both
    forbidden (lines 12-16) -> a new colour 2.  Back in the original order:
    columns [0, 1, 2, 3] have colours [1, 0, 1, 2], D = [[1], [0, 2], [3]]."""
    D = _sym(4, [(1, 0), (1, 2), (1, 3)]).star_coloring()
    assert (D.nrow, D.ncol) == (4, 2) and _cols_of(D) == [[1], [0, 2, 3]]
    D = _sym(4, [(0, 1), (1, 2), (2, 3)]).star_coloring()
    assert (D.nrow, D.ncol) == (4, 3) and _cols_of(D) == [[1], [0, 2], [3]]
    # without the ordering (ordering 0), GMP Alg. 4.1 in the natural order:
    # the hub of the arrow comes second and gets colour 1
    D = _sym(4, [(1, 0), (1, 2), (1, 3)]).star_coloring(0)
    assert _cols_of(D) == [[0, 2, 3], [1]]


def test_uni_coloring_hand_traced():
    """uni_coloring (casadi/core/sparsity_internal.cpp:2959-3046): column i
    takes the first colour no earlier column sharing a row with it holds.
    A = 3x4 staircase (column rows {0}, {0,1}, {1,2}, {2}): colours
    [0, 1, 0, 1]: column 3 shares row 2 with column 2 (colour 0) only, so
    colour 0 is forbidden and colour 1 is free.  D lists each colour's
    columns in increasing order."""
    A = Sparsity.triplet(3, 4, [0, 0, 1, 1, 2, 2], [0, 1, 1, 2, 2, 3])
    D = A.uni_coloring()
    assert (D.nrow, D.ncol) == (4, 2) and _cols_of(D) == [[0, 2], [1, 3]]
    # columns 0 and 1 share no row: both colour 0; column 2 shares a row with
    # each of them, so colour 0 is forbidden and it opens colour 1
    B = Sparsity.triplet(2, 3, [0, 1, 0, 1], [0, 1, 2, 2])
    assert _cols_of(B.uni_coloring()) == [[0, 1], [2]]
