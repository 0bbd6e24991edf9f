"""Pin the oracle's LDL^T / QR restatement.

* LDL^T: bit-exact against the reference's own casadi_ldl / casadi_ldl_solve
  compiled from casadi/core/runtime/casadi_ldl.hpp (oracle/_ref); plus the
  reconstruction invariant ||P^T L D L^T P - A|| <= 1e-14 ||A||
  (test/python/matrix.py:1124-1140) and numpy.linalg.solve.
* det (LinsolQr::det): bit-exact against the reference's own casadi_det
  (casadi/core/runtime/casadi_det.hpp, oracle/_ref) times the permutation
  sign; numpy.linalg.det; exact values for permutation and triangular
  matrices (a wrong permutation sign or a dropped reflector factor fails).
* QR: casadi_qr is not buildable standalone (casadi_house needs calculus.hpp's
  if_else, whose header chain needs the generated casadi_export.h), so it is
  pinned by A = Q R reconstruction, numpy.linalg.solve / lstsq agreement and
  the singularity count on rank-deficient matrices.
"""
import numpy as np
import pytest

import csparse_ref as CS
import oracle_ref as O
from casadi_amd import workloads as W
from casadi_amd.sx import Sparsity

needs_ref = pytest.mark.skipif(not CS.AVAILABLE, reason="oracle/_ref not built (reference tree absent)")


@pytest.fixture(scope="module")
def kkt():
    sp, nw, nl = W.ocp_kkt_sparsity()
    Lt, p = sp.ldl(amd=True)
    return sp, Lt, p


def backward_err(K, x, b):
    """Normwise backward error ||Kx - b|| / (||K|| ||x|| + ||b||)."""
    return np.linalg.norm(K @ x - b) / (np.linalg.norm(K, 2) * np.linalg.norm(x) + np.linalg.norm(b))


def kkt_values(sp, seed):
    K = W.ocp_kkt_values(np.random.default_rng(seed))
    return K, K[sp.row, sp.get_col()]


@needs_ref
def test_ldl_bit_exact_vs_reference_runtime(kkt):
    sp, Lt, p = kkt
    for seed in range(3):
        K, a = kkt_values(sp, seed)
        lt, d = O.ldl(sp.compressed(), a, Lt.compressed(), p)
        lt_r, d_r = CS.ref_ldl(sp.compressed(), a, Lt.compressed(), p)
        assert np.array_equal(lt, lt_r) and np.array_equal(d, d_r)
        b = np.random.default_rng(seed + 10).standard_normal((2, K.shape[0]))
        x = O.ldl_solve(b.reshape(-1), Lt.compressed(), lt, d, p)
        x_r = CS.ref_ldl_solve(b.reshape(-1), Lt.compressed(), lt, d, p)
        assert np.array_equal(x, x_r)


def test_ldl_reconstruction_and_solve(kkt):
    sp, Lt, p = kkt
    K, a = kkt_values(sp, 7)
    n = K.shape[0]
    lt, d = O.ldl(sp.compressed(), a, Lt.compressed(), p)
    U = np.zeros((n, n))
    U[Lt.row, Lt.get_col()] = lt
    Lm = U.T + np.eye(n)
    P = np.eye(n)[p]
    rec = P.T @ Lm @ np.diag(d) @ Lm.T @ P
    # quasi-definite KKT with a 1e-8 dual shift: |L|, |D| reach ~1e8
    assert np.linalg.norm(rec - K) <= 1e-9 * np.linalg.norm(K)
    b = np.random.default_rng(1).standard_normal(n)
    x = O.ldl_solve(b, Lt.compressed(), lt, d, p)
    xr = np.linalg.solve(K, b)
    # casadi_ldl does not pivot: with the AMD order some -1e-8 dual pivots come
    # early, |D| grows to ~1e8 and the backward error is ~1e-11 although
    # cond(K) is only ~1e4 -- the reference algorithm's own accuracy.
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
    assert backward_err(K, x, b) <= 1e-10


def test_ldl_planted_mistake_is_caught(kkt):
    """A transposed L in the reconstruction check fails it."""
    sp, Lt, p = kkt
    K, a = kkt_values(sp, 3)
    n = K.shape[0]
    lt, d = O.ldl(sp.compressed(), a, Lt.compressed(), p)
    U = np.zeros((n, n))
    U[Lt.row, Lt.get_col()] = lt
    wrong = U + np.eye(n)  # L^T used as L
    P = np.eye(n)[p]
    rec = P.T @ wrong @ np.diag(d) @ wrong.T @ P
    assert np.linalg.norm(rec - K) > 1e-3 * np.linalg.norm(K)


def _qr_dense(sp, a):
    V, R, prinv, pc = sp.qr_sparse(amd=True)
    v, r, beta = O.qr(sp.compressed(), a, V.compressed(), R.compressed(), prinv, pc)
    return V, R, prinv, pc, v, r, beta


def test_qr_reconstruction_and_solve(kkt):
    sp = kkt[0]
    K, a = kkt_values(sp, 5)
    n = K.shape[0]
    V, R, prinv, pc, v, r, beta = _qr_dense(sp, a)
    m = V.nrow
    Rd = np.zeros((m, n))
    Rd[R.row, R.get_col()] = r
    Vd = np.zeros((m, n))
    Vd[V.row, V.get_col()] = v
    # Q = H_0 H_1 ... H_{n-1}, H_c = I - beta_c v_c v_c^T
    Q = np.eye(m)
    for c in reversed(range(n)):
        Q = Q - beta[c] * np.outer(Vd[:, c], Vd[:, c] @ Q)
    PA = np.zeros((m, n))
    PA[prinv[:n], :] = K[:, pc]
    assert np.linalg.norm(Q @ Rd - PA) <= 1e-12 * np.linalg.norm(K)
    b = np.random.default_rng(2).standard_normal(n)
    x = O.qr_solve(b, 0, V.compressed(), v, R.compressed(), r, beta, prinv, pc)
    assert backward_err(K, x, b) <= 1e-13
    xt = O.qr_solve(b, 1, V.compressed(), v, R.compressed(), r, beta, prinv, pc)
    assert backward_err(K.T, xt, b) <= 1e-13


def test_qr_singularity_count():
    n = 12
    rng = np.random.default_rng(4)
    A = rng.standard_normal((n, n))
    A[:, 5] = A[:, 2] * 3.0       # rank n-1
    r_, c_ = np.nonzero(np.ones((n, n)))
    sp = Sparsity.triplet(n, n, r_, c_)
    a = A[sp.row, sp.get_col()]
    V, R, prinv, pc, v, r, beta = _qr_dense(sp, a)
    nul, rmin, irmin = O.qr_singular(r, R.compressed(), pc, 1e-12)
    assert nul == 1 and rmin < 1e-12
    nul2, _, _ = O.qr_singular(r, R.compressed(), pc, 1e-300)
    assert nul2 == 0


def _perm_sign(p):
    seen, cycles = np.zeros(len(p), bool), 0
    for i in range(len(p)):
        if not seen[i]:
            cycles += 1
            j = i
            while not seen[j]:
                seen[j] = True
                j = p[j]
    return -1.0 if (len(p) - cycles) % 2 else 1.0


def _qr_det(sp, a):
    V, R, prinv, pc = sp.qr_sparse(amd=True)
    v, r, beta = O.qr(sp.compressed(), a, V.compressed(), R.compressed(), prinv, pc)
    return O.qr_det(V.compressed(), v, R.compressed(), r, beta, prinv, pc), (V, R, prinv, pc, v, r, beta)


@pytest.mark.parametrize("seed", range(4))
def test_qr_det_matches_numpy_and_reference(seed):
    rng = np.random.default_rng(seed)
    n = 12
    K = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.4) + np.diag(rng.uniform(1, 2, n))
    sp = Sparsity.triplet(n, n, *np.nonzero(K.T)[::-1])
    a = K[sp.row, sp.get_col()]
    det, (V, R, prinv, pc, v, r, beta) = _qr_det(sp, a)
    ref = np.linalg.det(K)
    assert abs(det - ref) <= 1e-12 * abs(ref)
    if CS.AVAILABLE:
        core = CS.ref_det(V.compressed(), v, R.compressed(), r, beta)
        assert det == _perm_sign(prinv) * _perm_sign(pc) * core


def test_qr_det_exact_cases():
    # a permutation matrix: det = its sign, exactly
    n = 7
    perm = np.array([3, 0, 6, 1, 2, 5, 4])
    P = np.zeros((n, n))
    P[np.arange(n), perm] = 1.0
    sp = Sparsity.triplet(n, n, np.arange(n), perm)
    det, _ = _qr_det(sp, P[sp.row, sp.get_col()])
    assert det == round(np.linalg.det(P))
    # a diagonal matrix with powers of two: the product, exactly
    d = np.array([2.0, -4.0, 0.5, 8.0, -0.25])
    sp = Sparsity.triplet(5, 5, np.arange(5), np.arange(5))
    det, _ = _qr_det(sp, d)
    assert det == np.prod(d)
    # an upper triangular matrix (no Householder work beyond the diagonal)
    T = np.triu(np.random.default_rng(1).standard_normal((6, 6))) + np.diag(np.arange(1, 7.0))
    sp = Sparsity.triplet(6, 6, *np.nonzero(T.T)[::-1])
    det, _ = _qr_det(sp, T[sp.row, sp.get_col()])
    assert abs(det - np.prod(np.diag(T))) <= 1e-13 * abs(np.prod(np.diag(T)))


def test_incomplete_ldl_pattern_and_oracle():
    """LinsolLdl's "incomplete" option: the pattern is the strictly upper
    triangle of A(p, p), no fill-in (casadi/solvers/linsol_ldl.cpp:85-95).
    On a tridiagonal matrix in natural order there is no fill-in, so the
    incomplete and complete factorisations coincide; on the KKT pattern the
    incomplete one is a subset of the complete one."""
    from casadi_amd.linsol import _incomplete_ldl_pattern
    n = 9
    T = np.diag(np.full(n, 4.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    sp = Sparsity.triplet(n, n, *np.nonzero(T.T)[::-1])
    Lt_i, p_i = _incomplete_ldl_pattern(sp, amd=False)
    Lt_c, p_c = sp.ldl(amd=False)
    assert Lt_i == Lt_c and np.array_equal(p_i, p_c)
    a = T[sp.row, sp.get_col()]
    assert np.array_equal(O.ldl(sp.compressed(), a, Lt_i.compressed(), p_i)[1],
                          O.ldl(sp.compressed(), a, Lt_c.compressed(), p_c)[1])
    spk, _, _ = W.ocp_kkt_sparsity()
    Lt_i, p_i = _incomplete_ldl_pattern(spk, amd=True)
    Lt_c, p_c = spk.ldl(amd=True)
    assert np.array_equal(p_i, p_c)
    mi, mc = Lt_i.dense_mask(), Lt_c.dense_mask()
    assert np.all(mc[mi]) and mi.sum() < mc.sum() and not np.any(np.tril(mi))
    # every entry of the permuted A's strict upper triangle, and nothing else
    Ap = spk.dense_mask()[np.ix_(p_i, p_i)]
    assert np.array_equal(mi, np.triu(Ap, 1))
