"""Batched LDL^T / QR kernels against the oracle (casadi_ldl / casadi_qr
restatement, itself bit-exact against the reference's casadi_ldl runtime).
Only + - * / sqrt on both sides: the bar is bit-exact."""
import numpy as np
import pytest

import oracle_ref as O
from conftest import seed_of

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    from casadi_amd import workloads as W
    sp, vals = W.ocp_kkt_batch(np.random.default_rng(seed_of(4)), 8192)
    rhs = np.random.default_rng(seed_of(4) + 1).standard_normal((sp.nrow, 8192))
    return sp, vals, rhs


def test_ldl_batch_bit_exact(cuda, batch):
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    n_inst = vals.shape[1]
    print(f"[case] ldl batch n={sp.nrow} instances={n_inst}", flush=True)
    ls = Linsol("kkt", "ldl", sp)
    A = torch.tensor(vals, dtype=torch.float64, device="cuda")
    B = torch.tensor(rhs, dtype=torch.float64, device="cuda")
    X, status = ls.solve_batch(A, B)
    X = X.cpu().numpy()
    assert int(status.sum()) == 0
    spc, ltc = sp.compressed(), ls.sp_lt.compressed()
    for k in np.random.default_rng(0).choice(n_inst, 256, replace=False):
        lt, d = O.ldl(spc, vals[:, k], ltc, ls.p)
        xr = O.ldl_solve(rhs[:, k], ltc, lt, d, ls.p)
        assert np.array_equal(X[:, k], xr), k


def test_ldl_specialised_and_generic_engines_agree(cuda, batch):
    """The specialised factor + solve (straight-line program of the pattern)
    and the generic kernel give the same bits, zero-pivot counts included."""
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    n_inst = 4099
    v = vals[:, :n_inst].copy()
    v[:, 11] = 0.0
    A = torch.tensor(v, dtype=torch.float64, device="cuda")
    B = torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda")
    print(f"[case] ldl engines instances={n_inst}", flush=True)
    Xj, sj = Linsol("j", "ldl", sp, {"engine": "jit"}).solve_batch(A, B)
    Xi, si = Linsol("i", "ldl", sp, {"engine": "interp"}).solve_batch(A, B)
    assert torch.equal(sj, si) and int(sj[11]) == sp.nrow
    assert np.array_equal(Xj.cpu().numpy(), Xi.cpu().numpy(), equal_nan=True)


@pytest.mark.parametrize("tr", [False, True])
def test_qr_specialised_and_generic_engines_agree(cuda, batch, tr):
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    n_inst = 1031
    v = vals[:, :n_inst].copy()
    v[:, 5] = 0.0
    A = torch.tensor(v, dtype=torch.float64, device="cuda")
    B = torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda")
    print(f"[case] qr engines tr={tr} instances={n_inst}", flush=True)
    Xj, sj = Linsol("j", "qr", sp, {"engine": "jit"}).solve_batch(A, B, tr=tr)
    Xi, si = Linsol("i", "qr", sp, {"engine": "interp"}).solve_batch(A, B, tr=tr)
    assert torch.equal(sj, si) and int(sj[5]) == sp.nrow
    assert np.array_equal(Xj.cpu().numpy(), Xi.cpu().numpy(), equal_nan=True)


def test_qr_batch_bit_exact(cuda, batch):
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    n_inst = 2048
    ls = Linsol("kkt", "qr", sp)
    A = torch.tensor(vals[:, :n_inst], dtype=torch.float64, device="cuda")
    B = torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda")
    for tr in (False, True):
        print(f"[case] qr batch tr={tr}", flush=True)
        X, status = ls.solve_batch(A, B, tr=tr)
        X = X.cpu().numpy()
        assert int(status.sum()) == 0
        spc, svc, src = sp.compressed(), ls.sp_v.compressed(), ls.sp_r.compressed()
        for k in np.random.default_rng(1).choice(n_inst, 64, replace=False):
            v, r, beta = O.qr(spc, vals[:, k], svc, src, ls.prinv, ls.pc)
            xr = O.qr_solve(rhs[:, k], int(tr), svc, v, src, r, beta, ls.prinv, ls.pc)
            assert np.array_equal(X[:, k], xr), (k, tr)


def test_ldl_qr_agree_and_singular_flags(cuda, batch):
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    n_inst = 512
    v = vals[:, :n_inst].copy()
    v[:, 7] = 0.0                       # a zero matrix: every pivot vanishes
    A = torch.tensor(v, dtype=torch.float64, device="cuda")
    B = torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda")
    Xl, sl = Linsol("a", "ldl", sp).solve_batch(A, B)
    Xq, sq = Linsol("b", "qr", sp).solve_batch(A, B)
    sl, sq = sl.cpu().numpy(), sq.cpu().numpy()
    assert sl[7] == sp.nrow and sq[7] == sp.nrow
    assert np.all(np.delete(sl, 7) == 0) and np.all(np.delete(sq, 7) == 0)
    ok = np.delete(np.arange(n_inst), 7)
    Xl, Xq = Xl.cpu().numpy()[:, ok], Xq.cpu().numpy()[:, ok]
    # cond(KKT) with the 1e-8 dual shift and unpivoted LDL^T growth: 1e-6 agreement
    assert np.max(np.abs(Xl - Xq) / np.maximum(np.abs(Xq).max(axis=0), 1)) < 1e-6


def test_linsol_host_single_solve_and_contract(cuda):
    from casadi_amd import workloads as W
    from casadi_amd.linsol import Linsol
    sp, _, _ = W.ocp_kkt_sparsity(N=5)
    K = W.ocp_kkt_values(np.random.default_rng(3), N=5)
    a = K[sp.row, sp.get_col()]
    b = np.arange(sp.nrow, dtype=float)
    ls = Linsol("s", "ldl", sp)
    with pytest.raises(RuntimeError, match="nfact"):
        ls.solve(b)
    ls.nfact(a)
    x = ls.solve(b)
    berr = np.linalg.norm(K @ x - b) / (np.linalg.norm(K, 2) * np.linalg.norm(x) + np.linalg.norm(b))
    assert berr < 1e-11
    assert np.allclose(x, np.linalg.solve(K, b), rtol=1e-6, atol=1e-6 * np.abs(x).max())
