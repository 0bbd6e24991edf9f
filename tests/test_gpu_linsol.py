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
    for k in range(n_inst):              # every instance of config 5, bit for bit
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
    n_inst = vals.shape[1]                # config 5's 8192 instances, every one checked
    ls = Linsol("kkt", "qr", sp)
    A = torch.tensor(vals[:, :n_inst], dtype=torch.float64, device="cuda")
    B = torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda")
    for tr in (False, True):
        print(f"[case] qr batch tr={tr} instances={n_inst}", flush=True)
        X, status = ls.solve_batch(A, B, tr=tr)
        X = X.cpu().numpy()
        assert int(status.sum()) == 0
        spc, svc, src = sp.compressed(), ls.sp_v.compressed(), ls.sp_r.compressed()
        for k in range(n_inst):
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


@pytest.mark.parametrize("engine", ["jit", "interp"])
def test_ldl_nfact_then_solves_bit_exact(cuda, batch, engine):
    """cxh_ldl_nfact factors once, the factors stay on the device, and every
    cxh_ldl_solve after it (several right-hand-side counts, repeated) gives the
    bits of the fused factor + solve and of the oracle's casadi_ldl_solve on
    the oracle's L^T and D.  neig and rank count the oracle's D."""
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    n_inst = 8192 if engine == "jit" else 2051
    v = vals[:, :n_inst].copy()
    v[:, 9] = 0.0                        # zero pivots everywhere in one instance
    A = torch.tensor(v, dtype=torch.float64, device="cuda")
    print(f"[case] ldl nfact/solve engine={engine} instances={n_inst}", flush=True)
    ls = Linsol("kkt", "ldl", sp, {"engine": engine})
    ls.nfact(A)
    print("[step] ldl nfact done (factor program compiled)", flush=True)
    A.fill_(1.0)                         # the factors are kept, not the values
    st = ls.status()
    assert st[9] == sp.nrow and np.all(np.delete(st, 9) == 0)
    rng = np.random.default_rng(seed_of(4) + 5)
    spc, ltc = sp.compressed(), ls.sp_lt.compressed()
    sample = rng.choice(np.delete(np.arange(n_inst), 9), 64, replace=False)
    ref = {k: O.ldl(spc, v[:, k], ltc, ls.p) for k in sample}
    fused = Linsol("f", "ldl", sp, {"engine": engine})
    Av = torch.tensor(v, dtype=torch.float64, device="cuda")
    n = sp.nrow
    for nrhs in (1, 3, 1):
        B = rng.standard_normal((n * nrhs, n_inst))
        X = ls.solve(torch.tensor(B, dtype=torch.float64, device="cuda")).cpu().numpy()
        print(f"[step] ldl solve nrhs={nrhs} done", flush=True)
        for q in range(nrhs):            # each right-hand side is solved on its own
            Xf, _ = fused.solve_batch(Av, torch.tensor(B[q * n:(q + 1) * n], dtype=torch.float64, device="cuda"))
            assert np.array_equal(X[q * n:(q + 1) * n], Xf.cpu().numpy(), equal_nan=True), (nrhs, q)
        print(f"[step] ldl fused factor + solve nrhs={nrhs} done", flush=True)
        for k in sample[:16]:
            lt, d = ref[k]
            assert np.array_equal(X[:, k], O.ldl_solve(B[:, k], ltc, lt, d, ls.p)), (nrhs, k)
    neig, rank = ls.neig(), ls.rank()
    for k in sample:
        d = ref[k][1]
        assert neig[k] == int((d < 0).sum()) and rank[k] == int((d != 0).sum())
    # the all-zero instance: the first pivot is 0, every later one 0/0 = NaN,
    # and NaN != 0 counts towards the rank as in LinsolLdl::rank
    d9 = O.ldl(spc, v[:, 9], ltc, ls.p)[1]
    assert neig[9] == int((d9 < 0).sum()) and rank[9] == int((d9 != 0).sum())
    # inertia of [H J'; J -1e-8 I] with H positive definite and J of full row
    # rank: one negative eigenvalue per dual row
    from casadi_amd import workloads as W
    _, nw, nl = W.ocp_kkt_sparsity()
    assert nw + nl == sp.nrow and np.all(neig[sample] == nl)


def test_ldl_neig_is_the_inertia(cuda):
    """Sylvester's law of inertia: the number of negative D entries equals the
    number of negative eigenvalues of A (checked with numpy's eigenvalues)."""
    import torch
    from casadi_amd.linsol import Linsol
    from casadi_amd.sx import Sparsity
    rng = np.random.default_rng(seed_of(4) + 6)
    n, n_inst = 10, 70
    mats = []
    for k in range(n_inst):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        ev = rng.uniform(0.5, 2, n) * np.where(rng.random(n) < (k % 11) / 10, -1, 1)
        mats.append((Q * ev) @ Q.T)
    sp = Sparsity.dense(n, n)
    vals = np.stack([M.T.reshape(-1) for M in mats], axis=1)
    ls = Linsol("d", "ldl", sp)
    ls.nfact(torch.tensor(vals, dtype=torch.float64, device="cuda"))
    neig, rank = ls.neig(), ls.rank()
    for k, M in enumerate(mats):
        assert neig[k] == int((np.linalg.eigvalsh(M) < 0).sum()) and rank[k] == n


@pytest.mark.parametrize("engine", ["jit", "interp"])
def test_qr_nfact_solve_det(cuda, batch, engine):
    """cxh_qr_nfact / cxh_qr_solve (both directions) give the fused call's
    bits; cxh_qr_det gives the oracle's casadi_det times the permutation
    sign bit for bit."""
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    n_inst = 8192 if engine == "jit" else 1031
    A = torch.tensor(vals[:, :n_inst], dtype=torch.float64, device="cuda")
    print(f"[case] qr nfact/solve/det engine={engine} instances={n_inst}", flush=True)
    ls = Linsol("kkt", "qr", sp, {"engine": engine})
    ls.nfact(A)
    print("[step] qr nfact done (factor program compiled)", flush=True)
    spc, svc, src = sp.compressed(), ls.sp_v.compressed(), ls.sp_r.compressed()
    sample = np.random.default_rng(2).choice(n_inst, 16, replace=False)
    for tr in (False, True, False):
        B = torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda")
        X = ls.solve(B, tr).cpu().numpy()
        print(f"[step] qr solve tr={tr} done", flush=True)
        Xf, _ = Linsol("f", "qr", sp, {"engine": engine}).solve_batch(A, B, tr=tr)
        print(f"[step] qr fused factor + solve tr={tr} done", flush=True)
        assert np.array_equal(X, Xf.cpu().numpy())
    det = ls.det(A)
    for k in sample:
        v, r, beta = O.qr(spc, vals[:, k], svc, src, ls.prinv, ls.pc)
        assert det[k] == O.qr_det(svc, v, src, r, beta, ls.prinv, ls.pc), k
        x = O.qr_solve(rhs[:, k], 0, svc, v, src, r, beta, ls.prinv, ls.pc)
        assert np.array_equal(ls.solve(torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda"))
                              .cpu().numpy()[:, k], x)


def test_qr_nfact_fails_on_singular_and_cache(cuda, batch):
    """A singular instance makes nfact fail (LinsolQr::nfact); with the
    "cache" option, factorisations of earlier batches come back unchanged
    (the reference's test_cache pattern, test/python/linearsolver.py:506-532)."""
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    n_inst = 640
    ls = Linsol("kkt", "qr", sp, {"cache": 2})
    sing = vals[:, :n_inst].copy()
    sing[:, 3] = 0.0
    with pytest.raises(RuntimeError, match="nfact"):
        ls.nfact(torch.tensor(sing, dtype=torch.float64, device="cuda"))
    with pytest.raises(RuntimeError, match="nfact"):
        ls.solve(torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda"))
    batches = [vals[:, i * n_inst:(i + 1) * n_inst] for i in range(3)]
    B = torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda")

    def run(i):
        ls.nfact(torch.tensor(batches[i], dtype=torch.float64, device="cuda"))
        return ls.solve(B).cpu().numpy()
    S = [None] * 3
    S[0] = run(0)
    assert np.array_equal(run(0), S[0])
    S[1] = run(1)
    assert np.array_equal(run(1), S[1]) and np.array_equal(run(0), S[0])
    S[2] = run(2)
    assert np.array_equal(run(2), S[2]) and np.array_equal(run(0), S[0]) and np.array_equal(run(1), S[1])
    plain = Linsol("p", "qr", sp)
    for i in range(3):
        Xf, _ = plain.solve_batch(torch.tensor(batches[i], dtype=torch.float64, device="cuda"), B)
        assert np.array_equal(S[i], Xf.cpu().numpy())


def test_incomplete_ldl_matches_oracle(cuda, batch):
    """The "incomplete" option (no fill-in) runs casadi_ldl on the strictly
    upper pattern of A(p, p): bit for bit the oracle's casadi_ldl on that
    pattern."""
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    n_inst = 700
    ls = Linsol("inc", "ldl", sp, {"incomplete": True})
    assert ls.sp_lt.nnz() < sp.ldl()[0].nnz()
    ls.nfact(torch.tensor(vals[:, :n_inst], dtype=torch.float64, device="cuda"))
    X = ls.solve(torch.tensor(rhs[:, :n_inst], dtype=torch.float64, device="cuda")).cpu().numpy()
    spc, ltc = sp.compressed(), ls.sp_lt.compressed()
    for k in np.random.default_rng(3).choice(n_inst, 16, replace=False):
        lt, d = O.ldl(spc, vals[:, k], ltc, ls.p)
        assert np.array_equal(X[:, k], O.ldl_solve(rhs[:, k], ltc, lt, d, ls.p), equal_nan=True), k


def test_linsol_solve_contract_errors(cuda, batch):
    import torch
    from casadi_amd.linsol import Linsol
    sp, vals, rhs = batch
    ls = Linsol("kkt", "ldl", sp)
    with pytest.raises(RuntimeError, match="nfact"):
        ls.neig()
    ls.nfact(torch.tensor(vals[:, :100], dtype=torch.float64, device="cuda"))
    with pytest.raises(Exception, match="instances"):
        ls.solve(torch.tensor(rhs[:, :99], dtype=torch.float64, device="cuda"))
    with pytest.raises(NotImplementedError, match="det"):
        ls.det(vals[:, 0])
    with pytest.raises(NotImplementedError, match="neig"):
        Linsol("q", "qr", sp).neig()
