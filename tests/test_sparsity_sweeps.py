"""Dependency (sparsity) propagation: the oracle's sp_forward / sp_reverse
restatement pinned, and the host's jac_sparsity checked against it.

The reference's tests hold no stored patterns for these sweeps
(test/python/sparsity.py checks patterns through Jacobians).  What fixes a
pattern here: closed forms (an elementwise map is diagonal, a map through a
sum is dense, x*y depends on both operands), the numeric Jacobian's nonzeros
at a random point (a structural pattern contains them, and equals them when
nothing cancels), duality (the reverse sweep's pattern is the forward sweep's
transpose) and the host layer's own sweeps (SXFunction::get_jac_sparsity).
"""
import numpy as np
import pytest

import oracle_ref as O


@pytest.fixture(scope="module")
def ca():
    import casadi_amd
    return casadi_amd


def unit_words(n):
    nw = max(1, -(-n // 64))
    u = np.zeros((n, nw), np.uint64)
    r = np.arange(n)
    u[r, r // 64] = np.left_shift(np.uint64(1), (r % 64).astype(np.uint64))
    return u


def bits(words, n_cols):
    b = np.unpackbits(np.ascontiguousarray(words).view(np.uint8).reshape(words.shape[0], -1), axis=1,
                      bitorder="little")
    return b[:, :n_cols].astype(bool)


def oracle_pattern(f, oind=0, iind=0, reverse=False):
    """Dense bool (nnz_out, nnz_in) pattern from the oracle's sweeps."""
    n_in, n_out = f.nnz_in(iind), f.nnz_out(oind)
    if not reverse:
        seeds = [None] * f.n_in()
        seeds[iind] = unit_words(n_in)
        return bits(O.sp_forward(f, seeds)[oind], n_in)
    seeds = [None] * f.n_out()
    seeds[oind] = unit_words(n_out)
    return bits(O.sp_reverse(f, seeds)[iind], n_out).T


def dense_of(sp):
    m = np.zeros(sp.shape, bool)
    m[sp.row, sp.get_col()] = True
    return m


def numeric_jacobian(f, x_aos, oind=0, iind=0):
    """(nnz_out, nnz_in) Jacobian from the oracle's numeric reverse sweep, one
    unit seed per output nonzero."""
    n_out = f.nnz_out(oind)
    args = [np.repeat(a, n_out, axis=0) for a in x_aos]
    seeds = [np.zeros((n_out, f.nnz_out(j))) for j in range(f.n_out())]
    seeds[oind] = np.eye(n_out)
    _, sens = O.adj_map(f, n_out, args, seeds)
    return sens[iind]


def test_oracle_sweeps_closed_forms(ca):
    x = ca.SX.sym("x", 70)
    y = ca.SX.sym("y")
    # elementwise: diagonal, across the 64-bit word boundary
    f = ca.Function("f", [x], [ca.sin(x) * 2])
    for rev in (False, True):
        assert np.array_equal(oracle_pattern(f, reverse=rev), np.eye(70, dtype=bool))
    # through a sum: dense
    g = ca.Function("g", [x], [x * ca.sum1(x)])
    for rev in (False, True):
        assert oracle_pattern(g, reverse=rev).all()
    # a binary op depends on both operands, a unary op on its one, a constant on none
    h = ca.Function("h", [x, y], [ca.vertcat(x[0] * y, ca.cos(x[1]), ca.SX(3.0) + 0 * y, x[2] / x[3])])
    pat = oracle_pattern(h, 0, 0)
    want = np.zeros((4, 70), bool)
    want[0, 0] = want[1, 1] = want[3, 2] = want[3, 3] = True
    assert np.array_equal(pat, want)
    assert np.array_equal(oracle_pattern(h, 0, 1)[:, 0], np.array([True, False, False, False]))
    # a sweep that dropped the second operand would lose x[3] and y
    assert pat[3, 3] and oracle_pattern(h, 0, 1)[0, 0]


@pytest.mark.parametrize("name", ["cartpole_step", "quadrotor_step", "arm_dynamics", "rollout5"])
def test_oracle_pattern_holds_the_numeric_jacobian(ca, name):
    from casadi_amd import workloads as W
    f = W.cartpole_rollout(5) if name == "rollout5" else getattr(W, name)()
    rng = np.random.default_rng(11)
    x = [rng.uniform(-1, 1, (1, f.nnz_in(i))) for i in range(f.n_in())]
    for o in range(f.n_out()):
        for i in range(f.n_in()):
            fwd, rev = oracle_pattern(f, o, i), oracle_pattern(f, o, i, reverse=True)
            assert np.array_equal(fwd, rev)                     # duality
            J = numeric_jacobian(f, x, o, i)
            assert not np.any((J != 0) & ~fwd)                  # structural superset
            if name == "cartpole_step":
                assert np.array_equal(J != 0, fwd)              # nothing cancels: equal
            host = dense_of(f.jac_sparsity(o, i))
            assert np.array_equal(host, fwd)


def test_oracle_reverse_consumes_seeds_and_ors_into_inputs(ca):
    x = ca.SX.sym("x", 3)
    f = ca.Function("f", [x], [ca.vertcat(x[0] + x[1], x[2])])
    seeds = [np.array([[0b01], [0b10]], np.uint64)]
    (sens,) = O.sp_reverse(f, seeds)
    assert sens[:, 0].tolist() == [0b01, 0b01, 0b10]


def test_random_words_forward_is_or_over_dependencies(ca):
    """For random seed words, each output word is the OR of the seed words of
    the inputs it depends on (linearity of the sweep over OR)."""
    from casadi_amd import workloads as W
    f = W.quadrotor_step()
    rng = np.random.default_rng(12)
    nw = 5
    seeds = [rng.integers(0, 2 ** 63, (f.nnz_in(i), nw), dtype=np.int64).astype(np.uint64)
             for i in range(f.n_in())]
    out = O.sp_forward(f, seeds)
    for o in range(f.n_out()):
        want = np.zeros((f.nnz_out(o), nw), np.uint64)
        for i in range(f.n_in()):
            P = oracle_pattern(f, o, i)
            for r in range(f.nnz_out(o)):
                for c in np.nonzero(P[r])[0]:
                    want[r] |= seeds[i][c]
        assert np.array_equal(out[o], want)


def test_device_sweep_arguments_are_checked(ca):
    """HipTape.sp_forward / sp_reverse take int64 CUDA tensors of 64-bit words,
    (nnz, >= n_words) per block; anything else is refused before a launch."""
    import torch
    from casadi_amd import workloads as W
    from casadi_amd.hip import HipTape
    t = HipTape(W.cartpole_step())
    with pytest.raises(TypeError):
        t.sp_forward([torch.zeros((4, 2), dtype=torch.int64), None], 2)        # host tensor
    with pytest.raises(TypeError):
        t.sp_reverse([np.zeros((4, 2), np.int64)], 2)                           # not a tensor
    if torch.cuda.is_available():
        with pytest.raises(ValueError):
            t.sp_forward([torch.zeros((3, 2), dtype=torch.int64, device="cuda"), None], 2)
