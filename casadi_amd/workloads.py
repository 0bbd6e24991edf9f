"""The benchmark workloads, written with the SX API the way a CasADi user would.

* cart-pole RK4 step f(x, u), dt = 0.05 (states x, theta, xdot, thetadot;
  m_cart = 1, m_pole = 0.1, half pole length l = 0.5, g = 9.81);
* quadrotor RK4 step, dt = 0.02 (position, velocity, roll-pitch-yaw, body
  rates; mass 1, inertia diag(0.01, 0.01, 0.02); controls: thrust + 3 torques);
* two-link planar arm forward dynamics (link lengths 1, point masses 1 at the
  link ends, gravity 9.81);
* the 100-step cart-pole rollout cost sum ||x_t||^2 + 0.01 u_t^2;
* the 20-step multiple-shooting quadrotor OCP Lagrangian;
* the horizon-30 OCP KKT matrix (nx = 4, nu = 1) for the batched LDL solve.
"""
import numpy as np

from .function import Function
from .sx import (SX, cos, dot, horzcat, mtimes, sin, sumsqr, tan, vertcat)

# --- cart-pole ------------------------------------------------------------------
CP_MC, CP_MP, CP_L, CP_G = 1.0, 0.1, 0.5, 9.81


def cartpole_ode(x, u):
    """Cart-pole dynamics (Barto, Sutton & Anderson 1983 form)."""
    th, xd, thd = x[1], x[2], x[3]
    s, c = sin(th), cos(th)
    mt = CP_MC + CP_MP
    temp = (u + CP_MP * CP_L * thd * thd * s) / mt
    thacc = (CP_G * s - c * temp) / (CP_L * (4.0 / 3.0 - CP_MP * c * c / mt))
    xacc = temp - CP_MP * CP_L * thacc * c / mt
    return vertcat(xd, thd, xacc, thacc)


def rk4(ode, x, u, dt):
    k1 = ode(x, u)
    k2 = ode(x + dt / 2 * k1, u)
    k3 = ode(x + dt / 2 * k2, u)
    k4 = ode(x + dt * k3, u)
    return x + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


def cartpole_step(dt=0.05):
    x = SX.sym("x", 4)
    u = SX.sym("u")
    return Function("cartpole", [x, u], [rk4(cartpole_ode, x, u, dt)], ["x", "u"], ["xf"])


def cartpole_cost_step(dt=0.05, u_weight=0.01):
    """One RK4 step carrying the running rollout cost: ([x; c], u) ->
    [x_next; c + |x|^2 + u_weight u^2].  Its mapaccum over N steps, plus
    |x_N|^2, is cartpole_rollout(N)'s cost."""
    x, c, u = SX.sym("x", 4), SX.sym("c"), SX.sym("u")
    xn = rk4(cartpole_ode, x, u, dt)
    cn = c + sumsqr(x) + u_weight * u * u
    return Function("cost_step", [vertcat(x, c), u], [vertcat(xn, cn)], ["xc", "u"], ["xc_next"])


def cartpole_rollout(n_steps=100, dt=0.05, u_weight=0.01):
    """J(x0, u[0:n]) = sum_t ||x_t||^2 + u_weight * u_t^2 over an RK4 rollout,
    x_t the state before step t plus the final state."""
    x0 = SX.sym("x0", 4)
    us = SX.sym("u", n_steps)
    x = x0
    J = SX(0)
    for t in range(n_steps):
        ut = us[t]
        J = J + sumsqr(x) + u_weight * ut * ut
        x = rk4(cartpole_ode, x, ut, dt)
    J = J + sumsqr(x)
    return Function("rollout", [x0, us], [J], ["x0", "u"], ["J"])


# --- quadrotor ------------------------------------------------------------------
QR_M, QR_J, QR_G = 1.0, (0.01, 0.01, 0.02), 9.81


def quadrotor_ode(x, u):
    v = x[3:6]
    phi, th, psi = x[6], x[7], x[8]
    w = x[9:12]
    T, tau = u[0], u[1:4]
    cph, sph, cth, sth, cps, sps = cos(phi), sin(phi), cos(th), sin(th), cos(psi), sin(psi)
    # body z axis in the world frame (ZYX Euler angles)
    zb = vertcat(cps * sth * cph + sps * sph, sps * sth * cph - cps * sph, cth * cph)
    acc = zb * (T / QR_M) - vertcat(0, 0, QR_G)
    # Euler angle rates from body rates
    p, q, r = w[0], w[1], w[2]
    phid = p + (q * sph + r * cph) * tan(th)
    thd = q * cph - r * sph
    psid = (q * sph + r * cph) / cth
    Jx, Jy, Jz = QR_J
    pd = (tau[0] - (Jz - Jy) * q * r) / Jx
    qd = (tau[1] - (Jx - Jz) * r * p) / Jy
    rd = (tau[2] - (Jy - Jx) * p * q) / Jz
    return vertcat(v, acc, phid, thd, psid, pd, qd, rd)


def quadrotor_step(dt=0.02):
    x = SX.sym("x", 12)
    u = SX.sym("u", 4)
    return Function("quadrotor", [x, u], [rk4(quadrotor_ode, x, u, dt)], ["x", "u"], ["xf"])


def quadrotor_ocp_lagrangian(N=20, dt=0.02, q_weight=1.0, r_weight=0.1):
    """Lagrangian L(w, lam) = sum_k q|x_k|^2 + r|u_k|^2 + q|x_N|^2
    + sum_k lam_k^T (F(x_k, u_k) - x_{k+1}) of the multiple-shooting OCP with
    w = [x_0, u_0, x_1, u_1, ..., x_{N-1}, u_{N-1}, x_N] (21*12 + 20*4 = 332)."""
    nx, nu = 12, 4
    w = SX.sym("w", (N + 1) * nx + N * nu)
    lam = SX.sym("lam", N * nx)
    Lag = SX(0)
    xs = []
    off = 0
    for k in range(N):
        xk = w[off:off + nx]
        uk = w[off + nx:off + nx + nu]
        off += nx + nu
        xs.append((xk, uk))
    xN = w[off:off + nx]
    for k in range(N):
        xk, uk = xs[k]
        xn = xs[k + 1][0] if k + 1 < N else xN
        Lag = Lag + q_weight * sumsqr(xk) + r_weight * sumsqr(uk)
        Lag = Lag + dot(lam[k * nx:(k + 1) * nx], rk4(quadrotor_ode, xk, uk, dt) - xn)
    Lag = Lag + q_weight * sumsqr(xN)
    return w, lam, Lag


# --- two-link arm -------------------------------------------------------------------
ARM_L1 = ARM_L2 = 1.0
ARM_M1 = ARM_M2 = 1.0
ARM_G = 9.81


def arm_dynamics():
    """xdot = [dq; M(q)^-1 (tau - C(q, dq) - G(q))] for a 2-link planar arm."""
    z = SX.sym("z", 6)  # q1, q2, dq1, dq2, tau1, tau2
    q1, q2, dq1, dq2, t1, t2 = (z[i] for i in range(6))
    l1, l2, m1, m2, g = ARM_L1, ARM_L2, ARM_M1, ARM_M2, ARM_G
    c2 = cos(q2)
    s2 = sin(q2)
    M11 = (m1 + m2) * l1 * l1 + m2 * l2 * l2 + 2 * m2 * l1 * l2 * c2
    M12 = m2 * l2 * l2 + m2 * l1 * l2 * c2
    M22 = m2 * l2 * l2
    h = m2 * l1 * l2 * s2
    C1 = -h * (2 * dq1 * dq2 + dq2 * dq2)
    C2 = h * dq1 * dq1
    G1 = (m1 + m2) * g * l1 * cos(q1) + m2 * g * l2 * cos(q1 + q2)
    G2 = m2 * g * l2 * cos(q1 + q2)
    r1 = t1 - C1 - G1
    r2 = t2 - C2 - G2
    det = M11 * M22 - M12 * M12
    qdd1 = (M22 * r1 - M12 * r2) / det
    qdd2 = (M11 * r2 - M12 * r1) / det
    return Function("arm", [z], [vertcat(dq1, dq2, qdd1, qdd2)], ["z"], ["zdot"])


# --- KKT matrix for the batched sparse LDL -----------------------------------------
def _kkt_blocks(N, nx, nu):
    """Block layout of the horizon-N OCP KKT system: primal w = [x_0, u_0, ...,
    x_{N-1}, u_{N-1}] (N*(nx+nu)) and N+1 equality blocks of nx rows:
    x_0 = x_init, x_{k+1} = A_k x_k + B_k u_k for k < N-1, and the terminal
    condition A_{N-1} x_{N-1} + B_{N-1} u_{N-1} = x_target.
    Dimension N*(nx+nu) + (N+1)*nx (= 274 for N=30, nx=4, nu=1)."""
    nw = N * (nx + nu)
    nl = (N + 1) * nx
    return nw, nl


def ocp_kkt_sparsity(N=30, nx=4, nu=1):
    """Symmetric pattern [[H, G^T], [G, -reg I]] of the OCP KKT matrix."""
    from .sx import Sparsity
    nw, nl = _kkt_blocks(N, nx, nu)
    n = nw + nl
    rows, cols = [], []

    def block(r0, c0, nr, nc):
        for i in range(nr):
            for j in range(nc):
                rows.extend((r0 + i, c0 + j))
                cols.extend((c0 + j, r0 + i))

    for k in range(N):
        o = k * (nx + nu)
        block(o, o, nx + nu, nx + nu)          # stage Hessian
    block(nw, 0, nx, nx)                       # x_0 = x_init: -I
    for k in range(N):
        ro = nw + (k + 1) * nx
        o = k * (nx + nu)
        block(ro, o, nx, nx + nu)              # [A_k B_k]
        if k + 1 < N:
            block(ro, o + nx + nu, nx, nx)     # -I on x_{k+1}
    for i in range(nl):                        # dual diagonal (regularisation)
        rows.append(nw + i)
        cols.append(nw + i)
    return Sparsity.triplet(n, n, rows, cols), nw, nl


def ocp_kkt_values(rng, N=30, nx=4, nu=1, reg=-1e-8):
    """Dense KKT matrix: stage Hessians A^T A + I (A ~ N(0,1)), dynamics blocks
    [A B] ~ N(0, 0.3) plus identity on A, -I couplings, and reg on the dual
    diagonal."""
    nw, nl = _kkt_blocks(N, nx, nu)
    n = nw + nl
    K = np.zeros((n, n))
    m = nx + nu
    for k in range(N):
        o = k * m
        A = rng.standard_normal((m, m))
        K[o:o + m, o:o + m] = A.T @ A + np.eye(m)
    K[nw:nw + nx, 0:nx] = -np.eye(nx)
    for k in range(N):
        ro = nw + (k + 1) * nx
        o = k * m
        AB = rng.normal(0.0, 0.3, (nx, m))
        AB[:, :nx] += np.eye(nx)
        K[ro:ro + nx, o:o + m] = AB
        if k + 1 < N:
            K[ro:ro + nx, o + m:o + m + nx] = -np.eye(nx)
    K[:nw, nw:] = K[nw:, :nw].T
    K[np.arange(nw, n), np.arange(nw, n)] = reg
    return K


def ocp_kkt_batch(rng, n_inst, N=30, nx=4, nu=1, reg=-1e-8):
    """Nonzeros of n_inst random KKT matrices (same distribution as
    ocp_kkt_values), laid out [nonzero][instance]; returns (sp, values)."""
    sp, nw, nl = ocp_kkt_sparsity(N, nx, nu)
    m = nx + nu
    vals = np.zeros((sp.nnz(), n_inst))
    col = sp.get_col()
    index = {(int(r), int(c)): k for k, (r, c) in enumerate(zip(sp.row, col))}

    def put(r0, c0, block):  # block: (n_inst, nr, nc); writes it and its transpose
        nr, nc = block.shape[1:]
        for i in range(nr):
            for j in range(nc):
                vals[index[(r0 + i, c0 + j)]] = block[:, i, j]
                vals[index[(c0 + j, r0 + i)]] = block[:, i, j]

    A = rng.standard_normal((n_inst, N, m, m))
    H = np.einsum("skij,skil->skjl", A, A) + np.eye(m)
    for k in range(N):
        put(k * m, k * m, H[:, k])
    eye = np.broadcast_to(-np.eye(nx), (n_inst, nx, nx))
    put(nw, 0, eye)
    AB = rng.normal(0.0, 0.3, (n_inst, N, nx, m))
    AB[:, :, :, :nx] += np.eye(nx)
    for k in range(N):
        ro = nw + (k + 1) * nx
        put(ro, k * m, AB[:, k])
        if k + 1 < N:
            put(ro, k * m + m, eye)
    for i in range(nl):
        vals[index[(nw + i, nw + i)]] = reg
    return sp, vals
