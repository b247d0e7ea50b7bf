"""Pins the oracle's whole-stage behaviour against what the mathematics fixes.

The per-face functors are pinned in test_oracle_functors.py; these tests pin
the pieces between them -- ghost halo and block decomposition, constrained
transport, the Heun update, the CFL dt and the stop rule -- with checks a
plausible mistake would fail:

* Sod against Toro's exact Riemann solution (PLM + HLLC + RK2 + outflow
  sides), and first-to-second order convergence of the L1 error.
* Circularly polarised Alfven wave (an exact nonlinear MHD solution) after
  one period: second-order convergence.
* Periodic MHD turbulence: mass, momentum, energy and mean B conserved to
  roundoff; div B stays at roundoff (CT telescoping identity).
* A uniform state is a fixed point (E = -v x B on every edge, fluxes cancel).
* One block vs many blocks: bit-identical (halo from neighbours).
* A cyclic permutation of the axes permutes the solution.
* dt of a uniform state equals cfl * min_a dx_a / (c_f,a + |v_a|).
"""
import math

import numpy as np
import pytest

import kratos_amd
import oracle
from kratos_amd import problems

G = 5.0 / 3.0


# ------------------------------------------------------------ exact Riemann
def exact_riemann(wl, wr, gamma, xi):
    """Toro's exact solver for the ideal-gas Riemann problem; wl = (rho,u,p);
    returns (rho, u, p) sampled at xi = (x - x0) / t."""
    rl, ul, pl = wl
    rr, ur, pr = wr
    cl, cr = math.sqrt(gamma * pl / rl), math.sqrt(gamma * pr / rr)
    g1, g2 = (gamma - 1) / (2 * gamma), (gamma + 1) / (2 * gamma)

    def f(p, r, pk, ck):
        if p > pk:
            a, b = 2 / ((gamma + 1) * r), (gamma - 1) / (gamma + 1) * pk
            q = math.sqrt(a / (p + b))
            return (p - pk) * q, q * (1 - (p - pk) / (2 * (b + p)))
        return (2 * ck / (gamma - 1) * ((p / pk) ** g1 - 1),
                (p / pk) ** (-g2) / (r * ck))

    p = 0.5 * (pl + pr)
    for _ in range(100):
        fl, dl = f(p, rl, pl, cl)
        fr, dr = f(p, rr, pr, cr)
        dp = (fl + fr + ur - ul) / (dl + dr)
        p = max(p - dp, 1e-12)
        if abs(dp) < 1e-14 * p:
            break
    fl, _ = f(p, rl, pl, cl)
    fr, _ = f(p, rr, pr, cr)
    us = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    out = np.zeros((3, xi.size))
    for i, s in enumerate(xi):
        if s <= us:                                      # left of the contact
            if p > pl:
                sl = ul - cl * math.sqrt(g2 * p / pl + g1)
                st = (rl * ((p / pl + (gamma - 1) / (gamma + 1)) /
                            ((gamma - 1) / (gamma + 1) * p / pl + 1)), us, p)
                out[:, i] = (rl, ul, pl) if s < sl else st
            else:
                csl = cl * (p / pl) ** g1
                if s < ul - cl:
                    out[:, i] = (rl, ul, pl)
                elif s > us - csl:
                    out[:, i] = (rl * (p / pl) ** (1 / gamma), us, p)
                else:
                    c = 2 / (gamma + 1) * (cl + (gamma - 1) / 2 * (ul - s))
                    u = 2 / (gamma + 1) * (cl + (gamma - 1) / 2 * ul + s)
                    r = rl * (c / cl) ** (2 / (gamma - 1))
                    out[:, i] = (r, u, pl * (c / cl) ** (2 * gamma / (gamma - 1)))
        else:                                            # right of the contact
            if p > pr:
                sr = ur + cr * math.sqrt(g2 * p / pr + g1)
                st = (rr * ((p / pr + (gamma - 1) / (gamma + 1)) /
                            ((gamma - 1) / (gamma + 1) * p / pr + 1)), us, p)
                out[:, i] = (rr, ur, pr) if s > sr else st
            else:
                csr = cr * (p / pr) ** g1
                if s > ur + cr:
                    out[:, i] = (rr, ur, pr)
                elif s < us + csr:
                    out[:, i] = (rr * (p / pr) ** (1 / gamma), us, p)
                else:
                    c = 2 / (gamma + 1) * (cr - (gamma - 1) / 2 * (ur - s))
                    u = 2 / (gamma + 1) * (-cr + (gamma - 1) / 2 * ur + s)
                    r = rr * (c / cr) ** (2 / (gamma - 1))
                    out[:, i] = (r, u, pr * (c / cr) ** (2 * gamma / (gamma - 1)))
    return out


def test_exact_riemann_known_answer():
    # Toro (2009) Table 4.2, test 1: p* = 0.30313, u* = 0.92745
    xi = np.array([0.9274 - 1e-3, 0.9274 + 1e-2])
    s = exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, xi)
    assert s[2, 0] == pytest.approx(0.30313, abs=1e-5)
    assert s[1, 0] == pytest.approx(0.92745, abs=1e-5)
    assert s[0, 0] == pytest.approx(0.42632, abs=1e-5)      # left star density
    assert s[0, 1] == pytest.approx(0.26557, abs=1e-5)      # right star density


def sod_l1(nx):
    spec = kratos_amd.MeshSpec((nx, 2, 1), (nx, 2, 1), (0, -1, -1), (1, 1, 1),
                               ("out", "out", "out", "out", "per", "per"))
    phys = kratos_amd.Physics(mhd=False, gamma=1.4, cfl=0.4, t_lim=0.2, top_up=True)
    o = oracle.OracleMesh(spec, phys, n_threads=1)
    o.upload(problems.sod(N=(nx, 2, 1)))
    o.advance(100000)
    assert o.time()[0] == 0.2
    u, _ = o.download()
    xc = (np.arange(nx) + 0.5) / nx
    ex = exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, (xc - 0.5) / 0.2)
    rho = u[0, 0, 0]
    np.testing.assert_array_equal(u[:, 0, 1], u[:, 0, 0])   # y-uniform stays y-uniform
    return float(np.mean(np.abs(rho - ex[0])))


def test_sod_against_exact_solution():
    e256, e512 = sod_l1(256), sod_l1(512)
    assert e512 < 4e-3, e512
    assert 1.4 < e256 / e512 < 2.6, (e256, e512)         # shocks limit the order to ~1


# ------------------------------------------------------------ Alfven wave
def alfven(nx, amp=0.1, t_end=1.0):
    """Circularly polarised Alfven wave along x on nx x 2 x 2 cells."""
    N = (nx, 2, 2)
    spec = kratos_amd.MeshSpec(N, N)
    phys = kratos_amd.Physics(mhd=True, gamma=G, cfl=0.3, t_lim=t_end, top_up=True)
    xc = (np.arange(nx) + 0.5) / nx
    rho, pre, b0 = 1.0, 0.1, 1.0
    by_c = amp * np.sin(2 * np.pi * xc)
    bz_c = amp * np.cos(2 * np.pi * xc)
    va = b0 / math.sqrt(rho)
    vy, vz = -by_c / math.sqrt(rho), -bz_c / math.sqrt(rho)
    bx = np.full((2, 2, nx + 1), b0)
    by = np.broadcast_to(by_c, (2, 3, nx)).copy()
    bz = np.broadcast_to(bz_c, (3, 2, nx)).copy()
    u = np.zeros((5, 2, 2, nx))
    u[0] = rho
    u[3] = rho * vy
    u[4] = rho * vz
    u[1] = pre / (G - 1) + 0.5 * rho * (vy ** 2 + vz ** 2) + 0.5 * (b0 ** 2 + by_c ** 2 + bz_c ** 2)
    o = oracle.OracleMesh(spec, phys, n_threads=1)
    o.upload(u, (bx, by, bz))
    o.advance(1000000)
    assert o.time()[0] == t_end
    _, (bx1, by1, bz1) = o.download()
    # the wave moves with v_A = 1: one period later it is back in place
    assert va * t_end == 1.0
    return float(np.mean(np.abs(by1[0, 0] - by_c)) + np.mean(np.abs(bz1[0, 0] - bz_c)))


def test_alfven_wave_second_order():
    # the limiter clips the extrema at coarse resolution; the ratio of
    # successive L1 errors tends to 4 (second order) as N grows
    e = [alfven(n) for n in (16, 32, 64, 128)]
    r = [e[i] / e[i + 1] for i in range(3)]
    assert e[3] < 6e-4, e
    assert r[0] > 2.2 and r[1] > 3.3 and r[2] > 3.7, r


# ------------------------------------------------------------ turbulence box
def run_box(N, blk, n_cycles, u, b, phys=None):
    spec = kratos_amd.MeshSpec(N, blk)
    phys = phys or kratos_amd.Physics(mhd=True, gamma=G, cfl=0.3)
    o = oracle.OracleMesh(spec, phys, n_threads=4)
    o.upload(u, b)
    assert o.advance(n_cycles) == n_cycles
    return o.download(), o


def divb(bs, N):
    bx, by, bz = bs
    return np.abs((bx[:, :, 1:] - bx[:, :, :-1]) * N[0] + (by[:, 1:, :] - by[:, :-1, :]) * N[1] +
                  (bz[1:] - bz[:-1]) * N[2]).max()


def test_mhd_conservation_and_divb():
    N = (16, 16, 16)
    u, b, _ = problems.turbulence(N, seed=11)
    (u1, b1), _ = run_box(N, (8, 8, 8), 6, u, b)
    for v in range(5):
        s0, s1 = u[v].sum(), u1[v].sum()
        scale = np.abs(u[v]).sum()
        assert abs(s1 - s0) <= 1e-13 * scale, (v, s0, s1)
    for a in range(3):
        # mean field through every plane is conserved exactly by CT
        assert abs(b1[a].sum() - b[a].sum()) <= 1e-12 * np.abs(b[a]).sum()
    assert divb(b, N) < 1e-12
    assert divb(b1, N) < 1e-11
    assert not np.array_equal(u1, u)                     # something did move


def test_uniform_state_fixed_point():
    N = (8, 8, 8)
    v = (0.3, -0.2, 0.1)
    B = (0.7, 0.4, -0.5)
    rho, pre = 1.3, 0.6
    u = np.zeros((5,) + N[::-1])
    u[0] = rho
    for a in range(3):
        u[2 + a] = rho * v[a]
    u[1] = pre / (G - 1) + 0.5 * rho * sum(x * x for x in v) + 0.5 * sum(x * x for x in B)
    bs = (np.full((8, 8, 9), B[0]), np.full((8, 9, 8), B[1]), np.full((9, 8, 8), B[2]))
    (u1, b1), _ = run_box(N, (4, 4, 4), 4, u, bs)
    np.testing.assert_allclose(u1, u, rtol=1e-14, atol=0)
    for x, y in zip(b1, bs):
        np.testing.assert_array_equal(x, y)


def test_block_decomposition_bit_identical():
    N = (16, 16, 16)
    u, b, _ = problems.turbulence(N, seed=5)
    (ua, ba), _ = run_box(N, (16, 16, 16), 4, u, b)
    for blk in [(8, 8, 8), (4, 16, 8)]:
        (ub, bb), _ = run_box(N, blk, 4, u, b)
        np.testing.assert_array_equal(ub, ua)
        for x, y in zip(bb, ba):
            np.testing.assert_array_equal(x, y)


def test_axis_permutation():
    """(x, y, z) -> (y, z, x): relabelling the axes relabels the solution."""
    N = (16, 16, 16)
    u, (bx, by, bz), _ = problems.turbulence(N, seed=2)
    # new axes: x' = y, y' = z, z' = x ; arrays are indexed [z, y, x]
    perm = lambda f: np.transpose(f, (2, 0, 1))            # f[z,y,x] -> g[z'=x, y'=z, x'=y]
    up = np.stack([perm(u[0]), perm(u[1]), perm(u[3]), perm(u[4]), perm(u[2])])
    bp = (perm(by), perm(bz), perm(bx))
    (u1, b1), _ = run_box(N, (8, 8, 8), 3, u, (bx, by, bz))
    (u2, b2), _ = run_box(N, (8, 8, 8), 3, up, bp)
    exp = np.stack([perm(u1[0]), perm(u1[1]), perm(u1[3]), perm(u1[4]), perm(u1[2])])
    np.testing.assert_allclose(u2, exp, rtol=0, atol=1e-12 * np.abs(u1).max())
    for got, want in zip(b2, (perm(b1[1]), perm(b1[2]), perm(b1[0]))):
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_dt_of_uniform_state():
    N = (8, 8, 8)
    v = (0.3, -0.2, 0.1)
    B = (0.7, 0.4, -0.5)
    rho, pre = 1.3, 0.6
    u = np.zeros((5,) + N[::-1])
    u[0] = rho
    for a in range(3):
        u[2 + a] = rho * v[a]
    u[1] = pre / (G - 1) + 0.5 * rho * sum(x * x for x in v) + 0.5 * sum(x * x for x in B)
    bs = (np.full((8, 8, 9), B[0]), np.full((8, 9, 8), B[1]), np.full((9, 8, 8), B[2]))
    _, o = run_box(N, (8, 8, 8), 1, u, bs)
    cs2 = G * pre / rho
    b2 = sum(x * x for x in B)
    dts = []
    for a in range(3):
        va2, vt2 = B[a] ** 2 / rho, (b2 - B[a] ** 2) / rho
        cf = math.sqrt(0.5 * (va2 + vt2 + cs2 + math.sqrt((va2 + vt2 - cs2) ** 2 + 4 * cs2 * vt2)))
        dts.append((1 / 8) / (cf + abs(v[a])))
    assert o.time()[1] == pytest.approx(0.3 * min(dts), rel=1e-14)
