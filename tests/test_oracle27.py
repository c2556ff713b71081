"""Pins of the D3Q27 oracle (oracle/lbm27_body.inc, the restatement of
D3Q27_opencl/kernel.cl): lattice tables against geometry, cumulant-collision invariants
and an exact relaxation law, a viscosity measurement, bounce-back and interpolated
bounce-back limits."""
import itertools

import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases
import pyoracle as po
from conftest import rel_l2

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1])
W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)


def feq27(rho, u, v, w):
    cu = np.multiply.outer(CX, u) + np.multiply.outer(CY, v) + np.multiply.outer(CZ, w)
    usq = u * u + v * v + w * w
    return W.reshape((27,) + (1,) * np.ndim(u)) * (rho + 3 * cu + 4.5 * cu * cu - 1.5 * usq)


def central(f, u, a, b, c):
    return (f * (CX - u[0]) ** a * (CY - u[1]) ** b * (CZ - u[2]) ** c).sum()


def test_lattice_is_isotropic_to_sixth_order():
    """D3Q27 weights and vectors (K27:2740-2747): sum w = 1, sum w c_a c_b = delta/3,
    fourth moments (1/9)(delta delta + ...)."""
    assert abs(W.sum() - 1) < 1e-15
    for a, b in itertools.product(range(3), repeat=2):
        C = [CX, CY, CZ]
        assert abs((W * C[a] * C[b]).sum() - (a == b) / 3) < 1e-15
    assert abs((W * CX ** 2 * CY ** 2).sum() - 1 / 9) < 1e-15
    assert abs((W * CX ** 4).sum() - 1 / 3) < 1e-15


def test_upwind_and_mirror_tables():
    L = po.orc()
    for nx, ny, nz in [(4, 3, 3), (3, 4, 1)]:
        for b1, b2, b3 in itertools.product([0, 1, 3, 4], repeat=3):
            for k, j, i in itertools.product(range(nz), range(ny), range(nx)):
                for q in range(27):
                    si, sj, sk = i - CX[q], j - CY[q], k - CZ[q]
                    wall = any((s < 0 or s >= n) and b in (1, 4, 5, 6, 7)
                               for s, n, b in ((si, nx, b1), (sj, ny, b2), (sk, nz, b3)))
                    exp = -1 if wall else (si % nx) + nx * ((sj % ny) + ny * (sk % nz))
                    assert L.orc27_upwind_B(q, i, j, k, nx, ny, nz, b1, b2, b3) == exp
    # mirrorQx/y/z (K27:186-528) are tables: the axis component flipped, or the whole
    # direction reflected when it has none (entries quoted from the tables below)
    for q in range(27):
        for ax in range(3):
            c = [CX[q], CY[q], CZ[q]]
            c = [-x for x in c] if c[ax] == 0 else [(-x if a == ax else x) for a, x in enumerate(c)]
            m = L.orc27_mirror(q, ax)
            assert (CX[m], CY[m], CZ[m]) == tuple(c)
    quoted = {(3, 0): 4, (15, 0): 16, (19, 0): 21, (1, 1): 2, (11, 1): 12, (7, 1): 9, (7, 2): 8,
              (19, 2): 25, (21, 2): 24, (26, 2): 20}
    for (q, ax), m in quoted.items():
        assert L.orc27_mirror(q, ax) == m


@pytest.mark.parametrize("seed", range(5))
def test_cumulant_conserves_and_relaxes_shear_exactly(seed):
    """collisionCumulantGeier (K27:2352-2691): mass and momentum are kept, and the
    off-diagonal second central moments relax as C' = (1 - omega) C exactly."""
    rng = np.random.default_rng(seed)
    u = 0.05 * rng.standard_normal(3)
    rho = 1 + 0.02 * rng.standard_normal()
    f = feq27(rho, *u) * (1 + 2e-3 * rng.standard_normal(27))
    rho = f.sum()
    u = np.array([(f * CX).sum(), (f * CY).sum(), (f * CZ).sum()]) / rho
    tau, tauSGS, omegaB = 0.6 + 0.4 * rng.random(), 0.02 * rng.random(), 1.0 + 0.5 * rng.random()
    om = 1 / (tau + tauSGS)
    out = po.collide27(f, rho, *u, tau, tauSGS, omegaB, 0.01)
    assert abs(out.sum() - rho) < 1e-14
    for c in (CX, CY, CZ):
        assert abs((out * c).sum() - (f * c).sum()) < 1e-14
    for a, b, c in ((1, 1, 0), (1, 0, 1), (0, 1, 1)):
        assert abs(central(out, u, a, b, c) - (1 - om) * central(f, u, a, b, c)) < 1e-14


def test_cumulant_rest_state_is_fixed_point():
    f = feq27(1.03, 0.0, 0.0, 0.0)
    out = po.collide27(f, 1.03, 0.0, 0.0, 0.0, 0.8, 0.0, 1.2, 0.01)
    assert np.abs(out - f).max() < 1e-14


def test_cumulant_taylor_green_viscosity():
    n, tau, U = 24, 0.8, 0.01
    c = cases.periodic_box(n, n, 1, tau=tau, sgs=lb.SGS_NONE, q=27)
    c.params["omegaB"] = 0.1  # input.hpp:198 default
    k = 2 * np.pi / n
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    u = U * np.sin(k * X) * np.cos(k * Y)
    v = -U * np.cos(k * X) * np.sin(k * Y)
    rho = 1 - 0.75 * U * U * (np.cos(2 * k * X) + np.cos(2 * k * Y))
    st = po.initial_state(c, np.float64)
    st["f"] = feq27(rho.reshape(-1), u.reshape(-1), v.reshape(-1), 0 * u.reshape(-1)).reshape(-1).copy()
    out = po.run19(c, 300, precision="f64", state=st)
    amp = np.sqrt(np.mean(out["u"] ** 2 + out["v"] ** 2)) / np.sqrt(np.mean(u ** 2 + v ** 2))
    nu_fit = -np.log(amp) / (2 * k * k * 300)
    assert abs(nu_fit / ((tau - 0.5) / 3) - 1) < 0.01


def test_closed_box_conserves_mass_d3q27():
    c = cases.cavity(6, sgs=lb.SGS_WALE, q=27)
    c.arrays["u0"][:] = 0
    rng = np.random.default_rng(2)
    st = po.initial_state(c, np.float64)
    st["f"] = (feq27(1.0, np.zeros(c.N), np.zeros(c.N), np.zeros(c.N)) * (1 + 1e-3 * rng.standard_normal((27, c.N))))
    st["f"] = st["f"].reshape(-1).copy()
    m0 = st["f"].sum()
    out = po.run19(c, 40, precision="f64", state=st)
    assert abs(out["f"].sum() - m0) < 1e-11 * m0


def test_ibb_half_distance_is_simple_bounce_back():
    """With sdf = 10000 everywhere q = 1/2, chi = 0 and the interpolated bounce-back of
    K27:1846-1888 reduces to simple bounce-back: a cylinder run with and without IBB
    distances equals the half-way rule, and with distances it differs."""
    c = cases.cylinder_channel(nx=40, ny=24, nz=4, D=8, cx=14, cy=12, Re=40.0, spz_cells=4, sgs=lb.SGS_NONE, q=27)
    a = po.run19(c, 30, precision="f64")
    # a second run with the same q = 1/2 written explicitly as equal distances
    c2 = cases.cylinder_channel(nx=40, ny=24, nz=4, D=8, cx=14, cy=12, Re=40.0, spz_cells=4, sgs=lb.SGS_NONE, q=27)
    near = c2.arrays["neiSolid"].astype(bool) | c2.arrays["solid"].astype(bool)
    c2.arrays["sdf"][near] = np.where(c2.arrays["solid"][near] == 1, -0.7, 0.7).astype(np.float32)
    b = po.run19(c2, 30, precision="f64")
    assert np.abs(a["f"] - b["f"]).max() < 1e-14
    c3 = cases.cylinder_channel(nx=40, ny=24, nz=4, D=8, cx=14, cy=12, Re=40.0, spz_cells=4, sgs=lb.SGS_NONE, q=27,
                                sdf_from_geometry=True)
    d = po.run19(c3, 30, precision="f64")
    assert np.abs(a["f"] - d["f"]).max() > 1e-8


# ---------------------------------------------------------------- D3Q27 recursive regularized
# The reference's RR collision (K19:3156-3309) carried over to the D3Q27 Hermite set, as the
# D3Q27 call site can select it (K27:2824).

def feq27_product(rho, u, v, w):
    """The D3Q27 tensor-product equilibrium: per axis the 1-D populations whose raw moments
    are (1, u, u^2 + 1/3); its Hermite coefficients are u^a v^b w^c for every multi-index,
    the equilibrium the recursive-regularized collision relaxes towards."""
    def one(c, x):
        s = x * x + 1.0 / 3.0
        return {-1: 0.5 * (s - x), 0: 1.0 - s, 1: 0.5 * (s + x)}[c]
    return np.array([rho * one(CX[q], u) * one(CY[q], v) * one(CZ[q], w) for q in range(27)])


def hermite27(f, a, b, c):
    """sum_q h_a(c_x) h_b(c_y) h_c(c_z) f_q, h_0 = 1, h_1 = c, h_2 = c^2 - 1/3."""
    h = lambda n, x: np.ones_like(x, dtype=float) if n == 0 else (x.astype(float) if n == 1 else x * x - 1.0 / 3.0)
    return (f * h(a, CX) * h(b, CY) * h(c, CZ)).sum()


@pytest.mark.parametrize("seed", range(6))
def test_rr27_conserves_and_relaxes_stress(seed):
    """Mass and momentum are kept; the deviatoric part of the non-equilibrium stress
    relaxes with omega, its trace with omegaB; the body force adds rho*dpdx of x-momentum."""
    rng = np.random.default_rng(seed)
    u = 0.05 * rng.standard_normal(3)
    rho = 1 + 0.03 * rng.standard_normal()
    f = feq27(rho, *u) * (1 + 2e-3 * rng.standard_normal(27))
    rho = f.sum()
    m = np.array([(f * CX).sum(), (f * CY).sum(), (f * CZ).sum()])
    u = m / rho
    tau, tauSGS, omegaB = 0.6 + 0.4 * rng.random(), 0.03 * rng.random(), 0.3 + rng.random()
    om = 1 / (tau + tauSGS)
    out = po.collide27_rr(f, rho, *u, tau, tauSGS, omegaB, 0.0)
    assert abs(out.sum() - rho) < 1e-14
    m2 = np.array([(out * CX).sum(), (out * CY).sum(), (out * CZ).sum()])
    assert np.abs(m2 - m).max() < 1e-14
    C = (CX, CY, CZ)
    P = lambda g: np.array([[(g * C[a] * C[b]).sum() for b in range(3)] for a in range(3)])
    Nb = P(f) / rho - np.eye(3) / 3 - np.outer(u, u)
    Na = P(out) / rho - np.eye(3) / 3 - np.outer(u, u)
    dev = lambda A: A - np.trace(A) / 3 * np.eye(3)
    assert np.abs(dev(Na) - (1 - om) * dev(Nb)).max() < 1e-13
    assert abs(np.trace(Na) - (1 - omegaB) * np.trace(Nb)) < 1e-13
    dpdx = 2e-4
    outF = po.collide27_rr(f, rho, *u, tau, tauSGS, omegaB, dpdx)
    assert abs(outF.sum() - rho) < 1e-14
    mF = np.array([(outF * CX).sum(), (outF * CY).sum(), (outF * CZ).sum()])
    assert abs(mF[0] - m[0] - rho * dpdx) < 1e-14 and np.abs(mF[1:] - m[1:]).max() < 1e-14


@pytest.mark.parametrize("seed", range(4))
def test_rr27_product_equilibrium_is_fixed_point(seed):
    """The tensor-product equilibrium at any velocity comes back unchanged (all 27 moments
    of the reconstruction round-trip, the equilibrium coefficient u^a v^b w^c of every
    order is the one relaxed towards)."""
    rng = np.random.default_rng(10 + seed)
    rho, u = 1 + 0.05 * rng.standard_normal(), 0.08 * rng.standard_normal(3)
    f = feq27_product(rho, *u)
    assert abs(f.sum() - rho) < 1e-14
    out = po.collide27_rr(f, rho, *u, 0.7, 0.01, 1.3, 0.0)
    assert np.abs(out - f).max() < 1e-15


@pytest.mark.parametrize("seed", range(3))
def test_rr27_filters_higher_orders_at_rest(seed):
    """At u = 0 every equilibrium and recursive non-equilibrium coefficient of order >= 3
    vanishes, so the post-collision populations carry no third- to sixth-order Hermite
    content, whatever the input held (the regularization removes the ghost modes)."""
    rng = np.random.default_rng(20 + seed)
    f = feq27(1.0, 0.0, 0.0, 0.0) * (1 + 5e-3 * rng.standard_normal(27))
    # remove the momentum so that u = 0 exactly
    rho = f.sum()
    for c in (CX, CY, CZ):
        f = f - c * (f * c).sum() / (2 * 9 / 27 * 0 + (c * c).sum()) * (c != 0) * 0 if False else f
    m = np.array([(f * c).sum() for c in (CX, CY, CZ)])
    for d, c in enumerate((CX, CY, CZ)):
        f = f - (m[d] / 2) * ((c == 1) & (CX ** 2 + CY ** 2 + CZ ** 2 == 1)) + (m[d] / 2) * (
            (c == -1) & (CX ** 2 + CY ** 2 + CZ ** 2 == 1))
    assert np.abs([(f * c).sum() for c in (CX, CY, CZ)]).max() < 1e-16
    rho = f.sum()
    out = po.collide27_rr(f, rho, 0.0, 0.0, 0.0, 0.8, 0.0, 1.1, 0.0)
    high = [(a, b, c) for a in range(3) for b in range(3) for c in range(3) if a + b + c >= 3]
    assert max(abs(hermite27(f, *i)) for i in high) > 1e-5  # the input has some
    assert max(abs(hermite27(out, *i)) for i in high) < 1e-15


def test_rr27_commutes_with_the_cube_symmetries():
    """Rotating or reflecting the lattice before the collision gives the rotated result:
    each of the 48 signed axis permutations maps populations, velocity and the result
    alike (a misplaced index in one of the mixed third- to sixth-order terms breaks it)."""
    rng = np.random.default_rng(7)
    u = np.array([0.04, -0.03, 0.02])
    f = feq27(1.01, *u) * (1 + 3e-3 * rng.standard_normal(27))
    rho = f.sum()
    u = np.array([(f * c).sum() for c in (CX, CY, CZ)]) / rho
    out = po.collide27_rr(f, rho, *u, 0.65, 0.0, 0.9, 0.0)
    vec = np.stack([CX, CY, CZ], 1)
    index = {tuple(v): q for q, v in enumerate(vec)}
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            M = np.zeros((3, 3))
            for r in range(3):
                M[r, perm[r]] = signs[r]
            # population q goes to the direction M c_q
            dst = np.array([index[tuple((M @ vec[q]).astype(int))] for q in range(27)])
            g = np.empty(27)
            g[dst] = f
            ug = M @ u
            outg = po.collide27_rr(g, rho, *ug, 0.65, 0.0, 0.9, 0.0)
            assert np.abs(outg[dst] - out).max() < 1e-15


def test_rr27_taylor_green_viscosity():
    """nu = (tau - 1/2)/3 from the decay rate of a Taylor-Green vortex (1 %)."""
    n, tau, U = 24, 0.8, 0.01
    c = cases.periodic_box(n, n, 1, tau=tau, sgs=lb.SGS_NONE, q=27)
    c.collision = lb.COLL_RR
    c.params["omegaB"] = 1.0
    k = 2 * np.pi / n
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    u = U * np.sin(k * X) * np.cos(k * Y)
    v = -U * np.cos(k * X) * np.sin(k * Y)
    rho = 1 - 0.75 * U * U * (np.cos(2 * k * X) + np.cos(2 * k * Y))
    st = po.initial_state(c, np.float64)
    st["f"] = feq27(rho.reshape(-1), u.reshape(-1), v.reshape(-1), 0 * u.reshape(-1)).reshape(-1).copy()
    out = po.run19(c, 300, precision="f64", state=st)
    amp = np.sqrt(np.mean(out["u"] ** 2 + out["v"] ** 2)) / np.sqrt(np.mean(u ** 2 + v ** 2))
    nu_fit = -np.log(amp) / (2 * k * k * 300)
    assert abs(nu_fit / ((tau - 0.5) / 3) - 1) < 0.01


def test_rr27_poiseuille_profile():
    """Force-driven channel between half-way bounce-back walls with the D3Q27 RR
    collision: the steady profile is u(y) = F/(2 nu) (y + 1/2)(ny - 1/2 - y),
    nu = (tau - 1/2)/3, with no cross flow (the force enters through S_q and the
    half-step velocity shift, K19:3641)."""
    c = cases.channel(nx=4, ny=20, nz=4, tau=0.8, force=1e-5, q=27)
    c.collision = lb.COLL_RR
    st = po.run19(c, 8000, precision="f64")
    u = st["u"].reshape(4, 20, 4)[1, :, 2]
    y = np.arange(20)
    F, nu = c.params["dpdx"], (1 / c.omega0 - 0.5) / 3
    ua = F / (2 * nu) * (y + 0.5) * (20 - 0.5 - y)
    assert rel_l2(u, ua) < 2e-3
    assert np.abs(st["v"]).max() < 1e-14 and np.abs(st["w"]).max() < 1e-14
    # the cumulant collision on the same lattice differs from it (the selection reaches
    # the step), by far less than the profile error bound
    c.collision = None
    cu = po.run19(c, 8000, precision="f64")["u"].reshape(4, 20, 4)[1, :, 2]
    assert 0 < np.abs(cu - u).max() and rel_l2(cu, ua) < 2e-3
