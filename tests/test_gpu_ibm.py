"""Immersed-boundary moving walls on the GPU (k_Gwall / k_Gibm / k_Force,
D3Q19_opencl/kernel.cl:3648-3807, D3Q27_opencl/kernel.cl:2828-2986).

* the reference's own kernels, built for gfx950, against the oracle on one pass: marker
  forcing, marker motion (translation + rotation, with the reference's axis dot product)
  and the population forcing; the reference's spread is not compared (it writes the
  pre-add value back over the atomic sum, K19:3760-3762);
* the HIP step with moving walls against the oracle over many steps (same parity bar as
  the other lattice tests), on both lattices.
"""
import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases
from conftest import rel_l2

import pyoracle as po
from parity_util import parity, run_hip

pytestmark = pytest.mark.gpu


def ring(cx, cy, z0, R, n, nz_layers=1, dz=1.0):
    t = np.arange(n) * 2 * np.pi / n
    pts = []
    for layer in range(nz_layers):
        pts.append(np.column_stack([cx + R * np.cos(t), cy + R * np.sin(t), np.full(n, z0 + layer * dz)]))
    return np.concatenate(pts).reshape(-1)


MOTION = np.array([0.002, -0.001, 0.0, 16.0, 12.0, 4.0, 0.0, 0.0, 1.0, 0.01])


def _case(q):
    c = cases.channel(nx=32, ny=24, nz=8, tau=0.8, force=1e-5, sgs=lb.SGS_WALE, q=q)
    return c


@pytest.mark.parametrize("q", [19, 27])
def test_reference_ibm_kernels_one_pass(q):
    c = _case(q)
    st = po.run19(c, 30)
    xyz = ring(16.0, 12.0, 3.3, 5.0, 48, nz_layers=2)
    rng = np.random.default_rng(5)
    Gin = (1e-3 * rng.standard_normal(3 * c.N)).astype(np.float32)
    xr, Gr, fr = po.run_reference_ibm(c, xyz, MOTION, st["u"], st["v"], st["w"], st["rho"], Gin, st["f"])
    # oracle, same inputs in FP32
    L = po.orc()
    P = po._p
    n = xyz.size // 3
    m32 = MOTION.astype(np.float32)
    xo = xyz.astype(np.float32).copy()
    Go = np.zeros((3, n), np.float32)
    L.orc_ibm_gwall_f32(P(m32), n, P(xo), c.nx, c.ny, c.nz, P(st["u"]), P(st["v"]), P(st["w"]), P(Go[0]), P(Go[1]),
                        P(Go[2]))
    scratch = np.zeros((3, c.N), np.float32)
    L.orc_ibm_spread_move_f32(P(m32), n, P(xo), c.nx, c.ny, c.nz, P(Go[0]), P(Go[1]), P(Go[2]), P(scratch[0]),
                              P(scratch[1]), P(scratch[2]))
    fo = st["f"].astype(np.float32).copy()
    G3 = Gin.reshape(3, -1)
    L.orc_ibm_force_f32(q, c.N, P(fo), P(np.ascontiguousarray(c.arrays["solid"], np.uint8)), P(st["rho"]),
                        P(G3[0]), P(G3[1]), P(G3[2]))
    eg = rel_l2(Gr, Go.reshape(-1))
    ex = np.abs(xr - xo).max()
    ef = np.abs(fr.astype(np.float64) - fo).max()
    print(f"q{q} ref vs oracle: Gm rel {eg:.2e}, positions max {ex:.2e}, f max {ef:.2e}", flush=True)
    assert eg <= 1e-5
    assert ex <= 1e-5
    assert ef <= 1e-7


@pytest.mark.parametrize("q", [19, 27])
def test_ibm_step_vs_oracle(q):
    c = _case(q)
    ibm = {"xyz": ring(16.0, 12.0, 3.3, 5.0, 48, nz_layers=2), "motion": MOTION}
    h, o32, o64 = parity(c, 60, label=f"ibm q{q}", ibm=ibm)
    ex = np.abs(h["ibm_xyz"].astype(np.float64) - o64["ibm_xyz"]).max()
    eg = rel_l2(h["ibm_G"], o64["ibm_G"])
    eG = rel_l2(np.concatenate([h[g] for g in ("GxIBM", "GyIBM", "GzIBM")])[np.tile(c.arrays["solid"] == 0, 3)],
                np.concatenate([o64[g] for g in ("GxIBM", "GyIBM", "GzIBM")])[np.tile(c.arrays["solid"] == 0, 3)])
    print(f"ibm q{q}: positions {ex:.2e}, marker forcing {eg:.2e}, lattice forcing {eG:.2e}", flush=True)
    assert ex <= 1e-4
    assert eg <= 1e-3
    assert eG <= 1e-3


def test_ibm_chunked_calls_match_one_call():
    # the spread adds with float atomics, whose order varies from run to run: two runs
    # agree to FP32 rounding, not bit for bit
    c = _case(19)
    ibm = {"xyz": ring(16.0, 12.0, 3.3, 5.0, 48, nz_layers=2), "motion": MOTION}
    a = run_hip(c, 30, ibm=ibm)
    b = run_hip(c, 30, ibm=ibm, chunks=4)
    for k in ("f", "u", "ibm_xyz", "ibm_G", "GxIBM"):
        e = rel_l2(a[k], b[k])
        assert e <= 2e-5, (k, e)


def test_one_marker_switches_ibm_off():
    c = _case(19)
    a = run_hip(c, 20)
    b = run_hip(c, 20, ibm={"xyz": np.array([16.0, 12.0, 3.0]), "motion": MOTION})
    assert np.array_equal(a["f"], b["f"])
