"""Interpolated bounce-back with precomputed q-distances in the oracle (the engine's ibb
option: the Filippova-Hanel / Mei-Luo-Shyy formula of K27:1846-1888 on both lattices, q read
from the qlink array).

Pins: (1) q = 1/2 on every link reduces the formula to half-way bounce-back exactly, so the
D3Q19 run equals the reference's simple bounce-back bit for bit; (2) on D3Q27, q-distances
equal to the signed-distance ratio reproduce the reference's own IBB bit for bit; (3) a
force-driven channel whose walls sit off the half-way positions converges to the Poiseuille
profile between the walls the q-distances place, not the half-way ones.
"""
import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases

import pyoracle as po


def _with_q(c, qlink):
    c.arrays["qlink"] = np.ascontiguousarray(qlink, np.float32)
    c.params["ibb"] = 1
    return c


def test_half_q_is_bounce_back_d3q19():
    c = cases.cylinder_channel(nx=48, ny=24, nz=4, D=8, cx=16, cy=12, Re=60.0, spz_cells=4, q=19)
    ref = po.run19(c, 30)
    c2 = cases.cylinder_channel(nx=48, ny=24, nz=4, D=8, cx=16, cy=12, Re=60.0, spz_cells=4, q=19)
    _with_q(c2, np.full(19 * c2.N, 0.5, np.float32))
    out = po.run19(c2, 30)
    for k in ("f", "rho", "u", "Fwx", "Fwy"):
        assert np.array_equal(out[k], ref[k]), k
    assert np.abs(ref["Fwx"]).sum() > 0


def test_sdf_ratio_q_is_the_reference_ibb_d3q27():
    c = cases.cylinder_channel(nx=48, ny=24, nz=4, D=8, cx=16, cy=12, Re=60.0, spz_cells=4, q=27,
                               sdf_from_geometry=True)
    ref = po.run19(c, 20)
    # q = |sdf0| / (|sdf0| + |sdf1|) per solid link, as K27:1848-1849 forms it in FP32
    sdf = c.arrays["sdf"].astype(np.float32)
    cx, cy, cz = (np.array(v) for v in po.ORC27_C)
    nx, ny, nz = c.nx, c.ny, c.nz
    ql = np.full((27, c.N), 0.5, np.float32)
    s3 = sdf.reshape(nz, ny, nx)
    for q in range(1, 27):
        up = np.roll(np.roll(np.roll(s3, int(cz[q]), axis=0), int(cy[q]), axis=1), int(cx[q]), axis=2).reshape(-1)
        a0 = np.abs(sdf)
        with np.errstate(invalid="ignore"):  # 0/0 only where no link is solid
            ql[q] = a0 / (a0 + np.abs(up))
    c2 = cases.cylinder_channel(nx=48, ny=24, nz=4, D=8, cx=16, cy=12, Re=60.0, spz_cells=4, q=27,
                                sdf_from_geometry=True)
    _with_q(c2, ql.reshape(-1))
    out = po.run19(c2, 20)
    for k in ("f", "rho", "u", "Fwx"):
        assert np.array_equal(out[k], ref[k]), k


def _offset_channel(q, ny, q_lo, q_hi, tau=0.8, force=1e-6, nx=4):
    """Periodic channel along x; solid rows j = 0 and j = ny-1; the walls cross the links of
    the first and last fluid rows at fractions q_lo and q_hi."""
    N = nx * ny
    solid = np.zeros((ny, nx), np.uint8)
    solid[0] = 1
    solid[ny - 1] = 1
    codes = {f: ("Cyclic", "Cyclic", "Cyclic") for f in cases.FACES}
    nu = (tau - 0.5) / 3.0
    c = cases.make_case("offset", nx, ny, 1, 1.0, 1.0, 1.0, nu * (1.0 / (nx - 1)), codes, solid=solid.reshape(-1),
                        sgs=lb.SGS_NONE, q=q)
    # the lattice-unit viscosity and force the case was built for
    c.arrays["omega"][:] = np.float32(1.0 / tau)
    c.omega0 = float(np.float32(1.0 / tau))
    c.params["dpdx"] = force
    nei = np.zeros((ny, nx), np.uint8)
    nei[1] = 1
    nei[ny - 2] = 1
    c.arrays["neiSolid"] = nei.reshape(-1)
    cxs, cys, czs = (np.array(v) for v in (po.ORC27_C if q == 27 else po.ORC19_C))
    ql = np.full((q, ny, nx), 0.5, np.float32)
    for d in range(1, q):
        if cys[d] == 1:    # pulls from the row below: crosses the lower wall
            ql[d, 1] = q_lo
        if cys[d] == -1:   # pulls from the row above: crosses the upper wall
            ql[d, ny - 2] = q_hi
    return _with_q(c, ql.reshape(-1)), nu


@pytest.mark.parametrize("q", [19, 27])
def test_offset_walls_give_their_poiseuille_profile(q):
    ny, q_lo, q_hi, F = 22, 0.2, 0.6, 1e-6
    c, nu = _offset_channel(q, ny, q_lo, q_hi, force=F)
    st = po.run19(c, 6000, precision="f64")
    u = st["u"].reshape(ny, c.nx)[:, 1]
    y = np.arange(ny, dtype=np.float64)
    y_lo, y_hi = 1 - q_lo, (ny - 2) + q_hi
    want = F / (2 * nu) * (y - y_lo) * (y_hi - y)
    half = F / (2 * nu) * (y - 0.5) * (ny - 1.5 - y)  # where half-way walls would put them
    fl = slice(1, ny - 1)
    err = np.abs(u[fl] - want[fl]).max() / want.max()
    err_half = np.abs(u[fl] - half[fl]).max() / want.max()
    print(f"q{q}: max deviation from the offset-wall profile {err:.2e}, from the half-way one {err_half:.2e}")
    # the Filippova-Hanel scheme carries a viscosity-dependent slip of a few 1e-3 here
    # (0.4 % on D3Q19, 0.8 % on D3Q27 with the cumulant collision); the half-way profile is
    # 5 % away
    assert err < 1.5e-2
    assert err_half > 4 * err
