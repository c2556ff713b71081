"""Geometry preprocessing on the GPU (include/lbffs_geom.h) against the oracle.

Bar: the solid flags are bit-identical to the oracle's restatement of setSolid (and, on
the integer meshes, to exact arithmetic); the wall distance of setSDF agrees to FP32
rounding -- the reference takes r^-p from the host's powf, the kernel rounds it once from
FP64, so a few terms may differ in their last bit.
"""
import time

import numpy as np
import pytest

from lbffs_amd import cases
from lbffs_amd import geometry as geo

import pyoracle as po
from test_geom import ALIGNED, CYL_STL, L_CYL, L_STEP, STEP_STL, _plane_patch, exact_solid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(ALIGNED))
def test_solid_integer_meshes_bitwise(name):
    make, (nx, ny, nz) = ALIGNED[name]
    v0, v1, v2 = make()
    g = geo.solid_from_triangles(v0, v1, v2, nx, ny, nz)
    o = po.geom_solid(v0, v1, v2, nx, ny, nz)
    assert np.array_equal(g, o), name
    ex = exact_solid(v0, v1, v2, nx, ny, nz)
    assert np.array_equal(g[ex >= 0], ex[ex >= 0].astype(np.uint8))


def test_solid_keeps_and_inverts():
    make, (nx, ny, nz) = ALIGNED["box"]
    v0, v1, v2 = make()
    init = np.random.default_rng(5).integers(0, 3, nx * ny * nz).astype(np.uint8)
    for inv in (False, True):
        g = geo.solid_from_triangles(v0, v1, v2, nx, ny, nz, invert=inv, solid=init)
        o = po.geom_solid(v0, v1, v2, nx, ny, nz, invert=inv, solid=init)
        assert np.array_equal(g, o), inv


def test_solid_empty_surface():
    z = np.zeros((3, 0), np.float32)
    g = geo.solid_from_triangles(z, z, z, 5, 4, 3)
    assert not g.any()
    assert geo.solid_from_triangles(z, z, z, 5, 4, 3, invert=True).all()


@pytest.mark.parametrize("dims", [(1200, 400, 1), (240, 240, 12)])
def test_solid_cylinder_bitwise(dims):
    # flowAroundCylinder's own STL on its own 2-D lattice, and a 3-D cut through it
    nx, ny, nz = dims
    s = geo.read_stl(CYL_STL, L_CYL)
    t0 = time.perf_counter()
    g = geo.set_solid(s, nx, ny, nz)
    t1 = time.perf_counter()
    o = po.geom_solid(*s.triangles(), nx, ny, nz)
    t2 = time.perf_counter()
    print(f"cylinder {dims}: GPU {1e3 * (t1 - t0):.1f} ms (with transfers), oracle {1e3 * (t2 - t1):.0f} ms, "
          f"{int(g.sum())} solid", flush=True)
    assert g.sum() > 0
    assert np.array_equal(g, o)


def test_solid_backstep_bitwise():
    # backStepFlow's own STL on its full 1500 x 91 x 90 lattice; the oracle checks the first
    # 360 x 91 x 20 cells (a cell's flag depends only on its position and the triangles)
    nx, ny, nz = 1500, 91, 90
    s = geo.read_stl(STEP_STL, L_STEP)
    t0 = time.perf_counter()
    g = geo.set_solid(s, nx, ny, nz).reshape(nz, ny, nx)
    t1 = time.perf_counter()
    sub = (360, 91, 20)
    o = po.geom_solid(*s.triangles(), *sub).reshape(sub[2], sub[1], sub[0])
    print(f"backstep {nx}x{ny}x{nz}: GPU {1e3 * (t1 - t0):.1f} ms (with transfers), "
          f"{int(g.sum())} solid", flush=True)
    assert np.array_equal(g[:sub[2], :, :sub[0]], o)
    # the slab below y = 0.99 m, the gap above the ribs
    j_slab = int(0.99 / L_STEP) - 1
    assert g[5:85, 1:j_slab, 1:290].all()
    assert not g[:, 36:84, :].any()


def test_sdf_plane_and_cylinder():
    # a plane patch: the distance itself
    nx, ny, nz = 16, 16, 12
    c, nrm = _plane_patch(5.25)
    nei = np.zeros(nx * ny * nz, np.uint8)
    nei[8 + nx * (8 + ny * 5)] = 1
    patch = geo.Patch("plane", c, c, c, normal=nrm, c=c)
    g = geo.set_sdf(patch, np.full(nx * ny * nz, 10000.0, np.float32), nei, nx, ny, nz, dr=3.0)
    o = po.geom_sdf(c, nrm, np.full(nx * ny * nz, 10000.0, np.float32), nei, nx, ny, nz, dr=3.0)
    assert np.array_equal(g == 10000.0, o == 10000.0)
    assert np.abs(g - o).max() <= 1e-6

    # the cylinder: solid, neiSolid, then the IDW distance with the reference's dr and p
    nx, ny, nz = 240, 240, 12
    s = geo.read_stl(CYL_STL, L_CYL)
    solid = geo.set_solid(s, nx, ny, nz)
    nei = cases.neisolid(19, nx, ny, nz, solid)
    t0 = time.perf_counter()
    g = geo.set_sdf(s.patches[0], np.full(nx * ny * nz, 10000.0, np.float32), nei, nx, ny, nz)
    t1 = time.perf_counter()
    o = po.geom_sdf(s.patches[0].c, s.patches[0].normal, np.full(nx * ny * nz, 10000.0, np.float32), nei,
                    nx, ny, nz)
    t2 = time.perf_counter()
    near = o != 10000.0
    exact = float(np.mean(g[near] == o[near]))
    err = float(np.abs(g - o).max())
    print(f"cylinder sdf: {int(near.sum())} near cells, GPU {1e3 * (t1 - t0):.1f} ms, oracle {1e3 * (t2 - t1):.0f} ms,"
          f" bit-equal {exact:.4f}, max |diff| {err:.2e}", flush=True)
    assert np.array_equal(g == 10000.0, o == 10000.0)
    assert near.sum() > 10000
    assert exact > 0.9
    assert err <= 2e-5 * max(1.0, float(np.abs(o[near]).max()))
    # the distance is signed: negative inside the cylinder, positive in the fluid around it
    sol = solid.astype(bool)
    assert (g[near & sol] < 0.5).mean() > 0.95 and (g[near & ~sol] > -0.5).mean() > 0.95


def test_preprocess_backstep_les():
    # ns.cpp:188-226 with LES on: both patches' distances (the last patch wins in the cells
    # both reach), then the bounce-back faces' layers
    nx, ny, nz = 360, 91, 20
    s = geo.read_stl(STEP_STL, L_STEP)
    cin, un = cases.units(nx, ny, nz, float(L_STEP * (nx - 1)), 1.0, 0.05, 1.8e-4,
                          face_codes={"nyMin": ("Cyclic", "BounceBack", "Cyclic"),
                                      "nyMax": ("Cyclic", "BounceBack", "Cyclic")})
    (bnd, b1, b2, b3), _ = cases.build_faces(cin, un)
    t0 = time.perf_counter()
    g = geo.preprocess(s, 19, nx, ny, nz, True, b1, b2, b3)
    t1 = time.perf_counter()
    solid = po.geom_solid(*s.triangles(), nx, ny, nz)
    nei = cases.neisolid(19, nx, ny, nz, solid)
    sdf = np.full(nx * ny * nz, 10000.0, np.float32)
    for p in s.patches:
        sdf = po.geom_sdf(p.c, p.normal, sdf, nei, nx, ny, nz)
    geo.sdf_near_bwalls(sdf, b1, b2, b3, nx, ny, nz)
    print(f"backstep preprocess {nx}x{ny}x{nz}: GPU path {1e3 * (t1 - t0):.0f} ms", flush=True)
    assert np.array_equal(g["solid"], solid)
    assert np.array_equal(g["neiSolid"], nei)
    assert np.array_equal(g["sdf"] == 10000.0, sdf == 10000.0)
    assert np.abs(g["sdf"] - sdf).max() <= 2e-5 * max(1.0, float(np.abs(sdf[sdf != 10000.0]).max()))
