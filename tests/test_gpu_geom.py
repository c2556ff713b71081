"""Geometry preprocessing on the GPU (include/lbffs_geom.h) against the oracle.

Bar: the solid flags are bit-identical to the oracle's restatement of setSolid (and, on
the integer meshes, to exact arithmetic); the wall distance of setSDF agrees to FP32
rounding -- the reference takes r^-p from the host's powf, the kernel rounds it once from
FP64, so a few terms may differ in their last bit; the FP64 ray-triangle q-distances are
bit-identical to the oracle's brute force.
"""
import time

import numpy as np
import pytest

import lbffs_amd as lb
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


@pytest.mark.parametrize("q", [19, 27])
def test_qlinks_bitwise(q):
    # FP64 ray-triangle q-distances: GPU kernel (bucketed triangles) against the oracle's
    # brute force over every triangle, bit for bit
    from test_geom import uv_sphere
    nx = ny = nz = 24
    c, R = (11.3, 11.6, 12.2), 7.4
    v0, v1, v2 = uv_sphere(c, R)
    solid = geo.solid_from_triangles(v0, v1, v2, nx, ny, nz)
    nei = cases.neisolid(q, nx, ny, nz, solid)
    t0 = time.perf_counter()
    g, gm = geo.qlinks(geo.STL([geo.Patch("s", v0, v1, v2)]), q, nx, ny, nz, solid, nei)
    t1 = time.perf_counter()
    o, om = po.geom_qlinks(q, v0, v1, v2, nx, ny, nz, solid, nei)
    print(f"sphere q{q}: {int((g != 0.5).sum())} links, GPU {1e3 * (t1 - t0):.1f} ms, misses {gm}", flush=True)
    assert gm == om
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    assert (g != 0.5).sum() > 500


def test_qlinks_cylinder_stl():
    # the reference's cylinder STL (its side surface only: an open tube, as the 2-D case
    # needs), raised by 12 cells so that its ends lie inside the lattice: q-distances on
    # the GPU equal the oracle's, and every solid link through the side meets the mesh --
    # the links that miss are those through the open ends (fluid layers k = 2 and k = 22)
    nx, ny, nz = 240, 240, 26
    s = geo.read_stl(CYL_STL, L_CYL)
    for p in s.patches:
        for v in (p.v0, p.v1, p.v2):
            v[2] += np.float32(12.0)
    solid = geo.set_solid(s, nx, ny, nz)
    nei = cases.neisolid(27, nx, ny, nz, solid)
    g, gm = geo.qlinks(s, 27, nx, ny, nz, solid, nei, init=-1.0)
    o, om = po.geom_qlinks(27, *s.triangles(), nx, ny, nz, solid, nei, init=-1.0)
    n_links = int((g >= 0).sum())
    print(f"cylinder STL q27: {n_links} links, misses {gm}", flush=True)
    assert gm == om and np.array_equal(g.view(np.uint32), o.view(np.uint32))
    assert n_links > 5000
    # misses only through the open ends: recount the solid links off those layers
    g = g.reshape(27, nz, ny, nx)
    nei3 = nei.reshape(nz, ny, nx).astype(bool)
    sol3 = solid.reshape(nz, ny, nx)
    cx, cy, cz = (np.array(a) for a in po.ORC27_C)
    for d in range(1, 27):
        up = np.roll(np.roll(np.roll(sol3, cz[d], axis=0), cy[d], axis=1), cx[d], axis=2) == 1
        side = nei3 & up
        side[:3] = False
        side[22:] = False
        assert (g[d][side] >= 0).all(), d


def _small_bluff(**kw):
    return cases.bluff_body(nx=72, ny=40, nz=36, D=18.0, centre=(24.3, 20.4, 18.2), n_lat=40, n_lon=64,
                            spz_cells=8, Re=100.0, **kw)


@pytest.mark.parametrize("storage", [lb.STORE_F32, lb.STORE_BF16DEV, lb.STORE_F64])
def test_qlinks_into_lattice_equal_host_route(storage):
    # the q-distances cast straight into a context equal the host-array route bit for bit
    # (rounded to FP32 in every storage form), and the solid flags of the procedural body
    # equal the oracle's restatement of setSolid
    c = _small_bluff(qlinks_on_device=False)
    stl = c.meta["stl"]
    o = po.geom_solid(*stl.triangles(), c.nx, c.ny, c.nz)
    assert np.array_equal(c.arrays["solid"], o)
    host = c.arrays.pop("qlink")
    with lb.Lattice(c.nx, c.ny, c.nz, q=27, storage=storage, sgs=c.sgs, collision=lb.COLL_CUMULANT,
                    omega0=c.omega0) as lat:
        c.upload_to(lat)
        miss = geo.qlinks_into(lat, stl)
        d = lat.download("qlink")
    print(f"bluff q27 storage {storage}: {int((host != 0.5).sum())} links, misses {miss}", flush=True)
    assert miss == c.meta["qlink_misses"]
    assert np.array_equal(d, host.astype(d.dtype))
    assert (host != 0.5).sum() > 3000
