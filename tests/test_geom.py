"""Geometry preprocessing on the CPU: the STL reader of the host mirror, the oracle's
setSolid / setSDF pinned to exact arithmetic, closed forms and analytic shapes, and the
host setSDFnearBwalls.

The oracle's solid test is pinned three ways:
  - meshes with small integer vertices, where every FP32 operation of the crossing test
    is exact up to the final divisions (whose rounding cannot move a quotient across 0
    or 1), so the reference's decisions -- including the 0/0 = NaN case of a ray lying in
    a triangle's plane, and edges shared by two triangles -- follow from exact rational
    arithmetic (fractions.Fraction);
  - the reference's own cylinder STL (flowAroundCylinder) against the analytic cylinder
    away from its surface;
  - the reference's backward-step boxes (backStepFlow) against the analytic boxes.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases
from lbffs_amd import geometry as geo

import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CYL_STL = os.path.join(GOLD, "cylinder.stl.gz")
STEP_STL = os.path.join(GOLD, "backStepBoxes.stl.gz")
L_CYL = float(np.float32(6.0) / np.float32(1199.0))    # flowAroundCylinder: Lx 6, nx 1200
L_STEP = float(np.float32(50.0) / np.float32(1499.0))  # backStepFlow: Lx 50, nx 1500


# ---- the STL reader -------------------------------------------------------------------

def test_tokens_follow_boost_split():
    assert geo.tokens("solid A\n") == ["solid", "A"]
    # an indented line starts with an empty token, trailing blanks end with one
    assert geo.tokens("  facet normal 0 0 1 \n") == ["", "facet", "normal", "0", "0", "1", ""]
    assert geo.tokens("a\r\n\n b") == ["a", "", "", "", "b"]


def _facet(n, a, b, c):
    f = lambda v: " ".join(str(x) for x in v)
    return (f"  facet normal {f(n)}\n    outer loop\n      vertex {f(a)}\n      vertex {f(b)}\n"
            f"      vertex {f(c)}\n    endloop\n  endfacet\n")


def test_parse_patches_scaling_and_centroids():
    text = ("solid first\n" + _facet((0, 0, 1), (0, 0, 0), (3, 0, 0), (0, 3, 0)) + "endsolid first\n"
            "solid second\n" + _facet((1, 0, 0), (1.5, 3, 6), (1.5, 0, 0), (1.5, 6, 0))
            + _facet((1, 0, 0), (0, 0, 0), (0, 0, 0), (0, 0, 0)) + "endsolid second\n")
    s = geo.parse_stl(text, 0.5)
    assert [p.name for p in s.patches] == ["first", "second"]
    assert [p.n for p in s.patches] == [1, 2]
    p0, p1 = s.patches
    # vertices divided by L, normals untouched, centroids (v0 + v1 + v2) / 3
    assert p0.v1[:, 0].tolist() == [6.0, 0.0, 0.0]
    assert p0.c[:, 0].tolist() == [2.0, 2.0, 0.0]
    assert p1.normal[:, 0].tolist() == [1.0, 0.0, 0.0]
    assert p1.v0[:, 0].tolist() == [3.0, 6.0, 12.0]
    v0, v1, v2 = s.triangles()
    assert v0.shape == (3, 3) and v0.dtype == np.float32


def test_parse_repeated_patch_name_joins_first_patch():
    # the reader moves to the next patch only when the name differs (stl.hpp:478-485)
    text = ("solid A\n" + _facet((0, 0, 1), (0, 0, 0), (1, 0, 0), (0, 1, 0)) + "endsolid A\n"
            "solid A\n" + _facet((0, 0, 1), (0, 0, 1), (1, 0, 1), (0, 1, 1)) + "endsolid A\n")
    s = geo.parse_stl(text, 1.0)
    assert [p.n for p in s.patches] == [2, 0]


def test_stof_is_strtof():
    # std::stof parses the decimal string straight to the nearest float
    assert geo._stof("0.1") == np.float32(0.1)
    assert geo._stof("1e-50") == np.float32(0.0)
    assert geo._stof("3.4028236e38") == np.float32(np.inf)
    with pytest.raises(ValueError):
        geo._stof("abc")


def test_missing_file_reads_empty(tmp_path):
    assert geo.read_stl(str(tmp_path / "none.stl"), 1.0).patches == []


def test_cylinder_fixture():
    s = geo.read_stl(CYL_STL, L_CYL)
    assert [p.name for p in s.patches] == ["Mesh"] and s.patches[0].n == 3664
    v0, v1, v2 = s.triangles()
    v = np.concatenate([v0, v1, v2], axis=1).astype(np.float64) * L_CYL
    r = np.hypot(v[0] - 1.0, v[1] - 1.0)
    assert r.max() <= 0.05 + 1e-5 and abs(v[2]).max() <= 0.05 + 1e-6


# ---- exact model of the crossing test -------------------------------------------------

def _det(a, b, c):
    return (a[0] * b[1] * c[2] + a[1] * b[2] * c[0] + a[2] * b[0] * c[1]
            - a[0] * b[2] * c[1] - a[1] * b[0] * c[2] - a[2] * b[1] * c[0])


def exact_is_cross(v0, v1, v2, i, j, k):
    """True/False, or None where only the sign of a zero would decide (z = +-inf)."""
    F = lambda x: Fraction(float(x)) if not isinstance(x, Fraction) else x
    a = [F(v1[t]) - F(v0[t]) for t in range(3)]
    b = [F(v2[t]) - F(v0[t]) for t in range(3)]
    c = [F(-1), F(0), F(0)]
    d = [F(i) - F(v0[0]), F(j) - F(v0[1]), F(k) - F(v0[2])]
    den, nxx, nyy, nzz = _det(a, b, c), _det(d, b, c), _det(a, d, c), _det(a, b, d)
    if den == 0:
        # x or y infinite fails a range test; x, y, z all 0/0 = NaN pass every test
        if nxx != 0 or nyy != 0:
            return False
        return True if nzz == 0 else None
    x, y, z = nxx / den, nyy / den, nzz / den
    return 0 <= x <= 1 and 0 <= y <= 1 and 0 <= x + y <= 1 and z >= 0


def exact_solid(v0, v1, v2, nx, ny, nz):
    """setSolid by exact arithmetic; -1 marks an undetermined cell."""
    n = v0.shape[1]
    V = [[[Fraction(float(v[a, t])) for a in range(3)] for t in range(n)] for v in (v0, v1, v2)]
    lo = [min([Fraction(0)] + [V[s][t][a] for s in range(3) for t in range(n)]) for a in range(3)]
    hi = [max([Fraction(0)] + [V[s][t][a] for s in range(3) for t in range(n)]) for a in range(3)]
    out = np.zeros(nx * ny * nz, np.int8)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if not all(lo[a] <= x <= hi[a] for a, x in enumerate((i, j, k))):
                    continue
                cnt, unknown = 0, False
                for t in range(n):
                    ys = [V[s][t][1] for s in range(3)]
                    zs = [V[s][t][2] for s in range(3)]
                    if not (min(ys) <= j <= max(ys) and min(zs) <= k <= max(zs)):
                        continue
                    r = exact_is_cross(V[0][t], V[1][t], V[2][t], i, j, k)
                    if r is None:
                        unknown = True
                    cnt += bool(r)
                out[i + nx * (j + ny * k)] = -1 if unknown else (cnt % 2)
    return out


def _mesh(tris):
    t = np.array(tris, np.float32)  # (n, 3 vertices, 3 axes)
    return [np.ascontiguousarray(t[:, s, :].T) for s in range(3)]


def box_mesh(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    P = lambda a, b, c: (x1 if a else x0, y1 if b else y0, z1 if c else z0)
    quads = [(P(0, 0, 0), P(0, 1, 0), P(0, 1, 1), P(0, 0, 1)), (P(1, 0, 0), P(1, 0, 1), P(1, 1, 1), P(1, 1, 0)),
             (P(0, 0, 0), P(0, 0, 1), P(1, 0, 1), P(1, 0, 0)), (P(0, 1, 0), P(1, 1, 0), P(1, 1, 1), P(0, 1, 1)),
             (P(0, 0, 0), P(1, 0, 0), P(1, 1, 0), P(0, 1, 0)), (P(0, 0, 1), P(0, 1, 1), P(1, 1, 1), P(1, 0, 1))]
    tris = []
    for a, b, c, d in quads:
        tris += [(a, b, c), (a, c, d)]
    return _mesh(tris)


def octa_mesh(c, r):
    cx, cy, cz = c
    X = [(cx + r, cy, cz), (cx - r, cy, cz)]
    Y = [(cx, cy + r, cz), (cx, cy - r, cz)]
    Z = [(cx, cy, cz + r), (cx, cy, cz - r)]
    return _mesh([(x, y, z) for x in X for y in Y for z in Z])


def tetra_mesh():
    p = [(1, 1, 1), (8, 2, 1), (3, 7, 2), (4, 3, 6)]
    return _mesh([(p[0], p[1], p[2]), (p[0], p[1], p[3]), (p[0], p[2], p[3]), (p[1], p[2], p[3])])


ALIGNED = {
    "box": (lambda: box_mesh((2, 1, 0), (7, 5, 3)), (10, 7, 5)),
    "octahedron": (lambda: octa_mesh((5, 4, 3), 3), (11, 9, 7)),
    "tetrahedron": (tetra_mesh, (10, 9, 8)),
    "two_boxes": (lambda: [np.concatenate(p, axis=1) for p in zip(box_mesh((1, 1, 1), (4, 4, 4)),
                                                                   box_mesh((4, 1, 1), (7, 3, 3)))], (9, 6, 6)),
}


@pytest.mark.parametrize("name", list(ALIGNED))
def test_oracle_solid_matches_exact_arithmetic(name):
    make, (nx, ny, nz) = ALIGNED[name]
    v0, v1, v2 = make()
    ex = exact_solid(v0, v1, v2, nx, ny, nz)
    orc = po.geom_solid(v0, v1, v2, nx, ny, nz)
    known = ex >= 0
    assert known.mean() > 0.9
    assert np.array_equal(orc[known], ex[known].astype(np.uint8)), name
    assert orc.sum() > 0


def test_exact_model_sees_the_degenerate_cases():
    # the box's faces lie on lattice planes: rays in the planes y = 1, y = 5, z = 0, z = 3
    # meet triangles edge-on (0/0), and cells on x = 2 sit on a face
    v0, v1, v2 = box_mesh((2, 1, 0), (7, 5, 3))
    nan_hits = sum(exact_is_cross(v0[:, t], v1[:, t], v2[:, t], 0, 1, 1) is True for t in range(12))
    assert nan_hits > 0
    ex = exact_solid(v0, v1, v2, 10, 7, 5)
    interior = [(i, j, k) for i in range(3, 7) for j in range(2, 5) for k in range(1, 3)]
    assert all(ex[i + 10 * (j + 7 * k)] == 1 for i, j, k in interior)


def test_oracle_solid_invert_and_keep():
    v0, v1, v2 = box_mesh((2, 1, 0), (7, 5, 3))
    rng = np.random.default_rng(3)
    init = rng.integers(0, 3, 10 * 7 * 5).astype(np.uint8)
    kept = po.geom_solid(v0, v1, v2, 10, 7, 5, solid=init)
    plain = po.geom_solid(v0, v1, v2, 10, 7, 5)
    assert np.array_equal(kept, np.where(plain == 1, 1, init))
    inv = po.geom_solid(v0, v1, v2, 10, 7, 5, invert=True, solid=init)
    assert np.array_equal(inv, np.where(kept == 1, 0, 1))


def test_oracle_solid_cylinder_analytic():
    s = geo.read_stl(CYL_STL, L_CYL)
    v0, v1, v2 = s.triangles()
    nx, ny, nz = 240, 240, 12
    sol = po.geom_solid(v0, v1, v2, nx, ny, nz).reshape(nz, ny, nx)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rho = np.hypot(i - 1.0 / L_CYL, j - 1.0 / L_CYL)
    # the mesh's side vertices lie between radii 0.04848 and 0.05 m (a faceted, uneven
    # surface), its caps at z = +-0.05 m
    R_in, R_out, H = 0.04848 / L_CYL, 0.05 / L_CYL, 0.05 / L_CYL
    inside = (rho < R_in - 0.05) & (k < H - 0.05)
    outside = (rho > R_out + 0.05) | (k > H + 0.05)
    assert sol[inside].all() and not sol[outside].any()
    assert inside.sum() > 2000


def test_oracle_solid_backstep_analytic():
    # backStepBoxes.stl: patch Bottom is a slab y < 0.99 m for x < 10 m, open at the domain
    # walls and closed by its face x = 10 m, with ribs up to y = 1.14 m near x = 1.1-1.55 m;
    # patch Top holds ribs hanging from y = 3 m down to 2.85 m at the same x
    s = geo.read_stl(STEP_STL, L_STEP)
    v0, v1, v2 = s.triangles()
    nx, ny, nz = 360, 91, 20
    sol = po.geom_solid(v0, v1, v2, nx, ny, nz).reshape(nz, ny, nx)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = i * L_STEP, j * L_STEP, k * L_STEP
    m = 0.01  # metres, a third of a cell
    slab = (x > m) & (x < 10 - m) & (y > m) & (y < 0.99 - m) & (z > m)
    gap = (y > 1.14 + m) & (y < 2.85 - m)
    beyond = (x > 10 + m) & (y < 2.85 - m)
    assert slab.sum() > 100000 and gap.sum() > 100000
    assert sol[slab].all()
    assert not sol[gap | beyond].any()


# ---- the signed distance --------------------------------------------------------------

def _plane_patch(z0, n_side=12, spacing=1.0, offset=(2.3, 1.7)):
    """Triangles of the plane z = z0 with normal +z (two per square)."""
    tris = []
    for a in range(n_side):
        for b in range(n_side):
            x, y = offset[0] + a * spacing, offset[1] + b * spacing
            p = [(x, y, z0), (x + spacing, y, z0), (x + spacing, y + spacing, z0), (x, y + spacing, z0)]
            tris += [(p[0], p[1], p[2]), (p[0], p[2], p[3])]
    v0, v1, v2 = _mesh(tris)
    c = ((v0 + v1) + v2) / np.float32(3.0)
    nrm = np.zeros_like(c)
    nrm[2] = 1.0
    return c, nrm


def test_oracle_sdf_plane_is_the_distance():
    nx, ny, nz = 16, 16, 12
    z0 = 5.25
    c, nrm = _plane_patch(z0)
    nei = np.zeros(nx * ny * nz, np.uint8)
    nei[8 + nx * (8 + ny * 5)] = 1
    sdf = np.full(nx * ny * nz, 10000.0, np.float32)
    out = po.geom_sdf(c, nrm, sdf, nei, nx, ny, nz, dr=3.0).reshape(nz, ny, nx)
    k = np.arange(nz)[:, None, None] * np.ones((1, ny, nx))
    box = np.zeros((nz, ny, nx), bool)
    box[2:9, 5:12, 5:12] = True  # int(dr) = 3 around (8, 8, 5)
    changed = out != 10000.0
    assert not changed[~box].any()
    # the layers k = 3..7 (|k - z0| <= 2.25) have centroids within dr; k = 2 has none
    inner = box.copy()
    inner[[2, 8]] = False
    assert changed[inner].all() and not changed[2].any()
    # the weighted mean of equal plane distances is that distance
    assert np.allclose(out[changed], (k - z0)[changed], rtol=0, atol=2e-6)


def test_oracle_sdf_weighted_mean_closed_form():
    # two triangles, centroids at distances r1, r2 of a cell; sd from the normals
    c = np.array([[3.0, 5.5], [2.0, 2.0], [2.0, 2.0]], np.float32)
    nrm = np.array([[1.0, 0.6], [0.0, 0.0], [0.0, 0.8]], np.float32)
    nx, ny, nz = 8, 5, 5
    nei = np.zeros(nx * ny * nz, np.uint8)
    nei[4 + nx * (2 + ny * 2)] = 1
    out = po.geom_sdf(c, nrm, np.full(nx * ny * nz, 7.0, np.float32), nei, nx, ny, nz, dr=2.0, p=7.0)
    ic = 4 + nx * (2 + ny * 2)  # cell (4, 2, 2)
    d1 = np.array([1.0, 0.0, 0.0]); d2 = np.array([-1.5, 0.0, 0.0])
    w1, w2 = 1.0 ** -7, 1.5 ** -7
    sd1, sd2 = d1 @ [1.0, 0.0, 0.0], d2 @ [0.6, 0.0, 0.8]
    assert abs(out[ic] - (w1 * sd1 + w2 * sd2) / (w1 + w2)) < 1e-6
    # a near cell with no centroid within dr keeps sdfIni; cells outside the box keep theirs
    far = 6 + nx * (4 + ny * 4)
    assert out[far] == np.float32(10000.0)
    assert out[0] == np.float32(7.0)


# ---- setSDFnearBwalls (host) ----------------------------------------------------------

def test_sdf_near_bwalls_layers():
    nx, ny, nz = 9, 7, 6
    N = nx * ny * nz
    b1 = np.zeros(N, np.int32)
    b2 = np.zeros(N, np.int32)
    b3 = np.zeros(N, np.int32)
    idx = lambda i, j, k: i + nx * (j + ny * k)
    for j in range(ny):
        for k in range(nz):
            b1[idx(0, j, k)] = 1          # BounceBack on nxMin only
    for i in range(nx):
        for j in range(ny):
            b3[idx(i, j, nz - 1)] = 1     # and on nzMax
    sdf = np.full(N, 10000.0, np.float32)
    geo.sdf_near_bwalls(sdf, b1, b2, b3, nx, ny, nz, dr=3.0)
    s = sdf.reshape(nz, ny, nx)
    want = np.full((nz, ny, nx), 10000.0, np.float32)
    for d in range(4):
        want[:, :, d] = 0.5 + d                  # x layers 0..3
    for d in range(4):
        want[nz - 1 - d, :, :] = 0.5 + d         # z layers, written after x
    assert np.array_equal(s, want)


def test_sdf_near_bwalls_skips_the_opposite_face():
    nx, ny, nz = 3, 2, 1
    N = nx * ny * nz
    b1 = np.ones(N, np.int32)
    z = np.zeros(N, np.int32)
    sdf = np.full(N, -1.0, np.float32)
    geo.sdf_near_bwalls(sdf, b1, z, z, nx, ny, nz, dr=3.0)
    # from i=0: layers i < nx-1 -> 0, 1; from i=nx-1: cells nx-1-d > 0 -> 2 (d=0), 1 (d=1)
    assert sdf.reshape(ny, nx)[0].tolist() == [0.5, 1.5, 0.5]


# ---- q-distances from the surface (FP64) ---------------------------------------------

def uv_sphere(c, R, nu=96, nv=48):
    """Closed triangle mesh of a sphere (latitude-longitude), outward winding."""
    cx, cy, cz = c
    th = np.linspace(0.0, np.pi, nv + 1)
    ph = np.linspace(0.0, 2 * np.pi, nu + 1)
    P = lambda a, b: (cx + R * np.sin(th[a]) * np.cos(ph[b]), cy + R * np.sin(th[a]) * np.sin(ph[b]),
                      cz + R * np.cos(th[a]))
    tris = []
    for a in range(nv):
        for b in range(nu):
            p00, p01, p10, p11 = P(a, b), P(a, b + 1), P(a + 1, b), P(a + 1, b + 1)
            if a > 0:
                tris.append((p00, p10, p01))
            if a < nv - 1:
                tris.append((p01, p10, p11))
    return _mesh(tris)


def test_oracle_qlinks_plane_is_exact():
    # plane y = 5.3 across the lattice; fluid row j = 5 next to the solid rows j >= 6
    nx, ny, nz = 8, 10, 6
    y0 = 5.3
    v0, v1, v2 = _mesh([((-3, y0, -3), (12, y0, -3), (12, y0, 10)), ((-3, y0, -3), (12, y0, 10), (-3, y0, 10))])
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 6:, :] = 1
    nei = np.zeros((nz, ny, nx), np.uint8)
    nei[:, 5, :] = 1
    ql, miss = po.geom_qlinks(19, v0, v1, v2, nx, ny, nz, solid.reshape(-1), nei.reshape(-1))
    assert miss == 0
    ql = ql.reshape(19, nz, ny, nx)
    cy = np.array(po.ORC19_C[1])
    want = float(np.float32(y0)) - 5.0  # the link x - t c_q reaches the plane at t = y0 - 5
    for d in range(1, 19):
        if cy[d] == -1:   # upwind x - c_d is the solid row above
            assert np.allclose(ql[d, :, 5, :], want, rtol=0, atol=1e-7), d
        else:
            assert np.all(ql[d] == 0.5), d


@pytest.mark.parametrize("q", [19, 27])
def test_oracle_qlinks_sphere_analytic(q):
    nx = ny = nz = 22
    c, R = (10.3, 10.6, 10.2), 6.4
    v0, v1, v2 = uv_sphere(c, R)
    # solid cells inside the sphere itself (the ray-parity voxeliser of the reference flips
    # whole row segments where a row grazes a shared edge of this mesh)
    k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt((i3 - c[0]) ** 2 + (j3 - c[1]) ** 2 + (k3 - c[2]) ** 2).reshape(-1)
    solid = (r < R).astype(np.uint8)
    nei = cases.neisolid(q, nx, ny, nz, solid)
    ql, miss = po.geom_qlinks(q, v0, v1, v2, nx, ny, nz, solid, nei)
    # a solid cell closer to the sphere than the chords (0.014 cells) may lie outside the
    # mesh: its links meet no triangle
    assert miss <= 10
    C3 = [np.array(a) for a in (po.ORC27_C if q == 27 else po.ORC19_C)]
    ql = ql.reshape(q, -1)
    ic = np.nonzero(nei)[0]
    i, j, k = ic % nx, ic // nx % ny, ic // (nx * ny)
    checked = 0
    for d in range(1, q):
        dx, dy, dz = C3[0][d], C3[1][d], C3[2][d]
        up = (i - dx) + nx * ((j - dy) + ny * (k - dz))
        m = solid[up] == 1
        # |p - t c|^2 = R^2 with p = x - centre: the smaller root
        px, py, pz = i[m] - c[0], j[m] - c[1], k[m] - c[2]
        a = dx * dx + dy * dy + dz * dz
        b = -(px * dx + py * dy + pz * dz)
        cc = px * px + py * py + pz * pz - R * R
        t = (-b - np.sqrt(np.maximum(b * b - a * cc, 0))) / a
        # the mesh's facets lie inside the sphere by up to ~0.01 cells; a link that meets the
        # surface at grazing incidence turns that into a larger change of t
        hit = ql[d, ic[m]] != 0.5
        err = np.abs(ql[d, ic[m]][hit] - t[hit])
        assert err.max() < 0.1 and err.mean() < 0.01, (d, err.max(), err.mean())
        checked += int(hit.sum())
    assert checked > 500


def test_procedural_body_mesh():
    # closed and consistently wound, the triangle count of the latitude-longitude mesh,
    # the volume and frontal area of the sphere it perturbs, reproducible from its seed
    c, R = (20.3, 18.4, 16.2), 9.0
    s = geo.procedural_body(c, R, n_lat=30, n_lon=48, amp=0.02, seed=7)
    p = s.patches[0]
    assert p.n == 2 * 48 * 29
    tri = np.stack([p.v0.T, p.v1.T, p.v2.T], axis=1).astype(np.float64)
    _, ids = np.unique(tri.reshape(-1, 3), axis=0, return_inverse=True)
    ids = ids.reshape(-1, 3)
    edges = np.concatenate([ids[:, [0, 1]], ids[:, [1, 2]], ids[:, [2, 0]]])
    fwd = {tuple(e) for e in edges}
    assert len(fwd) == len(edges)                        # every directed edge once
    assert all((b, a) in fwd for a, b in fwd)            # and its reverse once: closed
    cc = np.array(c)
    vol = np.einsum("ij,ij->i", tri[:, 0] - cc, np.cross(tri[:, 1] - cc, tri[:, 2] - cc)).sum() / 6
    assert abs(vol / (4 / 3 * np.pi * R ** 3) - 1) < 0.03      # outward winding, sphere-like
    assert abs(geo.frontal_area(s) / (np.pi * R * R) - 1) < 0.03
    again = geo.procedural_body(c, R, n_lat=30, n_lon=48, amp=0.02, seed=7).patches[0]
    other = geo.procedural_body(c, R, n_lat=30, n_lon=48, amp=0.02, seed=8).patches[0]
    assert np.array_equal(again.v0, p.v0) and not np.array_equal(other.v0, p.v0)
    # the normals are unit and point away from the centre
    assert np.allclose(np.linalg.norm(p.normal, axis=0), 1.0, atol=1e-6)
    assert (np.einsum("ij,ij->j", p.normal.astype(np.float64), p.c - cc[:, None]) > 0).all()


def test_procedural_body_reference_flags():
    # the reference's crossing test (oracle) on the procedural mesh gives the body: solid
    # well inside the perturbed radius, fluid well outside, no streaks of broken parity
    nx, ny, nz = 40, 36, 32
    c, R, amp = (20.3, 18.4, 16.2), 9.0, 0.05
    s = geo.procedural_body(c, R, n_lat=30, n_lon=48, amp=amp, seed=3)
    o = po.geom_solid(*s.triangles(), nx, ny, nz).reshape(nz, ny, nx)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2)
    assert o[r < R * (1 - amp) - 0.5].all()
    assert not o[r > R * (1 + amp) + 0.5].any()
    assert 0.9 < o.sum() / (4 / 3 * np.pi * R ** 3) < 1.1


def test_write_stl_round_trip(tmp_path):
    s = geo.procedural_body((5.5, 6.25, 7.0), 3.0, n_lat=10, n_lon=14, seed=11)
    path = str(tmp_path / "body.stl")
    geo.write_stl(s, path)
    r = geo.read_stl(path, 1.0)
    assert [p.name for p in r.patches] == ["body"]
    a, b = s.patches[0], r.patches[0]
    for k in ("v0", "v1", "v2", "normal", "c"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def _patch(tris, normal):
    """A patch from a list of triangles (three vertices each) with one normal for all."""
    v = [np.array([t[k] for t in tris], np.float32).T.copy() for k in range(3)]
    n = np.repeat(np.array(normal, np.float32)[:, None], len(tris), axis=1)
    p = geo.Patch("p", v[0], v[1], v[2], normal=n)
    p.c = ((p.v0 + p.v1) + p.v2) / np.float32(3.0)
    return p


def _in_tri_exact(x, y, x1, y1, x2, y2):
    # stl.hpp:95-111 pointInElement in exact rational arithmetic
    from fractions import Fraction as Fr
    x, y, x1, y1, x2, y2 = (Fr(float(a)) for a in (x, y, x1, y1, x2, y2))
    d1, d2 = x1 * y2 - x2 * y1, x2 * y1 - x1 * y2
    if d1 == 0:
        return False
    s, t = (x * y2 - x2 * y) / d1, (x * y1 - x1 * y) / d2
    return 0 <= s + t <= 1 and 0 <= s <= 1 and 0 <= t <= 1


def test_find_points_rectangle_on_each_face():
    # two right triangles (legs 4 cells) spanning a 4 x 4 square on a lattice face: every
    # lattice point of the closed square is listed, the points of the shared diagonal twice;
    # s and t are multiples of 1/4 there, exact in FP32, so the exact rule decides
    nx, ny, nz = 9, 8, 7
    for ax, face in ((0, 0), (0, nx - 1), (1, 0), (1, ny - 1), (2, 0), (2, nz - 1)):
        u, w = (1, 2) if ax == 0 else ((0, 2) if ax == 1 else (0, 1))
        def P(a, b):
            p = [0.0, 0.0, 0.0]
            p[ax], p[u], p[w] = float(face), a, b
            return tuple(p)
        a0, b0 = 1.0, 2.0
        tris = [(P(a0, b0), P(a0 + 4, b0), P(a0 + 4, b0 + 4)), (P(a0, b0), P(a0 + 4, b0 + 4), P(a0, b0 + 4))]
        nrm = [0.0, 0.0, 0.0]
        nrm[ax] = -1.0 if face == 0 else 1.0
        pts = geo.find_points(_patch(tris, nrm), nx, ny, nz)
        dims = (nx, ny, nz)
        want = []
        for tri in tris:
            (pu0, pw0), (pu1, pw1), (pu2, pw2) = ((t[u], t[w]) for t in tri)
            for a in range(dims[u]):
                for b in range(dims[w]):
                    if _in_tri_exact(a - pu0, b - pw0, pu1 - pu0, pw1 - pw0, pu2 - pu0, pw2 - pw0):
                        idx = [0, 0, 0]
                        idx[ax], idx[u], idx[w] = face, a, b
                        want.append(idx[0] + nx * (idx[1] + ny * idx[2]))
        assert pts.tolist() == want, (ax, face)
        assert len(set(pts.tolist())) == 25 and len(pts) == 25 + 5  # 5x5 points, diagonal twice


def test_find_points_plane_selection():
    nx, ny, nz = 9, 8, 7
    # a triangle whose centroid is inside the lattice, not on a face: no points
    tri = [((4.0, 1.0, 1.0), (4.0, 5.0, 1.0), (4.0, 5.0, 5.0))]
    assert geo.find_points(_patch(tri, (1.0, 0.0, 0.0)), nx, ny, nz).size == 0
    # centroid just outside the lattice (x in (-1, 0]): still the i = 0 face
    tri = [((-0.4, 1.0, 1.0), (-0.4, 5.0, 1.0), (-0.4, 5.0, 5.0))]
    assert geo.find_points(_patch(tri, (-1.0, 0.0, 0.0)), nx, ny, nz).size > 0
    # a component below 1 in magnitude truncates to 0 in the reference's abs(): the normal
    # (0, 0.6, 0.8) selects the x plane (first axis), so a patch on the k = 0 face with that
    # normal lists nothing
    tri = [((1.0, 1.0, 0.0), (5.0, 1.0, 0.0), (5.0, 5.0, 0.0))]
    assert geo.find_points(_patch(tri, (0.0, 0.6, 0.8)), nx, ny, nz).size == 0
    assert geo.find_points(_patch(tri, (0.0, 0.0, -1.0)), nx, ny, nz).size > 0


def test_inlet_patch_reproduces_the_face_setup():
    # an inlet patch covering the whole nxMin face (two triangles beyond its corners) with
    # FixedVelocity and the inlet velocity gives the face codes and u0 that the face setup
    # gives for a FixedVelocity nxMin face; an outlet patch on nxMax with a sponge changes
    # omega only within spz cells of that face
    nx, ny, nz = 12, 8, 6
    ref = cases.cylinder_channel(nx=nx, ny=ny, nz=nz, D=2, cx=6, cy=4, spz_cells=2, noise=0.0)
    base_codes = {f: ("Cyclic", "Cyclic", "Cyclic") for f in cases.FACES}
    base_codes["nyMin"] = ("Cyclic", "BounceBack", "Cyclic")
    base_codes["nyMax"] = ("Cyclic", "BounceBack", "Cyclic")
    c = cases.make_case("patched", nx, ny, nz, 1.0, 1.0, ref.meta["u_in"], ref.units["nu"] * (nx - 1) * ref.meta["u_in"],
                        base_codes, q=19)
    lo, hi = -0.5, max(ny, nz) * 2.0
    tris = [((0.0, lo, lo), (0.0, hi, lo), (0.0, hi, hi)), ((0.0, lo, lo), (0.0, hi, hi), (0.0, lo, hi))]
    pts = geo.find_points(_patch(tris, (-1.0, 0.0, 0.0)), nx, ny, nz)
    assert set(pts.tolist()) == {nx * (j + ny * k) for j in range(ny) for k in range(nz)}
    # the face setup leaves the y-wall edge cells' x code as it is: patch only the others
    keep = np.isin(np.arange(nx * ny * nz), pts) & (ref.arrays["boundary1"] == lb.BC["FixedVelocity"])
    cases.apply_patch(c, np.nonzero(keep)[0], ("FixedVelocity", "Cyclic", "Cyclic"), vel=(ref.meta["u_in"], 0.0, 0.0))
    i0 = np.arange(nx * ny * nz) % nx == 0
    assert np.array_equal(c.arrays["boundary1"][i0], ref.arrays["boundary1"][i0])
    # the face setup gives the velocity to the face's interior cells (not its edges)
    ic = np.arange(nx * ny * nz)
    j, k = (ic // nx) % ny, ic // (nx * ny)
    inner = i0 & (j > 0) & (j < ny - 1) & (k > 0) & (k < nz - 1)
    assert np.array_equal(c.arrays["u0"][inner], ref.arrays["u0"][inner])
    assert (ref.arrays["u0"][inner] > 0).all()
    xe = float(nx - 1)
    tris_out = [((xe, lo, lo), (xe, hi, lo), (xe, hi, hi)), ((xe, lo, lo), (xe, hi, hi), (xe, lo, hi))]
    om0 = c.arrays["omega"].copy()
    out = geo.find_points(_patch(tris_out, (1.0, 0.0, 0.0)), nx, ny, nz)
    cases.apply_patch(c, out, ("Outlet", "Cyclic", "Cyclic"), spz=3)
    changed = (c.arrays["omega"] != om0).reshape(nz, ny, nx)
    assert changed.any() and not changed[:, :, : nx - 3].any()
    assert (c.arrays["boundary1"].reshape(nz, ny, nx)[:, :, nx - 1] == lb.BC["Outlet"]).all()
