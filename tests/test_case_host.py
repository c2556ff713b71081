"""Host case setup (lbffs_amd/csrc/lbffs_case.cpp) against the reference's own host
fragments, compiled where they lie (oracle/_ref/libref_host.so: setBoundaryConditions.hpp,
initialization.hpp), bit for bit; setNeiSolid against a brute-force restatement."""
import itertools
import os

import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases
import pyoracle as po

needs_ref = pytest.mark.skipif(not os.path.exists(po.REF_HOST_SO), reason="oracle/_ref not built (make -C oracle ref)")

CODES = [0, 1, 3, 4, 5, 6, 7]


@needs_ref
@pytest.mark.parametrize("seed,shape", [(0, (5, 4, 3)), (1, (4, 6, 1)), (2, (3, 3, 5)), (3, (7, 5, 4))])
def test_face_codes_and_velocities_bit_identical(seed, shape):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    codes = rng.choice(CODES, size=(6, 3)).astype(np.int32)
    vel = rng.choice([0.0, 0.05, -0.03, 0.1], size=(6, 3)).astype(np.float32)
    # product path: physical velocities are divided by c inside lbffs_case_units
    face_codes = {f: tuple(int(x) for x in codes[n]) for n, f in enumerate(cases.FACES)}
    U0 = 0.05
    c_scale = 1.0 / U0
    face_vels = {f: tuple(float(x) * c_scale for x in vel[n]) for n, f in enumerate(cases.FACES)}
    cin, un = cases.units(nx, ny, nz, 1.0, 1.0, U0, 0.01, face_codes=face_codes, face_vels=face_vels)
    (bnd, b1, b2, b3), (u0, v0, w0) = cases.build_faces(cin, un)
    lat_vel = np.array(un.face_vel, dtype=np.float32).reshape(6, 3)
    r = po.ref_set_boundary_conditions(nx, ny, nz, codes, lat_vel)
    for name, mine, ref in zip(("b1", "b2", "b3", "u0", "v0", "w0"), (b1, b2, b3, u0, v0, w0), r):
        assert np.array_equal(mine.view(np.uint32) if mine.dtype == np.float32 else mine,
                              ref.view(np.uint32) if ref.dtype == np.float32 else ref), name
    # face flag of ns.cpp:164-177
    k, j, i = np.meshgrid(range(nz), range(ny), range(nx), indexing="ij")
    exp = ((i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1) | ((nz > 1) & ((k == 0) | (k == nz - 1))))
    assert np.array_equal(bnd, exp.reshape(-1).astype(np.int32))


@needs_ref
@pytest.mark.parametrize("vel", [(0.0, 0.0, 0.0), (0.05, -0.02, 0.01), (-0.1, 0.0, 0.03)])
def test_initial_populations_bit_identical(vel):
    f = cases.init_feq(19, 4, 3, 2, 1.0, *[np.float32(x) for x in vel])
    ref, ref_tmp = po.ref_initialization(4, 3, 2, 1.0, *vel)
    assert np.array_equal(f.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(ref, ref_tmp)


def test_units_follow_ns_cpp():
    """ns.cpp:104-128 in float arithmetic (omega0 through double, as ns.cpp:127 writes it)."""
    cin, un = cases.units(100, 100, 1, 1.0, 1.0, 0.05, 0.01, dpdx=0.08)
    f32 = np.float32
    L = f32(1.0) / f32(99)
    c = f32(1.0) / f32(0.05)
    nu = f32(0.01) / (L * c)
    assert un.L == L and un.c == c and un.nu == nu
    assert un.omega0 == f32(1.0 / (3.0 * float(nu) + 0.5))
    assert un.dpdx == f32(0.08) * (L / (c * c))


@pytest.mark.parametrize("seed,shape", [(0, (6, 5, 4)), (1, (7, 6, 1)), (2, (5, 5, 5))])
def test_neisolid_brute_force(seed, shape):
    """setNeiSolid (walls.hpp:187-216): a fluid cell off the lattice faces whose upwind
    neighbour (periodic wrap) is solid."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    solid = (rng.random(nx * ny * nz) < 0.3).astype(np.uint8)
    got = cases.neisolid(19, nx, ny, nz, solid)
    cx = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    cy = [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1]
    cz = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]
    exp = np.zeros_like(solid)
    for k, j, i in itertools.product(range(nz), range(ny), range(nx)):
        ic = i + nx * (j + ny * k)
        ii, jj, kk = 0 < i < nx - 1, 0 < j < ny - 1, 0 < k < nz - 1
        if not ((nz == 1 and ii and jj) or (ii and jj and kk)):
            continue
        if solid[ic]:
            continue
        for q in range(19):
            s = ((i - cx[q]) % nx) + nx * (((j - cy[q]) % ny) + ny * ((k - cz[q]) % nz))
            if solid[s]:
                exp[ic] = 1
    assert np.array_equal(got, exp)


def test_outlet_patch_sponge_shape():
    """ns.cpp:343-375: omega rises from 1/tau0 (tau0 = 0.55) at the patch to omega0 with a
    1 - cos ramp over spz cells."""
    nx, ny, nz, spz = 40, 3, 1, 10
    om0 = np.float32(1.0 / (3 * 0.02 + 0.5))
    omega = np.full(nx * ny * nz, om0, np.float32)
    L = lb.lib()
    rc = L.lbffs_case_outlet_sponge(nx, ny, nz, nx - 1, 1, 0, 3, 0, 0, spz, om0, lb._ptr(omega))
    assert rc == 0
    row = omega.reshape(nz, ny, nx)[0, 1]
    assert np.all(row[: nx - spz - 1] == om0)
    ramp = row[nx - spz - 1:]
    fx = 1 - np.cos(0.5 * np.pi * (np.arange(nx - spz - 1, nx)[::-1] - 0 if False else (nx - 1 - np.arange(nx - spz - 1, nx)) / spz))
    tau = fx * (1 / om0) + (1 - fx) * 0.55
    assert np.allclose(ramp[1:], (1 / tau)[1:], rtol=1e-6)
    assert abs(row[-1] - 1 / 0.55) < 1e-6


def test_outlet_patch_sponge_rejects_cells_outside_the_lattice():
    """The patch cell (i, j, k) must lie in the lattice: only the swept coordinate is
    clamped, so an index outside it would write omega out of bounds.  The array is left
    untouched when the call is refused."""
    nx, ny, nz, spz = 12, 4, 3, 3
    om0 = np.float32(1.5)
    L = lb.lib()
    for (i, j, k) in ((-1, 1, 1), (nx, 1, 1), (5, -1, 1), (5, ny, 1), (5, 1, -1), (5, 1, nz)):
        omega = np.full(nx * ny * nz, om0, np.float32)
        assert L.lbffs_case_outlet_sponge(nx, ny, nz, i, j, k, 3, 3, 3, spz, om0, lb._ptr(omega)) == -1
        assert np.all(omega == om0)
    omega = np.full(nx * ny * nz, om0, np.float32)
    assert L.lbffs_case_outlet_sponge(0, ny, nz, 0, 0, 0, 3, 0, 0, spz, om0, lb._ptr(omega)) == -1
    assert L.lbffs_case_outlet_sponge(nx, ny, nz, nx - 1, 1, 1, 3, 0, 0, spz, om0, lb._ptr(omega)) == 0
