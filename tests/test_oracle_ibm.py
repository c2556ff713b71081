"""The oracle's immersed-boundary kernels (k_Gwall, k_Gibm, k_Force and Urot/Xrot/icBox,
D3Q19_opencl/kernel.cl:1112-1279, 3648-3807), pinned by what the mathematics fixes:
icBox by its definition, trilinear interpolation by exactness on linear fields, the
spread by the partition of unity, the motion by closed-form rotations (including the
reference's axis dot product OX_x*e0_z, K19:1213/1256, worked by hand for a z axis), and
the forcing by its moments (mass unchanged, momentum + rho G)."""
import numpy as np
import pytest

import pyoracle as po

P = po._p


def motion(trans=(0, 0, 0), center=(0, 0, 0), axis=(0, 0, 0), omega=0.0):
    return np.array([*trans, *center, *axis, omega], np.float64)


def test_icbox_definition():
    L = po.orc()
    nx, ny, nz = 4, 3, 5
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                ic = i + nx * (j + ny * k)
                for b, (di, dj, dk) in enumerate([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1),
                                                   (0, 1, 1), (1, 1, 1)]):
                    ii, jj, kk = i + di, j + dj, k + dk
                    exp = -1 if (ii >= nx or jj >= ny or kk >= nz) else ii + nx * (jj + ny * kk)
                    assert L.orc_ibm_icbox(ic, b, nx, ny, nz) == exp


def _gwall(m, xyz, nx, ny, nz, u, v, w):
    n = len(xyz) // 3
    G = np.zeros((3, n))
    po.orc().orc_ibm_gwall_f64(P(m), n, P(np.ascontiguousarray(xyz, np.float64)), nx, ny, nz, P(u), P(v), P(w),
                              P(G[0]), P(G[1]), P(G[2]))
    return G


def test_trilinear_interpolation_is_exact_on_linear_fields():
    rng = np.random.default_rng(0)
    nx, ny, nz = 6, 5, 7
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    a = rng.standard_normal((3, 4))
    u, v, w = [np.ascontiguousarray((a[c, 0] + a[c, 1] * i + a[c, 2] * j + a[c, 3] * k).reshape(-1), np.float64)
               for c in range(3)]
    xyz = np.column_stack([rng.uniform(0, nx - 1, 50), rng.uniform(0, ny - 1, 50), rng.uniform(0, nz - 1, 50)])
    # a pure translation (axis 0): U_wall = trans; G = trans - u(x)
    m = motion(trans=(0.01, -0.02, 0.03), center=(-50, -50, -50))
    G = _gwall(m, xyz.reshape(-1), nx, ny, nz, u, v, w)
    for c in range(3):
        exact = a[c, 0] + a[c, 1] * xyz[:, 0] + a[c, 2] * xyz[:, 1] + a[c, 3] * xyz[:, 2]
        assert np.allclose(G[c], m[c] - exact, atol=1e-12)


def test_spread_is_a_partition_of_unity_and_markers_translate():
    rng = np.random.default_rng(1)
    nx, ny, nz = 8, 7, 6
    N = nx * ny * nz
    n = 40
    xyz = np.column_stack([rng.uniform(0, nx - 1, n), rng.uniform(0, ny - 1, n), rng.uniform(0, nz - 1, n)]).reshape(-1)
    Gm = rng.standard_normal((3, n))
    Gc = np.zeros((3, N))
    m = motion(trans=(0.25, 0.0, -0.125), center=(-50, -50, -50))
    x2 = xyz.copy()
    po.orc().orc_ibm_spread_move_f64(P(m), n, P(x2), nx, ny, nz, P(Gm[0]), P(Gm[1]), P(Gm[2]), P(Gc[0]), P(Gc[1]),
                                     P(Gc[2]))
    assert np.allclose(Gc.sum(1), Gm.sum(1), atol=1e-12)  # every marker's weights sum to 1
    assert np.allclose(x2.reshape(-1, 3), xyz.reshape(-1, 3) + m[:3], atol=1e-12)


def test_rotation_about_z_closed_forms():
    L = po.orc()
    cx, cy, cz, Rr, th = 3.0, 4.0, 2.0, 1.5, 0.2
    m = motion(center=(cx, cy, cz), axis=(0, 0, 1), omega=th)
    o = [np.zeros(1) for _ in range(3)]
    out = o
    # marker with OX_x = 0: the quirk term vanishes, a true rotation
    L.orc_ibm_xrot_f64(P(m), cx, cy + Rr, cz, P(o[0]), P(o[1]), P(o[2]))
    assert np.allclose([o[0][0], o[1][0], o[2][0]], [cx - Rr * np.sin(th), cy + Rr * np.cos(th), cz], atol=1e-12)
    L.orc_ibm_urot_f64(P(m), cx, cy + Rr, cz, P(o[0]), P(o[1]), P(o[2]))
    assert np.allclose([o[0][0], o[1][0], o[2][0]], [-Rr * th, 0, 0], atol=1e-12)
    # marker with OX_x = R: OX.e0 is taken as OX_x (K19:1256), so the foot is (0,0,R) above the
    # centre, e1 = (1,0,-1)/sqrt2, e2 = (0,1,0)/sqrt2, r = R sqrt2: the point turns in x-y and
    # rises by R (1 - cos th)
    L.orc_ibm_xrot_f64(P(m), cx + Rr, cy, cz, P(o[0]), P(o[1]), P(o[2]))
    assert np.allclose([o[0][0], o[1][0], o[2][0]], [cx + Rr * np.cos(th), cy + Rr * np.sin(th), cz + Rr * (1 - np.cos(th))],
                       atol=1e-12)
    L.orc_ibm_urot_f64(P(m), cx + Rr, cy, cz, P(o[0]), P(o[1]), P(o[2]))
    assert np.allclose([o[0][0], o[1][0], o[2][0]], [0, Rr * th, 0], atol=1e-12)


@pytest.mark.parametrize("Q", [19, 27])
def test_force_adds_rho_G_momentum_and_no_mass(Q):
    rng = np.random.default_rng(2)
    N = 30
    f = rng.uniform(0.01, 0.3, (Q, N))
    f0 = f.copy()
    solid = (rng.random(N) < 0.2).astype(np.uint8)
    rho = rng.uniform(0.9, 1.1, N)
    G = rng.standard_normal((3, N)) * 1e-3
    po.orc().orc_ibm_force_f64(Q, N, P(f), P(solid), P(rho), P(G[0]), P(G[1]), P(G[2]))
    if Q == 27:
        cx, cy, cz = np.array(po.ORC27_C)
    else:
        cx, cy, cz = np.array(po.ORC19_C)
    df = f - f0
    fl = solid == 0
    assert np.allclose(df.sum(0), 0, atol=1e-15)
    for c, G_c in zip((cx, cy, cz), G):
        assert np.allclose((df * c[:, None]).sum(0)[fl], (rho * G_c)[fl], atol=1e-15)
    assert np.all(df[:, ~fl] == 0)
