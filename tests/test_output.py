"""The output step on the CPU: the VTK and restart file formats of write.hpp /
restart.hpp, the diagnostics' arithmetic, the running means, and the oracle's restated
host cal_rhoUVW pinned by closed forms."""
import struct

import numpy as np
import pytest

from lbffs_amd import cases
from lbffs_amd import output as out

import pyoracle as po


def _fields(N, **over):
    a = {n: np.zeros(N, np.float32) for n in ("u", "v", "w", "rho", "tauSGS", "sdf", "GxIBM", "GyIBM", "GzIBM",
                                              "omega", "uMean", "vMean", "wMean")}
    for n in ("boundary1", "boundary2", "boundary3"):
        a[n] = np.zeros(N, np.int32)
    a["solid"] = np.zeros(N, np.uint8)
    a["neiSolid"] = np.zeros(N, np.uint8)
    for k, v in over.items():
        a[k] = np.asarray(v, a[k].dtype)
    return a


def test_vtk_ascii_known_text(tmp_path):
    # 2 x 1 x 1 cells; values chosen so that every FP32 product below is exact
    f = _fields(2, u=[0.5, -0.25], v=[1.5e-5, 0.0], w=[-0.0, 1234567.0], rho=[1.0, 0.5],
                tauSGS=[0.25, 0.0], boundary1=[5, 0], boundary2=[0, 1], boundary3=[7, 3], sdf=[2.0, -4.0],
                solid=[1, 0], neiSolid=[0, 1], GxIBM=[0.125, 0], GyIBM=[0, -2], GzIBM=[3, 0], omega=[1.75, 0.5],
                uMean=[8.0, 4.0], vMean=[-8.0, 0.0], wMean=[0.0, 2.0])
    p = tmp_path / "r.vtk"
    out.write_vtk(str(p), 2, 1, 1, 0.5, 2.0, 4.0, f, out.VTK_ASCII)
    # c = 2, L = 0.5, timeMean = 4: u*c, 3*(3*tau*L*c), sdf*L, uMean*c/timeMean
    expect = ("# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET STRUCTURED_POINTS\n"
              "DIMENSIONS 2 1 1\nORIGIN 0 0 0\nSPACING 0.5 0.5 0.5\n\nPOINT_DATA 2\n"
              "SCALARS u float\nLOOKUP_TABLE default\n1\n-0.5\n\n"
              "SCALARS v float\nLOOKUP_TABLE default\n3e-05\n0\n\n"
              "SCALARS w float\nLOOKUP_TABLE default\n-0\n2.46913e+06\n\n"
              "SCALARS rho float\nLOOKUP_TABLE default\n1\n0.5\n\n"
              "SCALARS nuSGS float\nLOOKUP_TABLE default\n2.25\n0\n\n"
              "SCALARS boundary1 float\nLOOKUP_TABLE default\n5\n0\n\n"
              "SCALARS boundary2 float\nLOOKUP_TABLE default\n0\n1\n\n"
              "SCALARS boundary3 float\nLOOKUP_TABLE default\n7\n3\n\n"
              "SCALARS sdf float\nLOOKUP_TABLE default\n1\n-2\n\n"
              "SCALARS solid float\nLOOKUP_TABLE default\n1\n0\n\n"
              "SCALARS neiSolid float\nLOOKUP_TABLE default\n0\n1\n\n"
              "SCALARS GxIBM float\nLOOKUP_TABLE default\n0.125\n0\n\n"
              "SCALARS GyIBM float\nLOOKUP_TABLE default\n0\n-2\n\n"
              "SCALARS GzIBM float\nLOOKUP_TABLE default\n3\n0\n\n"
              "SCALARS omega float\nLOOKUP_TABLE default\n1.75\n0.5\n\n"
              "SCALARS uMean float\nLOOKUP_TABLE default\n4\n2\n\n"
              "SCALARS vMean float\nLOOKUP_TABLE default\n-4\n0\n\n"
              "SCALARS wMean float\nLOOKUP_TABLE default\n0\n1\n\n")
    assert p.read_text() == expect


def _binary_values(data, n_fields, N):
    """Split a binary VTK file into its header and the N*4-byte blocks of every field."""
    head_end = data.index(b"POINT_DATA")
    pos = data.index(b"\n", head_end) + 1
    blocks = []
    for _ in range(n_fields):
        pos = data.index(b"LOOKUP_TABLE default\n", pos) + len(b"LOOKUP_TABLE default\n")
        blocks.append(data[pos:pos + 4 * N])
        pos += 4 * N + 1
    assert pos == len(data)
    return data[:head_end], blocks


def test_vtk_binary_big_endian_and_reference_order(tmp_path):
    N = 6
    rng = np.random.default_rng(1)
    f = _fields(N, u=rng.standard_normal(N), rho=1 + 0.01 * rng.standard_normal(N), omega=rng.random(N),
                boundary1=[0, 1, 3, 4, 5, 7], solid=[0, 1, 0, 1, 1, 0])
    for mode in (out.VTK_BINARY, out.VTK_BINARY_REF):
        p = tmp_path / f"b{mode}.vtk"
        out.write_vtk(str(p), 3, 2, 1, 0.1, 3.0, 1.0, f, mode)
        head, blocks = _binary_values(p.read_bytes(), 18, N)
        assert b"BINARY\n" in head
        names = out.VTK_ORDER
        want = {"u": f["u"] * np.float32(3.0), "rho": f["rho"], "omega": f["omega"],
                "boundary1": f["boundary1"].astype(np.float32), "solid": f["solid"].astype(np.float32)}
        for name, w in want.items():
            le = np.asarray(w, np.float32).tobytes()
            b = blocks[names.index(name)]
            if mode == out.VTK_BINARY:
                assert np.array_equal(np.frombuffer(b, ">f4"), np.asarray(w, np.float32)), name
            else:
                # base.hpp:35-43: bytes 3, 2, 1, 3 of the little-endian float
                ref = bytes(x for i in range(N) for x in (le[4 * i + 3], le[4 * i + 2], le[4 * i + 1], le[4 * i + 3]))
                assert b == ref, name


def test_restart_layout_and_round_trip(tmp_path):
    q, nx, ny, nz = 19, 4, 3, 2
    f = np.random.default_rng(2).random(q * nx * ny * nz).astype(np.float32)
    p = tmp_path / "data.dat"
    out.write_restart(str(p), 1234, q, nx, ny, nz, f)
    raw = p.read_bytes()
    assert len(raw) == 16 + 4 * f.size
    assert struct.unpack("<4i", raw[:16]) == (1234, nx, ny, nz)
    assert raw[16:] == f.tobytes()
    nt, a, b, c, g = out.read_restart(str(p), q)
    assert (nt, a, b, c) == (1234, nx, ny, nz) and np.array_equal(g, f)
    # a D3Q27 reader expects 27 arrays: the file is too short
    with pytest.raises(Exception):
        out.read_restart(str(p), 27)
    with pytest.raises(Exception):
        out.read_restart(str(tmp_path / "missing.dat"), q)


def test_stats_reference_arithmetic():
    nx, ny, nz = 3, 2, 1
    rho = np.array([1.5, 0.75, 2.0, 1.0, 1.25, 0.5], np.float32)
    u = np.array([0.25, 9, 9, 0.75, 9, 9], np.float32)  # the i = 0 column: cells 0 and 3
    solid = np.array([0, 0, 1, 0, 0, 1], np.uint8)
    forces = {"Fwx": np.array([1, 2, 100, 3, 4, 100], np.float32), "Fwy": np.full(6, 0.5, np.float32),
              "Fwz": np.zeros(6, np.float32), "GxIBM": np.array([1, 0, 0, 0, 0, 0], np.float32),
              "GyIBM": np.zeros(6, np.float32), "GzIBM": np.zeros(6, np.float32)}
    s = out.stats(nx, ny, nz, 2.0, 0.5, 1.0, 0.125, rho, u, solid, forces)
    # fluid cells 0, 1, 3, 4: mean(rho - 1) = (0.5 - 0.25 + 0 + 0.25) / 4 = 0.125
    assert s["eleFluid"] == 4 and s["rho_av"] == np.float32(1.125)
    assert s["rhoMax"] == 2.0 and s["rhoMin"] == 0.5        # over every cell, solid included
    assert s["uAve"] == np.float32((0.25 + 0.75) * 2.0 / 2)  # u*c over the i = 0 plane / (ny*nz)
    assert s["sumFw"] == [10.0, 2.0, 0.0] and s["sumG"] == [1.0, 0.0, 0.0]
    den = 0.5 * 1.125 * 1.0 * 1.0 * 0.125 * 0.5
    assert s["Cx"] == pytest.approx(10.0 * 4 * 0.25 / den, rel=1e-7)
    assert s["CxIBM"] == pytest.approx(-1.0 * 4 * 0.25 / den, rel=1e-7)
    c64 = out.coeffs_fp64([10.0, 2.0, 0.0, 1.0, 0.0, 0.0, 0.5, 4.0], 2.0, 0.5, 1.0, 0.125)
    assert c64["rho_av"] == 1.125 and c64["Cx"] == pytest.approx(10.0 / den, rel=1e-15)
    assert c64["CxIBM"] == pytest.approx(-1.0 / den, rel=1e-15)


def test_running_means():
    from lbffs_amd import lib, _ptr
    u = np.array([0.5, -1.0], np.float32)
    v = np.array([2.0, 0.0], np.float32)
    w = np.array([0.0, 0.25], np.float32)
    m = [np.zeros(2, np.float32) for _ in range(3)]
    t = np.zeros(1, np.float32)
    for _ in range(3):
        assert lib().lbffs_out_accumulate_mean(2, 100, _ptr(u), _ptr(v), _ptr(w), *[_ptr(x) for x in m], _ptr(t)) == 0
    assert t[0] == 300.0
    assert m[0].tolist() == [150.0, -300.0] and m[2].tolist() == [0.0, 75.0]


@pytest.mark.parametrize("q", [19, 27])
def test_oracle_cal_rhouvw_recovers_the_equilibrium(q):
    # f = feq(rho0, u0) of initialization.hpp: the moments give back rho0 and u0
    nx, ny, nz = 3, 2, 2
    f = cases.init_feq(q, nx, ny, nz, 1.03, 0.04, -0.02, 0.01)
    m = po.cal_rhouvw(q, f)
    assert np.allclose(m["rho"], 1.03, atol=2e-7)
    assert np.allclose(m["u"], 0.04 / 1.03, atol=1e-7) and np.allclose(m["v"], -0.02 / 1.03, atol=1e-7)
    assert np.allclose(m["w"], 0.01 / 1.03, atol=1e-7)
    # the force half-step: u += 0.5 dpdx / rho
    m2 = po.cal_rhouvw(q, f, dpdx=1e-3)
    assert np.allclose(m2["u"] - m["u"], 0.5e-3 / 1.03, atol=1e-9)


def test_oracle_cal_rhouvw_direction_sums():
    # one cell, f_q = q + 1: rho = sum, momentum = sum of (q + 1) c_q over the D3Q19 set
    f = np.arange(1, 20, dtype=np.float32)
    m = po.cal_rhouvw(19, f)
    cx, cy, cz = (np.array(c) for c in po.ORC19_C)
    assert m["rho"][0] == 190.0
    assert m["u"][0] == np.float32(float((f * cx).sum()) / 190.0)
    assert m["v"][0] == np.float32(float((f * cy).sum()) / 190.0)
    assert m["w"][0] == np.float32(float((f * cz).sum()) / 190.0)
