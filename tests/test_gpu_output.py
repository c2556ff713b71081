"""The output step on the GPU: lbffs_output_macros against the oracle's restated host
cal_rhoUVW (bit-identical), a run continued from the restart file written by the output
step (bit-identical to the uninterrupted run where the step reads no stored moments), and
the diagnostics of an output step with force coefficients."""
import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases
from lbffs_amd import output as out

import pyoracle as po
from parity_util import run_hip

pytestmark = pytest.mark.gpu


def _lattice(c, storage=lb.STORE_F32):
    coll = lb.COLL_CUMULANT if c.q == 27 else lb.COLL_RR
    lat = lb.Lattice(c.nx, c.ny, c.nz, q=c.q, storage=storage, sgs=c.sgs, collision=coll, omega0=c.omega0)
    c.upload_to(lat)
    return lat


@pytest.mark.parametrize("name,q,storage", [("cylinder", 19, lb.STORE_F32), ("channel", 19, lb.STORE_F32),
                                            ("cavity", 27, lb.STORE_F32), ("cylinder", 19, lb.STORE_F16DEV)])
def test_output_macros_bitwise(name, q, storage):
    if name == "cylinder":
        c = cases.cylinder_channel(nx=64, ny=32, nz=4, D=8, cx=20, cy=16, q=q)
    elif name == "channel":
        c = cases.channel(nx=8, ny=20, nz=4, tau=0.8, force=1e-4, q=q)
    else:
        c = cases.cavity(16, q=q)
    with _lattice(c, storage) as lat:
        lat.step(25, **c.params)
        f = lat.download("f")
        dpdx = c.params.get("dpdx", 0.0)
        m = lat.output_macros(dpdx)
    o = po.cal_rhouvw(q, f, dpdx)
    for k in ("rho", "u", "v", "w"):
        assert np.array_equal(m[k].view(np.uint32), o[k].view(np.uint32)), k


def test_restart_continues_bitwise(tmp_path):
    # periodic channel, body force, no sub-grid model: the step reads no stored moments,
    # so a run restarted from data.dat (populations only, restart.hpp) is the same run
    c = cases.channel(nx=16, ny=24, nz=6, tau=0.7, force=2e-5, q=19)
    ref = run_hip(c, 40)
    with _lattice(c) as lat:
        lat.step(20, **c.params)
        st = out.OutputStep(lat, c.arrays, L=1.0, c=1.0, uMax=1.0, deltaT=1.0, out_dir=str(tmp_path),
                            out_interval=20, next_out=20, mode=out.VTK_BINARY, dpdx=c.params["dpdx"],
                            log=lambda s: None)
        assert st.step_begin(19) is None
        assert st.step_begin(20) is not None
    nt, nx, ny, nz, f = out.read_restart(str(tmp_path / "data.dat"), 19)
    assert (nt, nx, ny, nz) == (20, c.nx, c.ny, c.nz)
    with _lattice(c) as lat:
        lat.upload("f", f)
        lat.step(20, **c.params)
        g = lat.download("f")
    assert np.array_equal(g, ref["f"])
    assert (tmp_path / "res00000020.vtk").stat().st_size > 18 * 4 * c.N


def test_output_step_force_coefficients(tmp_path):
    c = cases.cylinder_channel(nx=96, ny=48, nz=4, D=12, cx=30, cy=24, Re=80.0, q=19)
    with _lattice(c) as lat:
        lat.step(200, **c.params, store_every=1)
        lines = []
        st = out.OutputStep(lat, c.arrays, L=0.01, c=2.0, uMax=0.1, deltaT=0.005, out_dir=str(tmp_path),
                            out_interval=100, next_out=200, mode=out.VTK_ASCII, force_coeffs=True, Dref=0.12,
                            log=lines.append)
        s = st.step_begin(200)
    print(lines, s["fp64"], flush=True)
    assert any(l.startswith("Cx: ") for l in lines)
    # the FP32 sums of the reference arithmetic and the FP64 device sums agree to FP32
    # rounding of sums whose terms are of the size of the drag (the lift is a small
    # difference of such terms, so its relative error is larger)
    scale = abs(s["fp64"]["Cx"])
    for k in ("Cx", "Cy"):
        assert abs(s[k] - s["fp64"][k]) <= 1e-4 * scale, k
    assert abs(s["rho_av"] - s["fp64"]["rho_av"]) <= 1e-6
    assert s["Cx"] > 0  # drag along the flow
    text = (tmp_path / "res00000200.vtk").read_text()
    assert text.count("SCALARS ") == 18 and "DIMENSIONS 96 48 4" in text
