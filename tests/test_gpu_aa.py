"""In-place (AA) streaming against the two-lattice step (lbffs_desc.streaming).

Both run the same per-cell arithmetic; in-place streaming only moves where each new
population is stored, and restores the two-lattice layout before anything reads the array
by cell.  Bar: bit-identical populations, moments, sub-grid viscosity, wall forces and FP64
sums after odd and even step counts, steps split across calls, in every storage form.
"""
import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases

pytestmark = pytest.mark.gpu

E_INVALID, E_UNSUPPORTED = -1, -4  # LBFFS_E_* of include/lbffs.h
KEYS = ("f", "rho", "u", "v", "w", "tauSGS", "Fwx", "Fwy", "Fwz")


def _channel_with_cylinder(nx=48, ny=24, nz=6, sgs=lb.SGS_WALE):
    # periodic x and z, bounce-back y walls, body force, a solid cylinder (simple bounce-back)
    c = cases.channel(nx=nx, ny=ny, nz=nz, tau=0.6, force=2e-5, sgs=sgs)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = (((i - 16) ** 2 + (j - 11.5) ** 2) < 4.2 ** 2).astype(np.uint8).reshape(-1)
    c.arrays["solid"] = solid
    c.arrays["neiSolid"] = cases.neisolid(19, nx, ny, nz, solid)
    return c


def _run(c, chunks, storage, streaming, autotune_at=None, rounding=lb.ROUND_STOCHASTIC):
    with lb.Lattice(c.nx, c.ny, c.nz, storage=storage, sgs=c.sgs, omega0=c.omega0, rounding=rounding,
                    streaming=streaming) as lat:
        c.upload_to(lat)
        for n, steps in enumerate(chunks):
            if autotune_at == n:
                lat.autotune(3, **c.params)
            lat.step(steps, **c.params)
        out = {k: lat.download(k) for k in KEYS}
        out["reduce"] = np.array(lat.reduce())
        out["macros"] = lat.output_macros(c.params.get("dpdx", 0.0))
    return out


def _same(a, b):
    for k in KEYS + ("reduce",):
        assert np.array_equal(a[k], b[k]), k
    for k in ("rho", "u", "v", "w"):
        assert np.array_equal(a["macros"][k], b["macros"][k]), k


@pytest.mark.parametrize("storage", [lb.STORE_F32, lb.STORE_F16DEV, lb.STORE_BF16DEV, lb.STORE_F64])
@pytest.mark.parametrize("chunks", [[7], [6], [3, 4, 1]])
def test_in_place_equals_two_lattice(storage, chunks):
    c = _channel_with_cylinder()
    two = _run(c, chunks, storage, lb.STREAM_TWO_LATTICE)
    aa = _run(c, chunks, storage, lb.STREAM_AA)
    _same(two, aa)
    assert np.abs(two["Fwx"]).sum() > 0


@pytest.mark.parametrize("sgs", [lb.SGS_SMAGORINSKY, lb.SGS_NONE])
def test_in_place_sgs_models_and_2d(sgs):
    c = _channel_with_cylinder(nz=1, sgs=sgs)
    _same(_run(c, [5], lb.STORE_F32, lb.STREAM_TWO_LATTICE), _run(c, [5], lb.STORE_F32, lb.STREAM_AA))
    b = cases.periodic_box(12, 10, 8, tau=0.7, ini=(0.03, -0.02, 0.01), sgs=sgs)
    _same(_run(b, [9], lb.STORE_F16DEV, lb.STREAM_TWO_LATTICE), _run(b, [9], lb.STORE_F16DEV, lb.STREAM_AA))


def test_in_place_through_the_autotuner():
    # the autotuner runs trial steps on the lattice and restores it
    c = _channel_with_cylinder()
    two = _run(c, [3, 4], lb.STORE_F32, lb.STREAM_TWO_LATTICE)
    aa = _run(c, [3, 4], lb.STORE_F32, lb.STREAM_AA, autotune_at=1)
    _same(two, aa)


def _step_rc(c, **kw):
    lat = lb.Lattice(c.nx, c.ny, c.nz, sgs=c.sgs, omega0=c.omega0, streaming=lb.STREAM_AA)
    try:
        c.upload_to(lat)
        p = lb._params(**{**c.params, **kw})
        import ctypes
        rc = lb.lib().lbffs_step(lat._h, 1, ctypes.byref(p))
        return rc, lb.lib().lbffs_last_error()
    finally:
        lat.close()


def test_in_place_refusals():
    # faces other than Cyclic / BounceBack
    rc, msg = _step_rc(cases.cavity(12))
    assert rc == E_INVALID and b"in-place streaming" in msg
    # different codes on the two sides of an axis
    codes = {f: ("Cyclic", "Cyclic", "Cyclic") for f in cases.FACES}
    codes["nyMin"] = ("Cyclic", "BounceBack", "Cyclic")
    c = cases.make_case("half", 10, 8, 6, 1.0, 1.0, 0.05, 1e-3, codes, sgs=lb.SGS_NONE)
    rc, msg = _step_rc(c)
    assert rc == E_INVALID and b"both sides" in msg
    # the interpolated bounce-back
    c = _channel_with_cylinder()
    c.arrays["qlink"] = np.full(19 * c.N, 0.5, np.float32)
    rc, msg = _step_rc(c, ibb=1)
    assert rc == E_UNSUPPORTED
    # D3Q27 and batches
    with pytest.raises(RuntimeError):
        lb.Lattice(8, 8, 8, q=27, collision=lb.COLL_CUMULANT, streaming=lb.STREAM_AA)
