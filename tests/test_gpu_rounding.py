"""Stochastic rounding of the 16-bit storage forms (lbffs_desc.rounding).

A wake behind a cylinder at Re 100 grows from the seeded inlet noise into periodic vortex
shedding.  Stored with round-to-nearest FP16 deviations the growing asymmetry stays below
half a unit in the last place and is rounded away every step, so the wake never sheds;
with stochastic rounding the FP16 run sheds like the FP32 one, at the same Strouhal number.
"""
import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases

pytestmark = pytest.mark.gpu

STEPS, EVERY = 50000, 50


def _history(c, storage, rounding):
    with lb.Lattice(c.nx, c.ny, c.nz, storage=storage, sgs=c.sgs, omega0=c.omega0, rounding=rounding) as lat:
        c.upload_to(lat)
        norm = 0.5 * c.meta["u_in"] ** 2 * c.meta["D"] * c.nz
        cd, cl = [], []
        for _ in range(STEPS // EVERY):
            lat.step(EVERY, **c.params)
            s = lat.reduce()
            cd.append(s[0] / norm)
            cl.append(s[1] / norm)
    return np.array(cd), np.array(cl)


def _strouhal(cl, D, u):
    c = cl[len(cl) // 2:]
    c = c - c.mean()
    spec = np.abs(np.fft.rfft(c * np.hanning(len(c))))
    k = int(np.argmax(spec[1:])) + 1
    a, b, g = spec[k - 1], spec[k], spec[k + 1]
    k = k + 0.5 * (a - g) / (a - 2 * b + g)
    return k / (len(c) * EVERY) * D / u


def test_fp16_stochastic_rounding_sheds_like_fp32():
    # the channel-past-a-cylinder workload: 400 x 160 x 16, D = 40, Re = 100, interpolated
    # bounce-back with analytic q; shedding is established within the 50k steps
    c = cases.cylinder_channel(ibb=True)
    D, u = c.meta["D"], c.meta["u_in"]
    res = {}
    for name, st, rnd in (("f32", lb.STORE_F32, 0), ("f16_sr", lb.STORE_F16DEV, lb.ROUND_STOCHASTIC),
                          ("f16_rn", lb.STORE_F16DEV, lb.ROUND_NEAREST)):
        cd, cl = _history(c, st, rnd)
        late = cl[-len(cl) // 4:]
        res[name] = {"Cd": float(cd[len(cd) // 2:].mean()), "Cl_rms_late": float(np.std(late)),
                     "St": float(_strouhal(cl, D, u))}
        print(name, res[name], flush=True)
    assert res["f32"]["Cl_rms_late"] > 0.05
    assert res["f16_sr"]["Cl_rms_late"] > 0.5 * res["f32"]["Cl_rms_late"]
    assert abs(res["f16_sr"]["St"] - res["f32"]["St"]) <= 0.01 * res["f32"]["St"]
    assert abs(res["f16_sr"]["Cd"] - res["f32"]["Cd"]) <= 0.01 * res["f32"]["Cd"]
    # round-to-nearest holds the wake symmetric: no shedding
    assert res["f16_rn"]["Cl_rms_late"] < 0.1 * res["f32"]["Cl_rms_late"]
