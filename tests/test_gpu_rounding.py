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


def _bracket16(d, bf16):
    """The two 16-bit values next to each float32 d (equal where d is representable)."""
    d = np.asarray(d, np.float32)
    if not bf16:
        rn = d.astype(np.float16)
        up = rn.astype(np.float32) < d
        dn = rn.astype(np.float32) > d
        lo = np.where(up, rn, np.where(dn, np.nextafter(rn, np.float16(-np.inf)), rn)).astype(np.float32)
        hi = np.where(up, np.nextafter(rn, np.float16(np.inf)), rn).astype(np.float32)
        return lo, hi
    b = d.view(np.uint32)
    t = (b & np.uint32(0xFFFF0000)).view(np.float32)          # toward zero
    away = ((b & np.uint32(0xFFFF0000)) + np.uint32(0x10000)).view(np.float32)
    exact = t == d
    lo = np.where(exact | (d > 0), t, away)
    hi = np.where(exact | (d < 0), t, away)
    return lo.astype(np.float32), hi.astype(np.float32)


@pytest.mark.parametrize("q,storage", [(19, lb.STORE_F16DEV), (27, lb.STORE_F16DEV), (19, lb.STORE_BF16DEV)])
def test_stochastic_rounding_is_unbiased(q, storage):
    # a uniform stream on a periodic lattice: every interior cell computes the same new
    # populations x_q, so the stored 16-bit deviations of one step are independent
    # roundings of one value.  Each must be one of the two neighbours of x_q's deviation,
    # and the share rounded up must equal x_q's position between them (the rounding is
    # unbiased); with nearest rounding every cell stores the nearest value.
    n = 32
    c = cases.periodic_box(n, n, n, tau=0.9, ini=(0.0123, 0.0071, -0.0049), q=q)
    coll = lb.COLL_CUMULANT if q == 27 else lb.COLL_RR
    bf16 = storage == lb.STORE_BF16DEV
    inner = np.zeros((n, n, n), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1)

    def lattice(st, rnd=0):
        return lb.Lattice(n, n, n, q=q, storage=st, sgs=c.sgs, collision=coll, omega0=c.omega0, rounding=rnd)

    with lattice(storage) as lat:
        c.upload_to(lat)
        f0 = lat.download("f")                        # the 16-bit state, decoded
    with lattice(lb.STORE_F32) as lat:
        c.upload_to(lat)
        lat.upload("f", f0)
        lat.step(1, **c.params)
        x = lat.download("f").reshape(q, -1)[:, inner]
        u = lat.download("u")
    assert abs(float(u[inner].mean())) > 1e-3
    assert (x == x[:, :1]).all(), "the FP32 step is not uniform"
    x = x[:, 0]
    w = np.array([float(lb_w) for lb_w in _weights(q)], np.float32)
    scale = np.array([_scale(q, d) for d in range(q)], np.float32)
    d = (x - w) * scale
    lo, hi = _bracket16(d, bf16)
    dec_lo, dec_hi = w + lo / scale, w + hi / scale
    for rnd in (lb.ROUND_STOCHASTIC, lb.ROUND_NEAREST):
        with lattice(storage, rnd) as lat:
            c.upload_to(lat)
            lat.step(1, **c.params)
            y = lat.download("f").reshape(q, -1)[:, inner]
        m = y.shape[1]
        # The 16-bit forms take their moments from the deviations, the FP32 storage form
        # from the populations: the two steps round differently at the level of an FP32
        # ulp of the weight, so the value the 16-bit form rounds is within a few such ulps
        # (scaled to the stored deviation) of d.  Where d lies that close to a 16-bit value
        # only the number of stored values is checked.
        for k in range(q):
            is_lo, is_hi = y[k] == dec_lo[k], y[k] == dec_hi[k]
            tol = 8 * float(np.spacing(w[k])) * float(scale[k])
            near_grid = min(abs(d[k] - lo[k]), abs(hi[k] - d[k])) <= tol
            if not near_grid:
                assert (is_lo | is_hi).all(), (rnd, k)
            else:
                vals = np.unique(y[k])
                assert len(vals) <= 2, (rnd, k, vals)
                continue
            if rnd == lb.ROUND_NEAREST:
                near = dec_hi[k] if abs(hi[k] - d[k]) < abs(d[k] - lo[k]) else dec_lo[k]
                mid_gap = abs(abs(hi[k] - d[k]) - abs(d[k] - lo[k]))
                if mid_gap > 2 * tol:
                    assert (y[k] == near).all(), k
                continue
            p = 0.0 if hi[k] == lo[k] else float((d[k] - lo[k]) / (hi[k] - lo[k]))
            share = float(is_hi.mean()) if hi[k] != lo[k] else 0.0
            sigma = np.sqrt(max(p * (1 - p), 1e-12) / m)
            dp = 0.0 if hi[k] == lo[k] else tol / float(hi[k] - lo[k])
            assert abs(share - p) < 6 * sigma + dp + 1e-9, (k, share, p)
        if rnd == lb.ROUND_STOCHASTIC:
            print(f"q{q} storage {storage}: {m} cells, rounded-up shares vs positions ok", flush=True)


def _weights(q):
    if q == 19:
        return [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12
    return [8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8


def _scale(q, d):
    if d == 0:
        return 4.0
    if d <= 6:
        return 16.0
    if q == 19 or d <= 18:
        return 32.0 if q == 19 else 64.0
    return 256.0


def test_stochastic_rounding_ignores_the_launch_shape():
    # the random bits come from the cell's own state and index, so the result does not
    # depend on the work-group size or order
    c = cases.cylinder_channel(nx=96, ny=48, nz=4, D=12, cx=30, cy=24, Re=150.0, q=19)
    out = []
    for block, tune in ((256, 0), (64, 2), (128, 0)):
        with lb.Lattice(c.nx, c.ny, c.nz, storage=lb.STORE_F16DEV, sgs=c.sgs, omega0=c.omega0) as lat:
            c.upload_to(lat)
            lat.set_launch(block, tune)
            lat.step(40, **c.params)
            out.append(lat.download("f"))
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
