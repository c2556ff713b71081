"""Batches: many lattices of uneven size advanced by one launch per step give, lattice by
lattice, the same populations and moments as running each lattice on its own -- for any
grid (one work-group per block, or a few persistent work-groups) and with the steps
replayed from captured hipGraphs."""
import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases

from parity_util import run_hip

pytestmark = pytest.mark.gpu

KEYS = ("f", "rho", "u", "v", "w", "tauSGS", "Fwx")


def _run_batch(cs, nsteps, storage=lb.STORE_F32, grid=0, graph_chunk=0, chunks=1):
    q = cs[0].q
    coll = lb.COLL_CUMULANT if q == 27 else lb.COLL_RR
    with lb.Batch([(c.nx, c.ny, c.nz, c.omega0) for c in cs], q=q, storage=storage, sgs=cs[0].sgs,
                  collision=coll) as b:
        for c, lat in zip(cs, b.lattice):
            c.upload_to(lat)
        b.set_launch(grid, graph_chunk)
        left = nsteps
        for k in range(chunks):
            n = left // (chunks - k)
            b.step(n, [c.params for c in cs])
            left -= n
        return [{k: lat.download(k) for k in KEYS} for lat in b.lattice]


@pytest.mark.parametrize("q,storage", [(19, lb.STORE_F32), (19, lb.STORE_F16DEV), (27, lb.STORE_F32)])
def test_batch_equals_single_lattices(q, storage):
    cs = cases.ensemble(6, edges=(12, 36), q=q)
    ref = [run_hip(c, 24, storage=storage) for c in cs]
    # grids: one work-group per block, persistent grids rounded up to a multiple of 8 (7, 13)
    # or already one (40), with and without captured graphs
    for grid, chunk, chunks in ((0, 0, 1), (7, 0, 1), (0, 4, 2), (40, 6, 1), (13, 3, 2)):
        out = _run_batch(cs, 24, storage=storage, grid=grid, graph_chunk=chunk, chunks=chunks)
        for d, (o, r) in enumerate(zip(out, ref)):
            for k in KEYS:
                assert np.array_equal(o[k], r[k]), (grid, chunk, d, k)


@pytest.mark.parametrize("storage", [lb.STORE_F32, lb.STORE_F16DEV])
def test_batch_face_table_follows_the_lattices_wave_summary_use(storage):
    # the face table leaves out the face cells that XMIX and WRAP waves run themselves; a
    # lattice whose launch option turns its wave summary off (16 = every wave reads its
    # flags first) needs them back in the table, and one that turns it on again needs them out
    cs = cases.ensemble(4, edges=(24, 40))
    ref = [run_hip(c, 20, storage=storage) for c in cs]
    with lb.Batch([(c.nx, c.ny, c.nz, c.omega0) for c in cs], q=19, storage=storage, sgs=cs[0].sgs) as b:
        for c, lat in zip(cs, b.lattice):
            c.upload_to(lat)
        params = [c.params for c in cs]
        b.lattice[1].set_launch(256, 16)
        b.step(7, params)
        b.lattice[1].set_launch(256, 0)
        b.lattice[2].set_launch(256, 16)
        b.step(6, params)
        b.lattice[2].set_launch(256, 0)
        b.step(7, params)
        for d, (lat, r) in enumerate(zip(b.lattice, ref)):
            for k in KEYS:
                assert np.array_equal(lat.download(k), r[k]), (d, k)


def test_batch_refuses_a_lattice_stepped_on_its_own():
    cs = cases.ensemble(2, edges=(8, 10))
    with lb.Batch([(c.nx, c.ny, c.nz, c.omega0) for c in cs], q=19, sgs=cs[0].sgs) as b:
        for c, lat in zip(cs, b.lattice):
            c.upload_to(lat)
        b.lattice[0].step(1, **cs[0].params)
        with pytest.raises(lb.LbffsError, match="parity"):
            b.step(1, [c.params for c in cs])


@pytest.mark.parametrize("storage", [lb.STORE_F32, lb.STORE_F16DEV])
def test_batch_autotune_leaves_every_lattice_unchanged(storage):
    # 12857 blocks of 256 cells: every persistent grid (2048, 4096, 8192) is tried; the
    # 16-bit form also checks that the step index feeding the stochastic rounding is put back
    cs = cases.ensemble(6, edges=(48, 96))
    ref = [run_hip(c, 20, storage=storage) for c in cs]
    with lb.Batch([(c.nx, c.ny, c.nz, c.omega0) for c in cs], q=19, storage=storage, sgs=cs[0].sgs) as b:
        for c, lat in zip(cs, b.lattice):
            c.upload_to(lat)
        params = [c.params for c in cs]
        b.step(7, params)
        grid, chunk, ms = b.autotune(params, 3)
        assert grid in (0, 2048, 4096, 8192) and chunk in (0, 16) and ms > 0
        b.step(13, params)
        for d, (lat, r) in enumerate(zip(b.lattice, ref)):
            for k in KEYS:
                assert np.array_equal(lat.download(k), r[k]), (grid, chunk, d, k)
        with pytest.raises(lb.LbffsError, match="trial_steps"):
            b.autotune(params, 0)
