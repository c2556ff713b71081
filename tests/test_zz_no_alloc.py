"""lbffs_reduce_noalloc (no counterpart in the reference): the FP64 sums of lbffs_reduce
with the reduction's partials in the library's fixed device array instead of context
memory.  Called through the C ABI; its results must equal lbffs_reduce's bit for bit, on
two calls in a row, the second on a smaller lattice, and it must refuse to rebuild flags
(which allocates)."""
import ctypes as C

import numpy as np
import pytest

import lbffs_amd as lb
from lbffs_amd import cases

pytestmark = pytest.mark.gpu


def _both(lat):
    L = lb.lib()
    a = (C.c_double * lb.REDUCE_N)()
    b = (C.c_double * lb.REDUCE_N)()
    assert L.lbffs_reduce(lat._h, a) == 0
    rc = L.lbffs_reduce_noalloc(lat._h, b)
    assert rc == 0, L.lbffs_last_error()
    return np.array(a[:]), np.array(b[:])


@pytest.mark.parametrize("storage", [lb.STORE_F32, lb.STORE_F16DEV, lb.STORE_F64])
def test_reduce_noalloc_equals_reduce_on_two_calls(storage):
    big = cases.cylinder_channel(nx=96, ny=48, nz=8, D=12, cx=30, cy=24, Re=60.0, spz_cells=8)
    small = cases.cavity(12)
    with lb.Lattice(big.nx, big.ny, big.nz, storage=storage, sgs=big.sgs, omega0=big.omega0) as A, \
            lb.Lattice(small.nx, small.ny, small.nz, storage=storage, sgs=small.sgs, omega0=small.omega0) as B:
        big.upload_to(A)
        small.upload_to(B)
        A.step(20, **big.params)
        B.step(7, **small.params)
        ra, na = _both(A)
        assert np.array_equal(ra, na), (ra, na)
        assert na[7] == int((big.arrays["solid"] == 0).sum()) and na[0] != 0  # fluid count, drag
        rb, nb = _both(B)  # the smaller lattice second: the shared partials hold the first call's rows
        assert np.array_equal(rb, nb), (rb, nb)
        assert nb[7] == small.N
        # the first lattice again, after the second overwrote the scratch
        ra2, na2 = _both(A)
        assert np.array_equal(ra2, na2) and np.array_equal(na2, na)


def test_reduce_noalloc_refuses_to_rebuild_flags():
    c = cases.cavity(10)
    L = lb.lib()
    out = (C.c_double * lb.REDUCE_N)()
    with lb.Lattice(c.nx, c.ny, c.nz, sgs=c.sgs, omega0=c.omega0) as lat:
        c.upload_to(lat)  # the flags are built by the next call that may allocate
        assert L.lbffs_reduce_noalloc(lat._h, out) == -5  # LBFFS_E_STATE
        assert b"rebuilt" in L.lbffs_last_error()
        lat.reduce()
        assert L.lbffs_reduce_noalloc(lat._h, out) == 0
    assert L.lbffs_reduce_noalloc(None, out) == -1
