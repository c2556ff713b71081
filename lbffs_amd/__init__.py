"""lbffs_amd -- MI355X lattice-Boltzmann time-step engine for the LBFFS solver.

Python view of the C ABI in ``include/lbffs.h`` (device step), ``include/lbffs_geom.h``
(geometry preprocessing on the device), ``include/lbffs_case.h`` (host case setup) and
``include/lbffs_output.h`` (output step: VTK and restart files, diagnostics).  The library ``lbffs_amd/liblbffs.so`` is
built by ``python -m lbffs_amd.build``; importing this package never builds it and
never falls back to another implementation: without the library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbffs.so")

ABI_VERSION = 1

# storage forms, SGS models, collisions (lbffs.h)
STORE_F32, STORE_F16DEV, STORE_BF16DEV, STORE_F64 = 0, 1, 2, 3
SGS_WALE, SGS_SMAGORINSKY, SGS_NONE = 0, 1, 2
COLL_RR, COLL_CUMULANT = 0, 1
ROUND_STOCHASTIC, ROUND_NEAREST = 0, 1  # 16-bit storage forms
STREAM_TWO_LATTICE, STREAM_AA = 0, 1   # population streaming: two arrays, or in place

# BC codes (input.hpp:221-256)
BC = {"Cyclic": 0, "BounceBack": 1, "Outlet": 3, "Symmetry": 4, "FixedVelocity": 5,
      "FixedDensity": 6, "WallFunction": 7}

FIELDS = ["f", "boundary", "boundary1", "boundary2", "boundary3", "sdf", "solid", "neiSolid",
          "u0", "v0", "w0", "Fwx", "Fwy", "Fwz", "tauSGS", "rho", "u", "v", "w", "omega",
          "GxIBM", "GyIBM", "GzIBM", "qlink"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
INT_FIELDS = {"boundary", "boundary1", "boundary2", "boundary3"}
U8_FIELDS = {"solid", "neiSolid"}
REDUCE_N = 8


class LbffsError(RuntimeError):
    pass


class Desc(C.Structure):
    _fields_ = [("abi_version", C.c_int), ("q", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("storage", C.c_int), ("sgs", C.c_int), ("collision", C.c_int), ("omega0", C.c_float),
                ("block", C.c_int), ("rounding", C.c_int), ("streaming", C.c_int), ("reserved", C.c_int * 6)]


class StepParams(C.Structure):
    _fields_ = [("dpdx", C.c_float), ("LES", C.c_float), ("omegaB", C.c_float), ("spzWidth", C.c_float),
                ("Cs", C.c_float), ("damping", C.c_int), ("store_every", C.c_int), ("clim", C.c_float),
                ("ibb", C.c_int), ("reserved", C.c_int * 6)]


class IbmMotion(C.Structure):
    _fields_ = [("trans", C.c_float * 3), ("rot_center", C.c_float * 3), ("rot_axis", C.c_float * 3),
                ("rot_omega", C.c_float), ("reserved", C.c_int * 6)]


IBM_POSITIONS, IBM_FORCING = 0, 1


class CaseIn(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("Lx", C.c_float), ("uMax", C.c_float),
                ("rho0", C.c_float), ("U0", C.c_float), ("nu", C.c_float), ("dpdx", C.c_float),
                ("uIni", C.c_float), ("vIni", C.c_float), ("wIni", C.c_float),
                ("face_code", (C.c_int * 3) * 6), ("face_vel", (C.c_float * 3) * 6)]


class Units(C.Structure):
    _fields_ = [("L", C.c_float), ("c", C.c_float), ("Re", C.c_float), ("nu", C.c_float), ("dpdx", C.c_float),
                ("uIni", C.c_float), ("vIni", C.c_float), ("wIni", C.c_float), ("omega0", C.c_float),
                ("deltaT", C.c_float), ("face_vel", (C.c_float * 3) * 6)]


_lib = None


def lib():
    """Load liblbffs.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LbffsError(f"{LIB_PATH} not found: run `python -m lbffs_amd.build` first")
        L = C.CDLL(LIB_PATH)
        vp, ip, sz = C.c_void_p, C.c_int, C.c_size_t
        L.lbffs_create.argtypes = [C.POINTER(Desc), C.POINTER(vp)]
        L.lbffs_upload.argtypes = [vp, ip, vp, sz]
        L.lbffs_download.argtypes = [vp, ip, vp, sz]
        L.lbffs_step.argtypes = [vp, ip, C.POINTER(StepParams)]
        L.lbffs_reduce.argtypes = [vp, C.POINTER(C.c_double)]
        L.lbffs_reduce_noalloc.argtypes = [vp, C.POINTER(C.c_double)]
        L.lbffs_synchronize.argtypes = [vp]
        L.lbffs_last_step_ms.argtypes = [vp, C.POINTER(C.c_float)]
        L.lbffs_destroy.argtypes = [vp]
        L.lbffs_strerror.argtypes = [ip]
        L.lbffs_strerror.restype = C.c_char_p
        L.lbffs_last_error.restype = C.c_char_p
        L.lbffs_case_units.argtypes = [C.POINTER(CaseIn), C.POINTER(Units)]
        L.lbffs_case_faces.argtypes = [C.POINTER(CaseIn), C.POINTER(Units)] + [vp] * 7
        L.lbffs_case_init_feq.argtypes = [ip, ip, ip, ip, C.c_float, C.c_float, C.c_float, C.c_float, vp]
        L.lbffs_case_neisolid.argtypes = [ip, ip, ip, ip, vp, vp]
        L.lbffs_case_outlet_sponge.argtypes = [ip] * 10 + [C.c_float, vp]
        L.lbffs_case_find_points.argtypes = [ip, vp, vp, vp, vp, vp, ip, ip, ip, vp, C.c_long, C.POINTER(C.c_long)]
        L.lbffs_probe_copy.argtypes = [sz, ip, C.POINTER(C.c_float)]
        L.lbffs_probe_bandwidth.argtypes = [ip, sz, ip, C.POINTER(C.c_float)]
        L.lbffs_probe_stream.argtypes = [ip, ip, ip, ip, ip, sz, ip, C.POINTER(C.c_float)]
        L.lbffs_probe_lattice.argtypes = [ip] * 7 + [C.POINTER(C.c_float)]
        L.lbffs_autotune.argtypes = [vp, C.POINTER(StepParams), ip, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_float)]
        L.lbffs_set_launch.argtypes = [vp, ip, ip]
        L.lbffs_ibm_set_markers.argtypes = [vp, vp, ip]
        L.lbffs_ibm_set_motion.argtypes = [vp, C.POINTER(IbmMotion)]
        L.lbffs_ibm_download.argtypes = [vp, ip, vp, ip]
        L.lbffs_batch_create.argtypes = [C.POINTER(Desc), ip, C.POINTER(vp)]
        L.lbffs_batch_context.argtypes = [vp, ip, C.POINTER(vp)]
        L.lbffs_batch_set_launch.argtypes = [vp, ip, ip]
        L.lbffs_batch_step.argtypes = [vp, ip, C.POINTER(StepParams)]
        L.lbffs_batch_last_step_ms.argtypes = [vp, C.POINTER(C.c_float)]
        L.lbffs_batch_autotune.argtypes = [vp, C.POINTER(StepParams), ip, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_float)]
        L.lbffs_batch_destroy.argtypes = [vp]
        fl = C.c_float
        L.lbffs_geom_solid.argtypes = [ip, vp, vp, vp, ip, ip, ip, ip, vp]
        L.lbffs_geom_sdf.argtypes = [ip, vp, vp, fl, fl, fl, ip, ip, ip, vp, vp]
        L.lbffs_geom_qlinks.argtypes = [ip, ip, vp, vp, vp, ip, ip, ip, vp, vp, vp, C.POINTER(ip)]
        L.lbffs_geom_qlinks_ctx.argtypes = [vp, ip, vp, vp, vp, fl, C.POINTER(ip)]
        L.lbffs_case_sdf_near_bwalls.argtypes = [fl, vp, vp, vp, ip, ip, ip, vp]
        L.lbffs_output_macros.argtypes = [vp, fl, vp, vp, vp, vp, sz]
        L.lbffs_out_write_vtk.argtypes = [C.c_char_p, vp, ip]
        L.lbffs_out_write_restart.argtypes = [C.c_char_p, ip, ip, ip, ip, ip, vp]
        L.lbffs_out_read_restart.argtypes = [C.c_char_p, ip] + [C.POINTER(ip)] * 4 + [vp, sz]
        L.lbffs_out_stats_compute.argtypes = [vp, vp]
        L.lbffs_out_coeffs_fp64.argtypes = [vp, fl, fl, fl, fl, vp]
        L.lbffs_out_accumulate_mean.argtypes = [C.c_long, ip, vp, vp, vp, vp, vp, vp, vp]
        for n in ("lbffs_create", "lbffs_upload", "lbffs_download", "lbffs_step", "lbffs_reduce", "lbffs_reduce_noalloc",
                  "lbffs_synchronize", "lbffs_last_step_ms", "lbffs_destroy", "lbffs_case_units",
                  "lbffs_case_faces", "lbffs_case_init_feq", "lbffs_case_neisolid", "lbffs_case_outlet_sponge",
                  "lbffs_case_find_points",
                  "lbffs_probe_copy", "lbffs_probe_bandwidth", "lbffs_probe_stream", "lbffs_probe_lattice",
                  "lbffs_autotune", "lbffs_set_launch",
                  "lbffs_ibm_set_markers", "lbffs_ibm_set_motion", "lbffs_ibm_download", "lbffs_batch_create",
                  "lbffs_batch_context", "lbffs_batch_set_launch", "lbffs_batch_step", "lbffs_batch_last_step_ms",
                  "lbffs_batch_autotune", "lbffs_batch_destroy", "lbffs_geom_solid", "lbffs_geom_sdf",
                  "lbffs_geom_qlinks", "lbffs_geom_qlinks_ctx",
                  "lbffs_case_sdf_near_bwalls",
                  "lbffs_output_macros", "lbffs_out_write_vtk", "lbffs_out_write_restart", "lbffs_out_read_restart",
                  "lbffs_out_stats_compute", "lbffs_out_coeffs_fp64", "lbffs_out_accumulate_mean"):
            getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        L = lib()
        raise LbffsError(f"{what}: {L.lbffs_strerror(rc).decode()} ({rc}): {L.lbffs_last_error().decode()}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _params(dpdx=0.0, LES=0.0, omegaB=0.1, spzWidth=0.1, Cs=0.1, damping=1, store_every=0, clim=0.01, ibb=0):
    return StepParams(dpdx=dpdx, LES=LES, omegaB=omegaB, spzWidth=spzWidth, Cs=Cs, damping=damping,
                      store_every=store_every, clim=clim, ibb=int(ibb))


def probe_copy(nbytes, reps=20):
    """Milliseconds of one device copy of nbytes (lbffs_probe_copy)."""
    ms = C.c_float()
    _check(lib().lbffs_probe_copy(int(nbytes), int(reps), C.byref(ms)), "lbffs_probe_copy")
    return ms.value


def probe_bandwidth(mode, nbytes, reps=10):
    """Milliseconds of one pass of probe `mode` over nbytes (lbffs_probe_bandwidth)."""
    ms = C.c_float()
    _check(lib().lbffs_probe_bandwidth(int(mode), int(nbytes), int(reps), C.byref(ms)), "lbffs_probe_bandwidth")
    return ms.value


def probe_stream(kind, unroll, nt, block, grid, nbytes, reps=10):
    """Milliseconds of one pass of the configurable streaming probe (lbffs_probe_stream):
    kind 0 copy, 1 read, 2 write."""
    ms = C.c_float()
    _check(lib().lbffs_probe_stream(int(kind), int(unroll), int(nt), int(block), int(grid), int(nbytes), int(reps),
                                    C.byref(ms)), "lbffs_probe_stream")
    return ms.value


def probe_lattice(q, store_bytes, variant, nx, ny, nz, reps=10):
    """Milliseconds of one pass of the lattice-step memory skeleton (lbffs_probe_lattice)."""
    ms = C.c_float()
    _check(lib().lbffs_probe_lattice(int(q), int(store_bytes), int(variant), int(nx), int(ny), int(nz), int(reps),
                                     C.byref(ms)), "lbffs_probe_lattice")
    return ms.value


class Lattice:
    """One device context (lbffs_create .. lbffs_destroy)."""

    def __init__(self, nx, ny, nz, q=19, storage=STORE_F32, sgs=SGS_WALE, collision=COLL_RR, omega0=1.0,
                 block=0, rounding=0, streaming=STREAM_TWO_LATTICE):
        self.nx, self.ny, self.nz, self.q = nx, ny, nz, q
        self.N = nx * ny * nz
        self.storage = storage
        self.real = np.float64 if storage == STORE_F64 else np.float32
        d = Desc(abi_version=ABI_VERSION, q=q, nx=nx, ny=ny, nz=nz, storage=storage, sgs=sgs,
                 collision=collision, omega0=omega0, block=block, rounding=rounding, streaming=streaming)
        h = C.c_void_p()
        _check(lib().lbffs_create(C.byref(d), C.byref(h)), "lbffs_create")
        self._h = h

    @classmethod
    def _view(cls, handle, nx, ny, nz, q, storage):
        """A Lattice on a context owned by a Batch (closing it does not destroy it)."""
        self = cls.__new__(cls)
        self.nx, self.ny, self.nz, self.q = nx, ny, nz, q
        self.N = nx * ny * nz
        self.storage = storage
        self.real = np.float64 if storage == STORE_F64 else np.float32
        self._h = handle
        self._owner = False
        return self

    def _dtype(self, name):
        if name in INT_FIELDS:
            return np.int32
        if name in U8_FIELDS:
            return np.uint8
        return self.real

    def upload(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=self._dtype(name))
        _check(lib().lbffs_upload(self._h, FIELD_ID[name], _ptr(a), a.nbytes), f"upload {name}")

    def download(self, name):
        n = self.q * self.N if name in ("f", "qlink") else self.N
        a = np.empty(n, dtype=self._dtype(name))
        _check(lib().lbffs_download(self._h, FIELD_ID[name], _ptr(a), a.nbytes), f"download {name}")
        return a

    def step(self, nsteps, **params):
        p = _params(**params)
        _check(lib().lbffs_step(self._h, int(nsteps), C.byref(p)), "lbffs_step")

    def autotune(self, trial_steps=10, **params):
        """Pick the fastest launch shape on this lattice (state is restored)."""
        p = _params(**params)
        b, t, ms = C.c_int(), C.c_int(), C.c_float()
        _check(lib().lbffs_autotune(self._h, C.byref(p), int(trial_steps), C.byref(b), C.byref(t), C.byref(ms)),
               "lbffs_autotune")
        return {"block": b.value, "tune": t.value, "ms_per_step": ms.value}

    def set_launch(self, block=256, tune=0):
        _check(lib().lbffs_set_launch(self._h, int(block), int(tune)), "lbffs_set_launch")

    def ibm_set_markers(self, xyz):
        """Marker positions (n x 3 or 3n, lattice coordinates); n <= 1 switches IBM off."""
        a = np.ascontiguousarray(np.asarray(xyz, dtype=self.real).reshape(-1))
        _check(lib().lbffs_ibm_set_markers(self._h, _ptr(a), a.size // 3), "lbffs_ibm_set_markers")
        self.n_markers = a.size // 3

    def ibm_set_motion(self, trans=(0.0, 0.0, 0.0), center=(0.0, 0.0, 0.0), axis=(0.0, 0.0, 0.0), omega=0.0):
        m = IbmMotion()
        for a in range(3):
            m.trans[a], m.rot_center[a], m.rot_axis[a] = trans[a], center[a], axis[a]
        m.rot_omega = omega
        _check(lib().lbffs_ibm_set_motion(self._h, C.byref(m)), "lbffs_ibm_set_motion")

    def ibm_download(self, what=IBM_POSITIONS):
        n = getattr(self, "n_markers", 0)
        a = np.empty(3 * n, dtype=self.real)
        _check(lib().lbffs_ibm_download(self._h, int(what), _ptr(a), n), "lbffs_ibm_download")
        return a

    def reduce(self):
        out = (C.c_double * REDUCE_N)()
        _check(lib().lbffs_reduce(self._h, out), "lbffs_reduce")
        return np.array(out[:], dtype=np.float64)

    def output_macros(self, dpdx=0.0):
        """rho, u, v, w from the current populations, as the host computes them at an
        output step (cal_rhoUVW, D3Q19_opencl/D3Q19.hpp:647-666)."""
        dt = self._dtype("rho")
        arrs = [np.empty(self.N, dt) for _ in range(4)]
        _check(lib().lbffs_output_macros(self._h, float(dpdx), *[_ptr(a) for a in arrs], arrs[0].nbytes),
               "lbffs_output_macros")
        return dict(zip(("rho", "u", "v", "w"), arrs))

    def synchronize(self):
        _check(lib().lbffs_synchronize(self._h), "lbffs_synchronize")

    def last_step_ms(self):
        ms = C.c_float()
        _check(lib().lbffs_last_step_ms(self._h, C.byref(ms)), "lbffs_last_step_ms")
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owner", True):
                lib().lbffs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Many independent lattices advanced by one launch per step (lbffs_batch_*).

    lattices: sequence of (nx, ny, nz, omega0); all share q, storage, sgs, collision.
    ``batch.lattice[i]`` is a Lattice view for uploads/downloads of lattice i."""

    def __init__(self, lattices, q=19, storage=STORE_F32, sgs=SGS_WALE, collision=COLL_RR):
        n = len(lattices)
        descs = (Desc * n)()
        for i, (nx, ny, nz, om) in enumerate(lattices):
            descs[i] = Desc(abi_version=ABI_VERSION, q=q, nx=nx, ny=ny, nz=nz, storage=storage, sgs=sgs,
                            collision=collision, omega0=om, block=0)
        h = C.c_void_p()
        _check(lib().lbffs_batch_create(descs, n, C.byref(h)), "lbffs_batch_create")
        self._h = h
        self.lattice = []
        for i, (nx, ny, nz, _) in enumerate(lattices):
            ch = C.c_void_p()
            _check(lib().lbffs_batch_context(self._h, i, C.byref(ch)), "lbffs_batch_context")
            self.lattice.append(Lattice._view(ch, nx, ny, nz, q, storage))
        self.N = sum(l.N for l in self.lattice)

    def set_launch(self, grid=0, graph_chunk=0):
        _check(lib().lbffs_batch_set_launch(self._h, int(grid), int(graph_chunk)), "lbffs_batch_set_launch")

    def step(self, nsteps, params):
        """params: one dict of step scalars per lattice (or one dict for all)."""
        n = len(self.lattice)
        plist = params if isinstance(params, (list, tuple)) else [params] * n
        arr = (StepParams * n)(*[_params(**p) for p in plist])
        _check(lib().lbffs_batch_step(self._h, int(nsteps), arr), "lbffs_batch_step")

    def last_step_ms(self):
        ms = C.c_float()
        _check(lib().lbffs_batch_last_step_ms(self._h, C.byref(ms)), "lbffs_batch_last_step_ms")
        return ms.value

    def autotune(self, params, trial_steps=10):
        """Time the batch's launch shapes, keep the fastest, leave every lattice unchanged.

        Returns (grid, graph_chunk, ms per step) of the kept shape."""
        n = len(self.lattice)
        plist = params if isinstance(params, (list, tuple)) else [params] * n
        arr = (StepParams * n)(*[_params(**p) for p in plist])
        g, ch, ms = C.c_int(), C.c_int(), C.c_float()
        _check(lib().lbffs_batch_autotune(self._h, arr, int(trial_steps), C.byref(g), C.byref(ch), C.byref(ms)),
               "lbffs_batch_autotune")
        return g.value, ch.value, ms.value

    def close(self):
        if getattr(self, "_h", None):
            for l in self.lattice:
                l._h = None
            lib().lbffs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
