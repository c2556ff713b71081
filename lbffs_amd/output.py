"""The output step of the LBFFS driver: VTK and restart files, diagnostics, running means.

The reference host runs write.hpp at the top of every time-loop iteration
(D3Q19_opencl/ns.cpp:509-511): when the output time is reached it downloads the
populations and side fields, recomputes rho and u from f, prints diagnostics (density
range and mean, inlet-plane mean velocity, force coefficients), adds u to the running
means, writes ``out/resNNNNNNNN.vtk`` and overwrites ``out/data.dat``.  ``OutputStep``
does the same around a ``Lattice``: rho and u come from ``lbffs_output_macros`` (device),
the files and the diagnostics' arithmetic from ``include/lbffs_output.h`` (host C++).
The FP64 device sums of ``lbffs_reduce`` give the force coefficients a second time,
free of the FP32 sums' rounding.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _check, _ptr, lib

VTK_ASCII, VTK_BINARY, VTK_BINARY_REF = 0, 1, 2

_fp = C.POINTER(C.c_float)


class VtkFields(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("L", C.c_float), ("c", C.c_float),
                ("timeMean", C.c_float)] + [(n, C.c_void_p) for n in (
                    "u", "v", "w", "rho", "tauSGS", "boundary1", "boundary2", "boundary3", "sdf", "solid",
                    "neiSolid", "GxIBM", "GyIBM", "GzIBM", "omega", "uMean", "vMean", "wMean")]


class StatsIn(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("c", C.c_float), ("L", C.c_float),
                ("uMax", C.c_float), ("Dref", C.c_float), ("forceCoeffs", C.c_int)] + [(n, C.c_void_p) for n in (
                    "rho", "u", "solid", "Fwx", "Fwy", "Fwz", "GxIBM", "GyIBM", "GzIBM")]


class Stats(C.Structure):
    _fields_ = [("rho_av", C.c_float), ("rhoMax", C.c_float), ("rhoMin", C.c_float), ("uAve", C.c_float),
                ("eleFluid", C.c_int), ("sumFw", C.c_float * 3), ("sumG", C.c_float * 3),
                ("Cx", C.c_double), ("Cy", C.c_double), ("Cz", C.c_double),
                ("CxIBM", C.c_double), ("CyIBM", C.c_double), ("CzIBM", C.c_double)]


VTK_ORDER = ("u", "v", "w", "rho", "nuSGS", "boundary1", "boundary2", "boundary3", "sdf", "solid", "neiSolid",
             "GxIBM", "GyIBM", "GzIBM", "omega", "uMean", "vMean", "wMean")

_DT = {"boundary1": np.int32, "boundary2": np.int32, "boundary3": np.int32, "solid": np.uint8, "neiSolid": np.uint8}


def write_vtk(path, nx, ny, nz, L, c, timeMean, arrays, mode=VTK_ASCII):
    """write.hpp:76-600.  arrays: u, v, w, rho, tauSGS, boundary1..3, sdf, solid, neiSolid,
    GxIBM, GyIBM, GzIBM, omega, uMean, vMean, wMean (lattice units, nx*ny*nz each)."""
    keep = {}
    F = VtkFields(nx=nx, ny=ny, nz=nz, L=L, c=c, timeMean=timeMean)
    for name, _ in VtkFields._fields_[6:]:
        a = np.ascontiguousarray(arrays[name], _DT.get(name, np.float32))
        assert a.size == nx * ny * nz, name
        keep[name] = a
        setattr(F, name, a.ctypes.data)
    _check(lib().lbffs_out_write_vtk(os.fsencode(path), C.byref(F), mode), f"write {path}")


def write_restart(path, nt, q, nx, ny, nz, f):
    """write.hpp:602-641 (data.dat)."""
    f = np.ascontiguousarray(f, np.float32)
    assert f.size == q * nx * ny * nz
    _check(lib().lbffs_out_write_restart(os.fsencode(path), int(nt), q, nx, ny, nz, _ptr(f)), f"write {path}")


def read_restart(path, q):
    """restart.hpp:1-28: returns (nt, nx, ny, nz, f)."""
    nt, nx, ny, nz = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _check(lib().lbffs_out_read_restart(os.fsencode(path), q, C.byref(nt), C.byref(nx), C.byref(ny), C.byref(nz),
                                        None, 0), f"read {path}")
    f = np.empty(q * nx.value * ny.value * nz.value, np.float32)
    _check(lib().lbffs_out_read_restart(os.fsencode(path), q, C.byref(nt), C.byref(nx), C.byref(ny), C.byref(nz),
                                        _ptr(f), f.size), f"read {path}")
    return nt.value, nx.value, ny.value, nz.value, f


def stats(nx, ny, nz, c, L, uMax, Dref, rho, u, solid, forces=None):
    """write.hpp:16-60, 93-99 with the reference's arithmetic.  forces: dict of Fwx, Fwy,
    Fwz, GxIBM, GyIBM, GzIBM when the force coefficients are wanted."""
    keep = [np.ascontiguousarray(rho, np.float32), np.ascontiguousarray(u, np.float32),
            np.ascontiguousarray(solid, np.uint8)]
    sin = StatsIn(nx=nx, ny=ny, nz=nz, c=c, L=L, uMax=uMax, Dref=Dref, forceCoeffs=int(forces is not None))
    sin.rho, sin.u, sin.solid = (a.ctypes.data for a in keep)
    if forces is not None:
        for n in ("Fwx", "Fwy", "Fwz", "GxIBM", "GyIBM", "GzIBM"):
            a = np.ascontiguousarray(forces[n], np.float32)
            keep.append(a)
            setattr(sin, n, a.ctypes.data)
    out = Stats()
    _check(lib().lbffs_out_stats_compute(C.byref(sin), C.byref(out)), "lbffs_out_stats_compute")
    return {k: (list(getattr(out, k)) if k in ("sumFw", "sumG") else getattr(out, k)) for k, _ in Stats._fields_}


def coeffs_fp64(sums, c, L, uMax, Dref):
    """Force coefficients from the FP64 device sums of lbffs_reduce."""
    s = (C.c_double * 8)(*[float(x) for x in sums[:8]])
    out = (C.c_double * 7)()
    _check(lib().lbffs_out_coeffs_fp64(s, c, L, uMax, Dref, out), "lbffs_out_coeffs_fp64")
    return dict(zip(("rho_av", "Cx", "Cy", "Cz", "CxIBM", "CyIBM", "CzIBM"), out[:]))


class MeanAccumulator:
    """The running means of write.hpp:68-74 over n values per component: add(u, v, w, interval)
    accumulates uMean += u * interval and timeMean += interval in FP32, as the reference
    host does; mean() returns (uMean, vMean, wMean) / timeMean."""

    def __init__(self, n):
        self.n = int(n)
        self.uMean, self.vMean, self.wMean = (np.zeros(self.n, np.float32) for _ in range(3))
        self.timeMean = np.zeros(1, np.float32)

    def add(self, u, v, w, interval):
        a = [np.ascontiguousarray(x, np.float32) for x in (u, v, w)]
        if any(x.size != self.n for x in a):
            raise ValueError("MeanAccumulator.add: arrays of the wrong size")
        _check(lib().lbffs_out_accumulate_mean(self.n, int(interval), _ptr(a[0]), _ptr(a[1]), _ptr(a[2]),
                                               _ptr(self.uMean), _ptr(self.vMean), _ptr(self.wMean),
                                               _ptr(self.timeMean)), "lbffs_out_accumulate_mean")

    def mean(self):
        t = float(self.timeMean[0])
        return tuple(m.astype(np.float64) / t for m in (self.uMean, self.vMean, self.wMean))


class OutputStep:
    """write.hpp around a Lattice: call ``step_begin(nt)`` at the top of every iteration of
    the time loop, as ns.cpp:509-511 includes write.hpp."""

    def __init__(self, lat, arrays, L, c, uMax, deltaT, out_dir="out", out_interval=10000, next_out=0,
                 mode=VTK_BINARY, force_coeffs=False, Dref=1.0, dpdx=0.0, log=print):
        self.lat, self.arrays = lat, arrays
        self.L, self.c, self.uMax, self.deltaT, self.Dref, self.dpdx = L, c, uMax, deltaT, Dref, dpdx
        self.out_dir, self.interval, self.next_out, self.mode = out_dir, out_interval, next_out, mode
        self.force_coeffs, self.log = force_coeffs, log
        N = lat.N
        self.uMean, self.vMean, self.wMean = (np.zeros(N, np.float32) for _ in range(3))
        self.timeMean = np.zeros(1, np.float32)
        os.makedirs(out_dir, exist_ok=True)

    def step_begin(self, nt):
        if not self.next_out < nt + 1:
            return None
        lat = self.lat
        m = lat.output_macros(self.dpdx)
        f = lat.download("f")
        tau, omega = lat.download("tauSGS"), lat.download("omega")
        forces = {n: lat.download(n) for n in ("Fwx", "Fwy", "Fwz", "GxIBM", "GyIBM", "GzIBM")}
        st = stats(lat.nx, lat.ny, lat.nz, self.c, self.L, self.uMax, self.Dref, m["rho"], m["u"],
                   self.arrays["solid"], forces if self.force_coeffs else None)
        _check(lib().lbffs_out_accumulate_mean(lat.N, self.interval, _ptr(m["u"]), _ptr(m["v"]), _ptr(m["w"]),
                                               _ptr(self.uMean), _ptr(self.vMean), _ptr(self.wMean),
                                               _ptr(self.timeMean)), "lbffs_out_accumulate_mean")
        self.log(f"uAve: {st['uAve']:g}")
        self.log(f"Time = {nt * self.deltaT:g} (s), {nt} steps")
        self.log(f"rhoMax: {st['rhoMax']:g}, rhoMin: {st['rhoMin']:g}, rhoAve: {st['rho_av']:g}")
        if self.force_coeffs:
            self.log(f"Cx: {st['Cx']:g}, Cy: {st['Cy']:g}, Cz: {st['Cz']:g}")
            self.log(f"CxIBM: {st['CxIBM']:g}, CyIBM: {st['CyIBM']:g}, CzIBM: {st['CzIBM']:g}")
            st["fp64"] = coeffs_fp64(lat.reduce(), self.c, self.L, self.uMax, self.Dref)
        fields = dict(m, tauSGS=tau, omega=omega, uMean=self.uMean, vMean=self.vMean, wMean=self.wMean,
                      GxIBM=forces["GxIBM"], GyIBM=forces["GyIBM"], GzIBM=forces["GzIBM"])
        for n in ("boundary1", "boundary2", "boundary3", "sdf", "solid", "neiSolid"):
            fields[n] = self.arrays[n]
        write_vtk(os.path.join(self.out_dir, f"res{nt:08d}.vtk"), lat.nx, lat.ny, lat.nz, self.L, self.c,
                  float(self.timeMean[0]), fields, self.mode)
        write_restart(os.path.join(self.out_dir, "data.dat"), nt, lat.q, lat.nx, lat.ny, lat.nz, f)
        self.next_out += self.interval
        return st
