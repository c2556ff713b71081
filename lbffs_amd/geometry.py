"""Geometry preprocessing: from the walls STL to the solid flags, the solid-neighbour
flags and the wall distances that the step kernels read.

The sequence is the reference host's (D3Q19_opencl/ns.cpp:188-226, the same in
D3Q27_opencl/ns.cpp): read the STL scaled by the cell size L (src/stl.hpp:456-557),
setSolid, setNeiSolid, setSDF for every patch when LES is on (dr 10, p 7, sdfIni 10000),
then setSDFnearBwalls with dr 3.  setSolid and setSDF run as HIP kernels
(``include/lbffs_geom.h``); setNeiSolid and setSDFnearBwalls are host functions
(``include/lbffs_case.h``).
"""
from __future__ import annotations

import ctypes as C
import gzip
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import _check, _ptr, lib
from .cases import neisolid

SDF_INI = 10000.0  # ns.cpp:201
SDF_DR = 10.0      # ns.cpp:202
SDF_P = 7.0        # ns.cpp:203
BWALL_DR = 3.0     # ns.cpp:221

_libc = C.CDLL(None)
_libc.strtof.restype = C.c_float
_libc.strtof.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]


def _stof(tok: str) -> np.float32:
    """std::stof: the C library's strtof, an error when no character converts."""
    b = tok.encode()
    end = C.c_char_p()
    v = _libc.strtof(b, C.byref(end))
    rest = end.value if end.value is not None else b""
    if len(rest) == len(b):
        raise ValueError(f"stof: no conversion of {tok!r}")
    return np.float32(v)


_SPACE = re.compile(r"[ \t\n\v\f\r]+")


def tokens(text: str) -> list:
    """input.hpp:90-110 readToLines: every line (std::getline) split at runs of white
    space (boost::split, token_compress_on), so an indented line starts with an empty
    token and a line with trailing blanks ends with one."""
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    out = []
    for line in lines:
        out.extend(_SPACE.split(line))
    return out


@dataclass
class Patch:
    """STLpatch (stl.hpp:85-315): rows x, y, z of the vertices, normals and centroids."""
    name: str
    v0: np.ndarray = None
    v1: np.ndarray = None
    v2: np.ndarray = None
    normal: np.ndarray = None
    c: np.ndarray = None

    @property
    def n(self):
        return self.v0.shape[1]


@dataclass
class STL:
    patches: list = field(default_factory=list)

    def triangles(self):
        """Vertices of every patch, concatenated: three (3, n) float32 arrays."""
        if not self.patches:
            z = np.zeros((3, 0), np.float32)
            return z, z.copy(), z.copy()
        return tuple(np.ascontiguousarray(np.concatenate([getattr(p, a) for p in self.patches], axis=1))
                     for a in ("v0", "v1", "v2"))


def parse_stl(text: str, L: float) -> STL:
    """stl.hpp:461-517 STL::readSTL on the file's text, then the scaling by 1/L, the
    centroids (v0 + v1 + v2) / 3 and (in the kernels) the bounding boxes, all FP32."""
    tok = tokens(text)
    patches = [Patch(tok[i + 1]) for i, t in enumerate(tok) if t == "solid"]
    rows = [{"n": [], "v0": [], "v1": [], "v2": []} for _ in patches]
    ip = 0
    i = 0
    while i < len(tok):
        if tok[i] == "solid" and tok[i + 1] != patches[ip].name:
            ip += 1
            i += 1
        if tok[i] == "facet":
            rows[ip]["n"].append([_stof(tok[i + 2]), _stof(tok[i + 3]), _stof(tok[i + 4])])
            i += 4
        elif tok[i] == "vertex":
            rows[ip]["v0"].append([_stof(tok[i + 1]), _stof(tok[i + 2]), _stof(tok[i + 3])])
            i += 5
            rows[ip]["v1"].append([_stof(tok[i + 1]), _stof(tok[i + 2]), _stof(tok[i + 3])])
            i += 5
            rows[ip]["v2"].append([_stof(tok[i + 1]), _stof(tok[i + 2]), _stof(tok[i + 3])])
            i += 3
        i += 1
    Lf = np.float32(L)
    for p, r in zip(patches, rows):
        def arr(key):
            return np.array(r[key], np.float32).reshape(-1, 3).T.copy()
        p.normal = arr("n")
        p.v0, p.v1, p.v2 = arr("v0") / Lf, arr("v1") / Lf, arr("v2") / Lf
        p.c = ((p.v0 + p.v1) + p.v2) / np.float32(3.0)
    return STL(patches)


def read_stl(path: str, L: float) -> STL:
    """STL(fileName, L) (stl.hpp:520-556): a missing file reads as a surface without
    patches.  A path ending in .gz is decompressed first."""
    if not os.path.exists(path):
        return STL([])
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as fh:
        return parse_stl(fh.read().decode("latin-1"), L)


def write_stl(stl: STL, path: str, L: float = 1.0) -> None:
    """ASCII STL of the patches, vertices and normals times L, in the indented layout the
    reader's index steps expect (one solid ... endsolid block per patch)."""
    with open(path, "w") as fh:
        for p in stl.patches:
            fh.write(f"solid {p.name}\n")
            v = [np.asarray(x, np.float64) * L for x in (p.v0, p.v1, p.v2)]
            nrm = np.asarray(p.normal, np.float64)
            for t in range(p.n):
                fh.write("  facet normal %.9g %.9g %.9g\n    outer loop\n" % tuple(nrm[:, t]))
                for a in v:
                    fh.write("      vertex %.9g %.9g %.9g\n" % tuple(a[:, t]))
                fh.write("    endloop\n  endfacet\n")
            fh.write(f"endsolid {p.name}\n")


def procedural_body(centre, R: float, n_lat: int = 250, n_lon: int = 400, amp: float = 0.02, modes: int = 8,
                    seed: int = 20251202, name: str = "body") -> STL:
    """A closed, seeded bluff body in lattice units: a sphere of radius R whose radius
    varies smoothly with the direction n, r(n) = R (1 + amp g(n)) with
    g = sum_k a_k sin(w_k d_k.n + p_k) / sum_k |a_k| (random unit d_k, w_k in [1, 4]),
    meshed on n_lat latitude bands (poles on the z axis, longitudes turned by a seeded
    angle) into 2 n_lon (n_lat - 1) triangles with outward winding; one patch."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((modes, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.uniform(1.0, 4.0, modes)
    ph = rng.uniform(0.0, 2 * np.pi, modes)
    a = rng.uniform(-1.0, 1.0, modes)
    a /= np.abs(a).sum()
    turn = rng.uniform(0.0, 2 * np.pi / n_lon)
    th = np.linspace(0.0, np.pi, n_lat + 1)[1:-1]
    lon = turn + 2 * np.pi * np.arange(n_lon) / n_lon
    T, P = np.meshgrid(th, lon, indexing="ij")
    n = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1)  # (n_lat-1, n_lon, 3)
    poles = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])

    def radius(u):
        return R * (1.0 + amp * np.sin(np.tensordot(u, d.T, axes=1) * w + ph) @ a)

    c = np.asarray(centre, np.float64)
    ring = c + n * radius(n)[..., None]
    pole = c + poles * radius(poles)[:, None]
    b = np.arange(n_lon)
    b1 = (b + 1) % n_lon
    tris = [np.stack([np.broadcast_to(pole[0], (n_lon, 3)), ring[0, b], ring[0, b1]], axis=1)]
    for r in range(n_lat - 2):
        tris.append(np.stack([ring[r, b], ring[r + 1, b], ring[r + 1, b1]], axis=1))
        tris.append(np.stack([ring[r, b], ring[r + 1, b1], ring[r, b1]], axis=1))
    tris.append(np.stack([ring[-1, b], np.broadcast_to(pole[1], (n_lon, 3)), ring[-1, b1]], axis=1))
    t = np.concatenate(tris).astype(np.float32)       # (n, 3 vertices, 3 axes)
    v0, v1, v2 = (np.ascontiguousarray(t[:, k, :].T) for k in range(3))
    cr = np.cross((v1 - v0).T.astype(np.float64), (v2 - v0).T.astype(np.float64))
    nrm = (cr / np.linalg.norm(cr, axis=1, keepdims=True)).T.astype(np.float32)
    p = Patch(name, v0, v1, v2, normal=np.ascontiguousarray(nrm))
    p.c = ((p.v0 + p.v1) + p.v2) / np.float32(3.0)
    return STL([p])


def frontal_area(stl: STL, axis: int = 0) -> float:
    """Area of a closed surface's projection on the plane normal to `axis`:
    half the sum of |n_axis| times the triangle areas (FP64)."""
    tot = 0.0
    for p in stl.patches:
        v0, v1, v2 = (np.asarray(x, np.float64) for x in (p.v0, p.v1, p.v2))
        cr = np.cross((v1 - v0).T, (v2 - v0).T)
        tot += 0.25 * float(np.abs(cr[:, axis]).sum())
    return tot


def solid_from_triangles(v0, v1, v2, nx: int, ny: int, nz: int, invert: bool = False, solid=None) -> np.ndarray:
    """walls.hpp:159-184 setSolid on the GPU for (3, n) float32 vertex rows; returns the
    nx*ny*nz uint8 flags (a copy of `solid` updated, or fresh zeros)."""
    v = [np.ascontiguousarray(x, np.float32) for x in (v0, v1, v2)]
    assert all(x.ndim == 2 and x.shape[0] == 3 and x.shape == v[0].shape for x in v)
    s = np.zeros(nx * ny * nz, np.uint8) if solid is None else np.ascontiguousarray(solid, np.uint8).copy()
    assert s.size == nx * ny * nz
    _check(lib().lbffs_geom_solid(v[0].shape[1], *[_ptr(x) for x in v], nx, ny, nz, int(bool(invert)), _ptr(s)),
           "lbffs_geom_solid")
    return s


def set_solid(stl: STL, nx: int, ny: int, nz: int, invert: bool = False, solid=None) -> np.ndarray:
    """setSolid(solid, stl, nx, ny, nz, invert) over every patch of the STL."""
    return solid_from_triangles(*stl.triangles(), nx, ny, nz, invert, solid)


def set_sdf(patch: Patch, sdf: np.ndarray, nei: np.ndarray, nx: int, ny: int, nz: int, dr: float = SDF_DR,
            p: float = SDF_P, sdf_ini: float = SDF_INI) -> np.ndarray:
    """walls.hpp:218-297 setSDF (patch form) on the GPU; updates and returns sdf."""
    assert sdf.dtype == np.float32 and sdf.flags.c_contiguous and sdf.size == nx * ny * nz
    nei = np.ascontiguousarray(nei, np.uint8)
    assert nei.size == nx * ny * nz
    c = np.ascontiguousarray(patch.c, np.float32)
    nrm = np.ascontiguousarray(patch.normal, np.float32)
    _check(lib().lbffs_geom_sdf(patch.n, _ptr(c), _ptr(nrm), sdf_ini, dr, p, nx, ny, nz, _ptr(nei), _ptr(sdf)),
           "lbffs_geom_sdf")
    return sdf


def qlinks(stl: STL, q: int, nx: int, ny: int, nz: int, solid, nei, init: float = 0.5):
    """q-distances of the solid links from the surface in FP64 (lbffs_geom_qlinks), for the
    interpolated bounce-back (Lattice field "qlink", step parameter ibb=1).  Returns the
    q*N float32 array (init where a link has no solid neighbour or meets no triangle) and
    the number of solid links that met no triangle."""
    v = [np.ascontiguousarray(x, np.float32) for x in stl.triangles()]
    s = np.ascontiguousarray(solid, np.uint8)
    n8 = np.ascontiguousarray(nei, np.uint8)
    assert s.size == n8.size == nx * ny * nz
    ql = np.full(q * nx * ny * nz, init, np.float32)
    miss = C.c_int()
    _check(lib().lbffs_geom_qlinks(q, v[0].shape[1], *[_ptr(x) for x in v], nx, ny, nz, _ptr(s), _ptr(n8), _ptr(ql),
                                   C.byref(miss)), "lbffs_geom_qlinks")
    return ql, miss.value


def find_points(patch: Patch, nx: int, ny: int, nz: int) -> np.ndarray:
    """stl.hpp:317-453 findPoints (lbffs_case_find_points): the lattice face cells the
    patch's triangles cover, triangle by triangle in the reference's order (a cell inside
    two triangles appears twice)."""
    rows = [np.ascontiguousarray(x, np.float32) for x in (patch.v0, patch.v1, patch.v2, patch.normal, patch.c)]
    n = rows[0].shape[1]
    cnt = C.c_long()
    _check(lib().lbffs_case_find_points(n, *[_ptr(r) for r in rows], nx, ny, nz, None, 0, C.byref(cnt)),
           "lbffs_case_find_points")
    out = np.empty(cnt.value, np.int32)
    if cnt.value:
        _check(lib().lbffs_case_find_points(n, *[_ptr(r) for r in rows], nx, ny, nz, _ptr(out), cnt.value,
                                            C.byref(cnt)), "lbffs_case_find_points")
    return out


def qlinks_into(lat, stl: STL, init: float = 0.5) -> int:
    """The same q-distances cast straight into the lattice's "qlink" field on the device
    (lbffs_geom_qlinks_ctx; the lattice's uploaded solid and neiSolid flags), without a
    q*N host array.  Returns the number of solid links that met no triangle."""
    v = [np.ascontiguousarray(x, np.float32) for x in stl.triangles()]
    miss = C.c_int()
    _check(lib().lbffs_geom_qlinks_ctx(lat._h, v[0].shape[1], *[_ptr(x) for x in v], init, C.byref(miss)),
           "lbffs_geom_qlinks_ctx")
    return miss.value


def sdf_near_bwalls(sdf: np.ndarray, b1, b2, b3, nx: int, ny: int, nz: int, dr: float = BWALL_DR) -> np.ndarray:
    """walls.hpp:372-469 setSDFnearBwalls (host); updates and returns sdf."""
    assert sdf.dtype == np.float32 and sdf.flags.c_contiguous and sdf.size == nx * ny * nz
    b = [np.ascontiguousarray(x, np.int32) for x in (b1, b2, b3)]
    _check(lib().lbffs_case_sdf_near_bwalls(dr, *[_ptr(x) for x in b], nx, ny, nz, _ptr(sdf)),
           "lbffs_case_sdf_near_bwalls")
    return sdf


def preprocess(stl: STL, q: int, nx: int, ny: int, nz: int, LES: bool, b1, b2, b3, invert: bool = False) -> dict:
    """ns.cpp:188-226: solid, neiSolid and sdf of a case from its walls STL."""
    solid = set_solid(stl, nx, ny, nz, invert)
    nei = neisolid(q, nx, ny, nz, solid)
    sdf = np.full(nx * ny * nz, SDF_INI, np.float32)
    if LES:
        for p in stl.patches:
            set_sdf(p, sdf, nei, nx, ny, nz)
        sdf_near_bwalls(sdf, b1, b2, b3, nx, ny, nz)
    return {"solid": solid, "neiSolid": nei, "sdf": sdf}
