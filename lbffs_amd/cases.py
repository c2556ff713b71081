"""Case builders: the per-cell arrays of the reference host for the workloads this
engine is measured and tested on.

Every builder goes through the host case setup of ``include/lbffs_case.h`` (units as
ns.cpp:104-128, setBoundaryConditions.hpp, the face flag of ns.cpp:164-177,
initialization.hpp, setNeiSolid) and then adds what the workload defines on top
(solid geometry, inlet noise, wall distance).  A ``Case`` holds the arrays in the
reference layout, ready for ``Lattice.upload`` or for the CPU oracle.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import BC, CaseIn, Units, lib, _check, _ptr, SGS_WALE, SGS_SMAGORINSKY, SGS_NONE, COLL_RR, COLL_CUMULANT

FACES = ["nxMin", "nxMax", "nyMin", "nyMax", "nzMin", "nzMax"]

SEED = 20251202  # seed of the stochastic inlet profiles


@dataclass
class Case:
    name: str
    nx: int
    ny: int
    nz: int
    arrays: dict
    omega0: float
    sgs: int = SGS_WALE
    params: dict = field(default_factory=dict)  # lbffs_step_params scalars
    units: dict = field(default_factory=dict)
    meta: dict = field(default_factory=dict)
    q: int = 19
    surface: object = None  # an STL whose q-distances are cast on the device after the upload
    collision: object = None  # LBFFS_COLL_*; None = the reference's (RR on D3Q19, cumulant on D3Q27)

    @property
    def coll(self):
        """The collision operator the case runs (an LBFFS_COLL_* value)."""
        if self.collision is not None:
            return int(self.collision)
        return COLL_CUMULANT if self.q == 27 else COLL_RR

    @property
    def N(self):
        return self.nx * self.ny * self.nz

    def upload_to(self, lat):
        for k, v in self.arrays.items():
            lat.upload(k, v)
        if self.surface is not None:
            from .geometry import qlinks_into
            self.meta["qlink_misses"] = qlinks_into(lat, self.surface)


def units(nx, ny, nz, Lx, uMax, U0, nu, rho0=1.0, dpdx=0.0, ini=(0.0, 0.0, 0.0), face_codes=None, face_vels=None):
    cin = CaseIn(nx=nx, ny=ny, nz=nz, Lx=Lx, uMax=uMax, rho0=rho0, U0=U0, nu=nu, dpdx=dpdx,
                 uIni=ini[0], vIni=ini[1], wIni=ini[2])
    for f, name in enumerate(FACES):
        codes = (face_codes or {}).get(name, ("Cyclic", "Cyclic", "Cyclic"))
        vel = (face_vels or {}).get(name, (0.0, 0.0, 0.0))
        for d in range(3):
            cin.face_code[f][d] = BC[codes[d]] if isinstance(codes[d], str) else int(codes[d])
            cin.face_vel[f][d] = vel[d]
    un = Units()
    _check(lib().lbffs_case_units(C.byref(cin), C.byref(un)), "lbffs_case_units")
    return cin, un


def build_faces(cin, un):
    N = cin.nx * cin.ny * cin.nz
    b = [np.zeros(N, np.int32) for _ in range(4)]
    v = [np.zeros(N, np.float32) for _ in range(3)]
    _check(lib().lbffs_case_faces(C.byref(cin), C.byref(un), *[_ptr(a) for a in b + v]), "lbffs_case_faces")
    return b, v


def init_feq(q, nx, ny, nz, rho, u, v, w):
    f = np.empty(q * nx * ny * nz, np.float32)
    _check(lib().lbffs_case_init_feq(q, nx, ny, nz, rho, u, v, w, _ptr(f)), "lbffs_case_init_feq")
    return f


def neisolid(q, nx, ny, nz, solid):
    s = np.ascontiguousarray(solid, np.uint8)
    n = np.zeros_like(s)
    _check(lib().lbffs_case_neisolid(q, nx, ny, nz, _ptr(s), _ptr(n)), "lbffs_case_neisolid")
    return n


def make_case(name, nx, ny, nz, Lx, uMax, U0, nu, face_codes, face_vels=None, rho0=1.0, dpdx=0.0,
              ini=(0.0, 0.0, 0.0), solid=None, sgs=SGS_WALE, q=19, **params):
    cin, un = units(nx, ny, nz, Lx, uMax, U0, nu, rho0, dpdx, ini, face_codes, face_vels)
    (bnd, b1, b2, b3), (u0, v0, w0) = build_faces(cin, un)
    N = nx * ny * nz
    f = init_feq(q, nx, ny, nz, rho0, un.uIni, un.vIni, un.wIni)
    sol = np.zeros(N, np.uint8) if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(-1)
    arrays = dict(f=f, boundary=bnd, boundary1=b1, boundary2=b2, boundary3=b3,
                  sdf=np.full(N, 10000.0, np.float32), solid=sol, neiSolid=neisolid(q, nx, ny, nz, sol),
                  u0=u0, v0=v0, w0=w0, omega=np.full(N, un.omega0, np.float32))
    p = dict(dpdx=float(un.dpdx), LES=0.0, omegaB=0.1, spzWidth=0.1, Cs=0.1, damping=1)
    p.update(params)
    if q == 27:
        p.setdefault("clim", 0.01)
    uu = {k: getattr(un, k) for k in ("L", "c", "Re", "nu", "dpdx", "omega0", "deltaT")}
    return Case(name, nx, ny, nz, arrays, float(un.omega0), sgs, p, uu, q=q)


def cavity(n=64, nz=None, u_lid=0.1, Re=100.0, sgs=SGS_WALE, q=19):
    """Lid-driven cavity: lid = nyMax FixedVelocity along +x, other faces half-way
    bounce-back (testCases/cavityFlow/boundaryConditions.txt:23-28).  Physical inputs are
    chosen so that the lattice lid speed is u_lid and nu = u_lid*n/Re (Lx=1, uMax=1)."""
    nz = n if nz is None else nz
    U0 = u_lid  # uMax/c = U0 -> lid 1 m/s maps to u_lid
    nu_lat = u_lid * n / Re
    nu = nu_lat * (1.0 / (n - 1)) * (1.0 / U0)  # inverse of nu/(L c), ns.cpp:111
    codes = {f: ("BounceBack", "BounceBack", "BounceBack") for f in FACES}
    codes["nyMax"] = ("BounceBack", "FixedVelocity", "BounceBack")
    if nz == 1:
        codes["nzMin"] = codes["nzMax"] = ("Cyclic", "Cyclic", "Cyclic")
    vels = {"nyMax": (1.0, 0.0, 0.0)}
    return make_case(f"cavity{n}", n, n, nz, 1.0, 1.0, U0, nu, codes, vels, sgs=sgs, q=q)


def periodic_box(nx, ny, nz, tau=0.8, ini=(0.0, 0.0, 0.0), sgs=SGS_NONE, q=19):
    """All faces Cyclic: a fully periodic lattice (faces still take the face path)."""
    nu_lat = (tau - 0.5) / 3.0
    U0 = 0.1
    nu = nu_lat * (1.0 / (nx - 1)) * (1.0 / U0)
    # lattice velocity = physical / c with c = uMax/U0 (ns.cpp:105, 113-115)
    return make_case(f"box{nx}x{ny}x{nz}", nx, ny, nz, 1.0, 1.0, U0, nu, {}, ini=tuple(x / U0 for x in ini),
                     sgs=sgs, q=q)


def poiseuille(nx=100, ny=20, nz=1, dpdx=0.08, nu=0.01, Lx=5.0, U0=0.05, sgs=SGS_NONE):
    """testCases/poiseuilleFlow: periodic x, half-way bounce-back walls in y, body force."""
    codes = {f: ("Cyclic", "Cyclic", "Cyclic") for f in FACES}
    codes["nyMin"] = ("Cyclic", "BounceBack", "Cyclic")
    codes["nyMax"] = ("Cyclic", "BounceBack", "Cyclic")
    return make_case("poiseuille", nx, ny, nz, Lx, 1.0, U0, nu, codes, dpdx=dpdx, sgs=sgs)


def channel(nx=8, ny=20, nz=1, tau=0.8, force=1e-5, sgs=SGS_NONE, q=19):
    """Force-driven periodic channel in lattice units: half-way bounce-back walls in y,
    body force `force` along x (inverse of the conversions of ns.cpp:104-112)."""
    U0 = 0.1
    L = 1.0 / (nx - 1)
    c = 1.0 / U0
    nu = (tau - 0.5) / 3.0 * L * c
    dpdx = force * c * c / L
    codes = {f: ("Cyclic", "Cyclic", "Cyclic") for f in FACES}
    codes["nyMin"] = ("Cyclic", "BounceBack", "Cyclic")
    codes["nyMax"] = ("Cyclic", "BounceBack", "Cyclic")
    return make_case("channel", nx, ny, nz, 1.0, 1.0, U0, nu, codes, dpdx=dpdx, sgs=sgs, q=q)


def apply_patch(case, points, codes, vel=(0.0, 0.0, 0.0), spz=None):
    """The patch loops of ns.cpp:278-377 for one inlet or outlet patch: the listed face
    cells (geometry.find_points) get the patch's face codes (names or numbers, per axis)
    and velocity (lattice units, as inputBoundaryCondition returns it after dividing by c);
    an outlet patch with spz cells gets the cosine omega sponge around each cell
    (lbffs_case_outlet_sponge, whose tau0 is the 0.55 of ns.cpp:327).  Updates case.arrays
    in place."""
    a = case.arrays
    p = [BC[cd] if isinstance(cd, str) else int(cd) for cd in codes]
    pts = np.asarray(points, np.int64)
    for d, key in enumerate(("boundary1", "boundary2", "boundary3")):
        a[key][pts] = p[d]
    for d, key in enumerate(("u0", "v0", "w0")):
        a[key][pts] = np.float32(vel[d])
    if spz is not None:
        nx, ny = case.nx, case.ny
        om = a["omega"]
        for ic in pts.tolist():
            i, j, k = ic % nx, (ic // nx) % ny, ic // (nx * ny)
            _check(lib().lbffs_case_outlet_sponge(case.nx, case.ny, case.nz, i, j, k, p[0], p[1], p[2], int(spz),
                                                  float(case.omega0), _ptr(om)), "lbffs_case_outlet_sponge")
    return case


def cylinder_qlinks(q, nx, ny, nz, solid, cx, cy, R):
    """q-distances of the links that leave a fluid cell towards a solid cell of the cylinder
    (axis along z through (cx, cy), radius R), in FP64: the link from the fluid node x to
    x - c_q meets the circle at x - t c_q with t the smaller root of |p - t d|^2 = R^2
    (p = x - centre, d = c_q in the x-y plane); q = t.  Q*N float32, q-major, 0.5 where no
    solid link (the reference's qfIni, ns.cpp:215)."""
    from . import BC  # noqa: F401  (keeps the import list of this module in one place)
    cxs, cys = (np.array(v) for v in _dirs(q)[:2])
    N = nx * ny * nz
    out = np.full((q, N), 0.5, np.float64)
    sol = np.asarray(solid, np.uint8).reshape(nz, ny, nx).astype(bool)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    px, py = i - float(cx), j - float(cy)
    fluid = ~sol
    for d in range(1, q):
        dx, dy = float(cxs[d]), float(cys[d])
        if dx == 0.0 and dy == 0.0:
            continue
        # the upwind cell x - c_q (periodic wrap as setNeiSolid) must be solid
        up = np.roll(np.roll(sol, int(cys[d]), axis=1), int(cxs[d]), axis=2)
        m = fluid & up
        a = dx * dx + dy * dy
        b = -(px * dx + py * dy)
        cc = px * px + py * py - R * R
        disc = np.maximum(b * b - a * cc, 0.0)
        t = (-b - np.sqrt(disc)) / a
        out[d][m.reshape(-1)] = np.clip(t, 0.0, 1.0)[m]
    return out.astype(np.float32).reshape(-1)


def _dirs(q):
    c19 = ([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
           [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1],
           [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
    if q == 19:
        return c19
    return (c19[0] + [1, -1, -1, 1, 1, -1, 1, -1], c19[1] + [1, -1, 1, -1, -1, 1, 1, -1],
            c19[2] + [1, -1, 1, -1, 1, -1, -1, 1])


def cylinder_channel(nx=400, ny=160, nz=16, D=40, cx=100, cy=80, u_in=0.05, noise=1e-3, Re=100.0,
                     spz_cells=20, sgs=SGS_WALE, seed=SEED, q=19, sdf_from_geometry=False, ibb=False):
    """Channel flow past a circular cylinder: FixedVelocity inlet (u_in plus seeded uniform
    noise per inlet cell), Outlet (Geier convective) at nxMax with a kernel sponge of
    spz_cells, bounce-back walls in y, periodic z; cylinder cells (i-cx)^2+(j-cy)^2 < (D/2)^2
    are solid.  ibb=True adds analytic q-distances (FP64) and selects the interpolated
    bounce-back on the cylinder."""
    U0 = u_in
    nu_lat = u_in * D / Re
    nu = nu_lat * (1.0 / (nx - 1)) * (1.0 / U0)
    codes = {f: ("Cyclic", "Cyclic", "Cyclic") for f in FACES}
    codes["nxMin"] = ("FixedVelocity", "Cyclic", "Cyclic")
    codes["nxMax"] = ("Outlet", "Cyclic", "Cyclic")
    codes["nyMin"] = ("Cyclic", "BounceBack", "Cyclic")
    codes["nyMax"] = ("Cyclic", "BounceBack", "Cyclic")
    vels = {"nxMin": (1.0, 0.0, 0.0)}
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = (((i - cx) ** 2 + (j - cy) ** 2) < (D / 2.0) ** 2).astype(np.uint8).reshape(-1)
    c = make_case("cylinder", nx, ny, nz, 1.0, 1.0, U0, nu, codes, vels, solid=solid, sgs=sgs,
                  spzWidth=float(spz_cells) / nx, q=q)
    if sdf_from_geometry:
        # signed distance to the circle (negative inside), used by the interpolated bounce-back
        d = (np.sqrt((i - cx) ** 2 + (j - cy) ** 2) - D / 2.0).reshape(-1)
        near = np.abs(d) < 3.0
        c.arrays["sdf"][near] = d[near].astype(np.float32)
    rng = np.random.default_rng(seed)
    inlet = (i.reshape(-1) == 0) & (c.arrays["boundary1"] == BC["FixedVelocity"])
    u0 = c.arrays["u0"]
    u0[inlet] = (u0[inlet].astype(np.float64)
                 + rng.uniform(-noise, noise, size=int(inlet.sum()))).astype(np.float32)
    if ibb:
        # Filippova-Hanel interpolated bounce-back with analytic q-distances
        c.arrays["qlink"] = cylinder_qlinks(q, nx, ny, nz, solid, cx, cy, D / 2.0)
        c.params["ibb"] = 1
    c.meta.update(D=D, cx=cx, cy=cy, u_in=u_in)
    return c


def bluff_body(nx=1024, ny=384, nz=320, D=80.0, centre=None, u_in=0.04, Re=300.0, spz_cells=64, sgs=SGS_WALE,
               seed=SEED, n_lat=250, n_lon=400, amp=0.02, qlinks_on_device=True):
    """A seeded procedural bluff body (about 200k triangles, lbffs_amd.geometry.procedural_body)
    in a D3Q27 wind tunnel: FixedVelocity inlet u_in at nxMin, Outlet + kernel sponge of
    spz_cells at nxMax, Symmetry (free-slip) side faces, cumulant collision with the
    sub-grid model `sgs`.  Solid flags from the surface by the reference's setSolid on the
    GPU, checked for rows whose crossing parity broke (a solid cell outside the body's box,
    or a fluid cell well inside it, raises); interpolated bounce-back with FP64 q-distances
    from the surface, cast into the lattice after the upload (qlinks_on_device) or held as a
    host array (arrays["qlink"]).  Re = u_in D / nu with D the base diameter."""
    from . import geometry as geo
    if centre is None:
        centre = (nx / 4 + 0.3, ny / 2 + 0.4, nz / 2 + 0.2)
    R = D / 2.0
    stl = geo.procedural_body(centre, R, n_lat=n_lat, n_lon=n_lon, amp=amp, seed=seed)
    solid = geo.set_solid(stl, nx, ny, nz)
    # rows whose crossing parity broke would leave solid streaks or fluid holes
    v = np.concatenate([np.asarray(x, np.float64) for x in stl.triangles()], axis=1)
    lo, hi = np.floor(v.min(axis=1)) - 1, np.ceil(v.max(axis=1)) + 1
    s3 = solid.reshape(nz, ny, nx)
    kk, jj, ii = np.nonzero(s3)
    stray = int(((ii < lo[0]) | (ii > hi[0]) | (jj < lo[1]) | (jj > hi[1]) | (kk < lo[2]) | (kk > hi[2])).sum())
    r_in = R * (1.0 - amp) - 1.5
    ci = [np.arange(int(np.floor(centre[a] - r_in)), int(np.ceil(centre[a] + r_in)) + 1) for a in range(3)]
    sub = s3[ci[2][0]:ci[2][-1] + 1, ci[1][0]:ci[1][-1] + 1, ci[0][0]:ci[0][-1] + 1]
    Z, Y, X = np.meshgrid(ci[2] - centre[2], ci[1] - centre[1], ci[0] - centre[0], indexing="ij")
    holes = int(((X * X + Y * Y + Z * Z < r_in * r_in) & (sub == 0)).sum())
    if stray or holes:
        raise RuntimeError(f"solid flags of the procedural body: {stray} solid cells outside its box, "
                           f"{holes} fluid cells inside it (broken crossing parity)")
    U0 = u_in
    nu_lat = u_in * D / Re
    nu = nu_lat * (1.0 / (nx - 1)) * (1.0 / U0)
    codes = {f: ("Cyclic", "Cyclic", "Cyclic") for f in FACES}
    codes["nxMin"] = ("FixedVelocity", "Cyclic", "Cyclic")
    codes["nxMax"] = ("Outlet", "Cyclic", "Cyclic")
    for f in ("nyMin", "nyMax"):
        codes[f] = ("Cyclic", "Symmetry", "Cyclic")
    for f in ("nzMin", "nzMax"):
        codes[f] = ("Cyclic", "Cyclic", "Symmetry")
    vels = {"nxMin": (1.0, 0.0, 0.0)}
    # the fluid starts at the inlet velocity (physical 1 -> lattice U0, as the inlet's)
    c = make_case("bluff", nx, ny, nz, 1.0, 1.0, U0, nu, codes, vels, ini=(1.0, 0.0, 0.0), solid=solid, sgs=sgs,
                  spzWidth=float(spz_cells) / nx, q=27)
    c.params["ibb"] = 1
    if qlinks_on_device:
        c.surface = stl
    else:
        ql, miss = geo.qlinks(stl, 27, nx, ny, nz, c.arrays["solid"], c.arrays["neiSolid"])
        c.arrays["qlink"] = ql
        c.meta["qlink_misses"] = miss
    c.meta.update(D=D, R=R, centre=tuple(centre), u_in=u_in, Re=Re, triangles=stl.patches[0].n,
                  frontal_area=geo.frontal_area(stl), stl=stl)
    return c


def backward_step(nx=1536, ny=192, nz=128, H=64, step_len=None, u_in=0.05, turb=0.05, Re=5500.0,
                  spz_cells=64, sgs=SGS_SMAGORINSKY, Cs=0.1, seed=SEED):
    """Backward-facing step: solid block i < step_len, j < H spanning z; FixedVelocity
    inlet above the step with seeded 5 % Gaussian turbulence on u0; Outlet + sponge at
    nxMax; bounce-back y walls; periodic z; Smagorinsky LES with wall damping (wall
    distance from the step and the y walls).  Solid links of face cells are handled by
    the reference rule (setNeiSolid leaves faces out)."""
    step_len = 3 * H if step_len is None else step_len
    U0 = u_in
    nu_lat = u_in * H / Re
    nu = nu_lat * (1.0 / (nx - 1)) * (1.0 / U0)
    codes = {f: ("Cyclic", "Cyclic", "Cyclic") for f in FACES}
    codes["nxMin"] = ("FixedVelocity", "Cyclic", "Cyclic")
    codes["nxMax"] = ("Outlet", "Cyclic", "Cyclic")
    codes["nyMin"] = ("Cyclic", "BounceBack", "Cyclic")
    codes["nyMax"] = ("Cyclic", "BounceBack", "Cyclic")
    vels = {"nxMin": (1.0, 0.0, 0.0)}
    N = nx * ny * nz
    ic = np.arange(N, dtype=np.int64)
    i = ic % nx
    j = (ic // nx) % ny
    solid = ((i < step_len) & (j < H)).astype(np.uint8)
    c = make_case("backstep", nx, ny, nz, 1.0, 1.0, U0, nu, codes, vels, solid=solid, sgs=sgs, LES=1.0, Cs=Cs,
                  damping=1, spzWidth=float(spz_cells) / nx)
    rng = np.random.default_rng(seed)
    inlet = (i == 0) & (c.arrays["boundary1"] == BC["FixedVelocity"]) & (solid == 0)
    u0 = c.arrays["u0"]
    u0[inlet] = (u0[inlet].astype(np.float64) * (1.0 + turb * rng.standard_normal(int(inlet.sum())))).astype(np.float32)
    # wall distance (half-way walls) for cells within 3 cells of the y walls or the step
    yb = j.astype(np.float64) + 0.5                     # to the nyMin wall
    yt = (ny - 1 - j).astype(np.float64) + 0.5          # to the nyMax wall
    ys = np.where(i >= step_len, np.hypot(np.maximum(step_len - 0.5 - i, 0), np.maximum(j - (H - 0.5), 0)),
                  (j - (H - 0.5)).astype(np.float64))   # to the step surfaces
    ys = np.where((i >= step_len) & (j >= H), np.hypot(i - (step_len - 0.5), j - (H - 0.5)), ys)
    ys = np.where((i >= step_len) & (j < H), (i - (step_len - 0.5)).astype(np.float64), ys)
    d = np.minimum(np.minimum(yb, yt), np.abs(ys))
    sdf = c.arrays["sdf"]
    near = (d < 3.0) & (solid == 0)
    sdf[near] = d[near].astype(np.float32)
    c.meta.update(H=H, step_len=step_len, u_in=u_in)
    return c


def ensemble_specs(n_domains=512, seed=SEED, edges=(24, 96), re_range=(50.0, 400.0)):
    """(kind, edge, Re) of the independent cubic domains of a batch: lid-driven cavities and
    cylinder channels in turn, edge uniform in [edges[0], edges[1]], Reynolds number
    uniform in re_range, from a seeded generator."""
    rng = np.random.default_rng(seed)
    out = []
    for d in range(n_domains):
        n = int(rng.integers(edges[0], edges[1] + 1))
        Re = float(rng.uniform(*re_range))
        out.append(("cavity" if d % 2 == 0 else "cylinder", n, Re))
    return out


def ensemble_case(spec, index=0, seed=SEED, sgs=SGS_WALE, q=19):
    kind, n, Re = spec
    if kind == "cavity":
        c = cavity(n, Re=Re, sgs=sgs, q=q)
    else:
        c = cylinder_channel(nx=n, ny=n, nz=n, D=max(4, n // 5), cx=n // 4, cy=n // 2, Re=Re,
                             spz_cells=max(2, n // 10), sgs=sgs, seed=seed + index, q=q)
    c.meta.update(edge=n, Re=Re)
    return c


def ensemble_omega0(spec):
    """omega0 of ensemble_case(spec) without building its arrays (the builders' units)."""
    kind, n, Re = spec
    if kind == "cavity":
        U0 = 0.1
        nu = (U0 * n / Re) * (1.0 / (n - 1)) * (1.0 / U0)
    else:
        U0 = 0.05
        nu = (U0 * max(4, n // 5) / Re) * (1.0 / (n - 1)) * (1.0 / U0)
    _, un = units(n, n, n, 1.0, 1.0, U0, nu)
    return float(un.omega0)


def ensemble(n_domains=512, seed=SEED, edges=(24, 96), re_range=(50.0, 400.0), sgs=SGS_WALE, q=19):
    """The domains of ensemble_specs, built."""
    return [ensemble_case(sp, d, seed, sgs, q) for d, sp in enumerate(ensemble_specs(n_domains, seed, edges, re_range))]
