"""The C-ABI library: it loads without a GPU, exports every entry point that
include/*.h declares, and reports errors as codes (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

import lbffs_amd as lb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(?:int|const char \*)\s*(lbffs_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = lb.lib()
    names = declared()
    assert len(names) >= 15
    for n in sorted(names):
        assert hasattr(L, n), n


def test_strerror_and_invalid_arguments():
    L = lb.lib()
    assert L.lbffs_strerror(0) == b"ok"
    assert L.lbffs_strerror(-1) == b"invalid argument"
    assert L.lbffs_create(None, None) == -1
    assert b"null" in L.lbffs_last_error()
    d = lb.Desc(abi_version=99, q=19, nx=4, ny=4, nz=4)
    h = ctypes.c_void_p()
    assert L.lbffs_create(ctypes.byref(d), ctypes.byref(h)) == -1
    d = lb.Desc(abi_version=lb.ABI_VERSION, q=19, nx=1, ny=4, nz=4)
    assert L.lbffs_create(ctypes.byref(d), ctypes.byref(h)) == -1
    d = lb.Desc(abi_version=lb.ABI_VERSION, q=19, nx=4, ny=4, nz=4, storage=9)
    assert L.lbffs_create(ctypes.byref(d), ctypes.byref(h)) == -1
    # lattice/collision pairs: D3Q19 runs the RR collision (K19:3637); D3Q27 the cumulant
    # (K27:2824) or the RR collision on the D3Q27 Hermite set; D3Q19 has no cumulant
    for q, coll in ((19, lb.COLL_CUMULANT), (21, lb.COLL_RR), (27, 7)):
        d = lb.Desc(abi_version=lb.ABI_VERSION, q=q, nx=4, ny=4, nz=4, collision=coll)
        assert L.lbffs_create(ctypes.byref(d), ctypes.byref(h)) == -4
    assert L.lbffs_ibm_set_markers(None, None, 0) == -1
    assert L.lbffs_ibm_set_motion(None, None) == -1
    assert L.lbffs_ibm_download(None, 0, None, 0) == -1
    # a batch shares the lattice type (validated before any device call)
    descs = (lb.Desc * 2)(lb.Desc(abi_version=lb.ABI_VERSION, q=19, nx=8, ny=8, nz=8, omega0=1.0),
                          lb.Desc(abi_version=lb.ABI_VERSION, q=27, nx=8, ny=8, nz=8, omega0=1.0,
                                  collision=lb.COLL_CUMULANT))
    assert L.lbffs_batch_create(descs, 2, ctypes.byref(h)) == -1
    assert b"share" in L.lbffs_last_error()
    assert L.lbffs_batch_create(descs, 0, ctypes.byref(h)) == -1
    assert L.lbffs_batch_step(None, 1, None) == -1
    assert L.lbffs_batch_destroy(None) == 0
    assert L.lbffs_upload(None, 0, None, 0) == -1
    # a population array addressed with 32-bit byte offsets: 1100^3 FP32 cells are refused
    d = lb.Desc(abi_version=lb.ABI_VERSION, q=19, nx=1100, ny=1100, nz=1100, storage=lb.STORE_F32)
    assert L.lbffs_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"4 GiB" in L.lbffs_last_error()
    # and so is a macroscopic (FP32) array: the same lattice in FP16 storage
    d = lb.Desc(abi_version=lb.ABI_VERSION, q=19, nx=1100, ny=1100, nz=1100, storage=lb.STORE_F16DEV)
    assert L.lbffs_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"4 GiB" in L.lbffs_last_error()
    # geometry: argument checks come before any device call
    assert L.lbffs_geom_solid(1, None, None, None, 4, 4, 4, 0, None) == -1
    assert L.lbffs_geom_sdf(1, None, None, 1.0, 1.0, 7.0, 4, 4, 4, None, None) == -1
    buf = (ctypes.c_uint8 * 8)()
    assert L.lbffs_geom_solid(0, None, None, None, 0, 2, 2, 0, buf) == -1
    assert L.lbffs_geom_solid(-1, None, None, None, 2, 2, 2, 0, buf) == -1
    assert L.lbffs_step(None, 1, None) == -1
    assert L.lbffs_destroy(None) == 0


def test_missing_library_raises(monkeypatch):
    monkeypatch.setattr(lb, "_lib", None)
    monkeypatch.setattr(lb, "LIB_PATH", "/nonexistent/liblbffs.so")
    with pytest.raises(lb.LbffsError):
        lb.lib()


def test_streaming_descriptor_checks():
    # in-place streaming: validated at create before any device call
    L = lb.lib()
    h = ctypes.c_void_p()
    d = lb.Desc(abi_version=lb.ABI_VERSION, q=19, nx=8, ny=8, nz=8, omega0=1.0, streaming=7)
    assert L.lbffs_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"streaming" in L.lbffs_last_error()
    d = lb.Desc(abi_version=lb.ABI_VERSION, q=27, nx=8, ny=8, nz=8, omega0=1.0, collision=lb.COLL_CUMULANT,
                streaming=lb.STREAM_AA)
    assert L.lbffs_create(ctypes.byref(d), ctypes.byref(h)) == -4
    descs = (lb.Desc * 1)(lb.Desc(abi_version=lb.ABI_VERSION, q=19, nx=8, ny=8, nz=8, omega0=1.0,
                                  streaming=lb.STREAM_AA))
    assert L.lbffs_batch_create(descs, 1, ctypes.byref(h)) == -4


def test_geometry_and_patch_argument_checks():
    # the context q-distance cast and findPoints refuse bad arguments before any device call
    L = lb.lib()
    miss = ctypes.c_int()
    assert L.lbffs_geom_qlinks_ctx(None, 0, None, None, None, ctypes.c_float(0.5), ctypes.byref(miss)) == -1
    cnt = ctypes.c_long()
    assert L.lbffs_case_find_points(-1, None, None, None, None, None, 4, 4, 4, None, 0, ctypes.byref(cnt)) != 0
    assert L.lbffs_case_find_points(1, None, None, None, None, None, 4, 4, 4, None, 0, ctypes.byref(cnt)) != 0
    assert L.lbffs_case_find_points(0, None, None, None, None, None, 4, 4, 4, None, 0, ctypes.byref(cnt)) == 0
    assert cnt.value == 0
    # launch options: a null context is refused before the options are read
    assert L.lbffs_set_launch(None, 256, 8) == -1


def test_headers_compile_as_c_and_link(tmp_path):
    # the drop-in boundary is plain C: every header compiles as C99 (gcc), and a C host that
    # calls one entry point of each header links against the library without HIP headers
    inc = os.path.join(ROOT, "include")
    src = tmp_path / "host.c"
    src.write_text(
        '#include "lbffs.h"\n#include "lbffs_case.h"\n#include "lbffs_geom.h"\n#include "lbffs_output.h"\n'
        "#include <stdio.h>\n"
        "int main(void) {\n"
        "  lbffs_ctx *c = 0; lbffs_desc d = {0}; long n = 0;\n"
        "  int e = lbffs_create(0, &c);\n"
        "  int f = lbffs_case_find_points(0, 0, 0, 0, 0, 0, 4, 4, 4, 0, 0, &n);\n"
        "  int g = lbffs_geom_qlinks_ctx(0, 0, 0, 0, 0, 0.5f, 0);\n"
        "  int h = lbffs_out_write_vtk(0, 0, LBFFS_VTK_ASCII);\n"
        "  (void)d; printf(\"%d %d %d %d %s\\n\", e, f, g, h, lbffs_strerror(e));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "host"
    lib_dir = os.path.join(ROOT, "lbffs_amd")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, str(src), "-L", lib_dir, "-llbffs",
                        "-Wl,-rpath," + lib_dir, "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    e, f, g, h = (int(x) for x in out.stdout.split()[:4])
    assert e == -1 and f == 0 and g == -1 and h != 0
