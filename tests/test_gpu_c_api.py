"""The boundary from C: tests/c_api/boundary.c, compiled with the system C
compiler against include/ and linked to libopenvkl_hip.so, runs the object
API, vklComputeSampleN / vklComputeGradientN on host pointers and the
per-ray interval and hit iterators; its output must equal, bit for bit, the
oracle's results committed in tests/golden/c_boundary_expected.bin
(tests/golden/make_c_boundary.py)."""
import os
import subprocess

import numpy as np
import pytest

import openvkl_amd as vkl

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "c_api", "boundary.c")
EXE = os.path.join(HERE, "c_api", "_build", "boundary")
EXPECTED = os.path.join(HERE, "golden", "c_boundary_expected.bin")
NPTS, NGRAD, NRAYS = 4096, 256, 16


def build_boundary_program():
    """cc -std=c99 boundary.c -I include -L openvkl_amd/lib -lopenvkl_hip
    (rebuilt when the source or the library is newer than the program)."""
    lib_dir = os.path.dirname(vkl.LIB_PATH)
    if os.path.exists(EXE) and os.path.getmtime(EXE) >= max(os.path.getmtime(SRC), os.path.getmtime(vkl.LIB_PATH)):
        return EXE
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    cmd = ["cc", "-std=c99", "-O2", "-ffp-contract=off", "-Wall", SRC, "-I", os.path.join(ROOT, "include"),
           "-L", lib_dir, "-lopenvkl_hip", "-lm", f"-Wl,-rpath,{lib_dir}", "-o", EXE]
    subprocess.run(cmd, check=True, capture_output=True)
    return EXE


def parse(buf):
    off = 0

    def take(n, shape):
        nonlocal off
        a = np.frombuffer(buf, np.float32, n, off).reshape(shape)
        off += 4 * n
        return a

    out = {"dense": take(NPTS, (NPTS,)), "grad": take(3 * NGRAD, (NGRAD, 3)), "vdb": take(NPTS, (NPTS,))}
    for key, width in (("intervals", 5), ("hits", 3)):
        rays = []
        for _ in range(NRAYS):
            n = int(np.frombuffer(buf, np.uint32, 1, off)[0])
            off += 4
            rays.append(take(n * width, (n, width)))
        out[key] = rays
    assert off == len(buf)
    return out


def same(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


def test_c_program_matches_oracle_fixture(tmp_path):
    exe = build_boundary_program()
    out = tmp_path / "boundary.bin"
    print("running the C boundary program", flush=True)
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got, want = parse(out.read_bytes()), parse(open(EXPECTED, "rb").read())
    for key in ("dense", "grad", "vdb"):
        assert same(got[key], want[key]), key
    for key in ("intervals", "hits"):
        for i, (g, w) in enumerate(zip(got[key], want[key])):
            assert same(g, w), (key, i)
    assert sum(len(h) for h in want["hits"]) > 0 and sum(len(v) for v in want["intervals"]) > 0
