"""The C++ drop-in header: a program written against the reference's
interface compiles and links against libvoro_gpu.so (CPU), and on the GPU
prints the same cells / volume sum / faces / neighbour checksum as the
oracle gives for the same particles."""
import os
import subprocess

import numpy as np
import pytest

import voro_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "loop_example.cc")
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "loop_example")


def build_example():
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    libdir = os.path.dirname(voro_amd.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"), SRC, "-L" + libdir,
                    "-lvoro_gpu", "-Wl,-rpath," + libdir, "-o", EXE], check=True)


def test_dropin_header_compiles_and_links():
    build_example()
    assert os.path.exists(EXE)


def _example_particles():
    g = np.uint64(0x9E3779B97F4A7C15)
    m1, m2 = np.uint64(0xBF58476D1CE4E5B9), np.uint64(0x94D049BB133111EB)

    def mix(z):
        with np.errstate(over="ignore"):
            z = (z ^ (z >> np.uint64(30))) * m1
            z = (z ^ (z >> np.uint64(27))) * m2
        return z ^ (z >> np.uint64(31))

    s0 = mix(np.array([12345], dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        k = np.arange(1, 3001, dtype=np.uint64)
        u = mix(s0 + k * g)
    return ((u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)).reshape(1000, 3)


@pytest.mark.gpu
def test_dropin_program_matches_oracle(oracle_mod):
    O = oracle_mod
    build_example()
    out = subprocess.run([EXE], check=True, capture_output=True, text=True).stdout.split()
    xyz = _example_particles()
    ref = O.run("restatement", O.make_spec(0, (0, 1, 0, 1, 0, 1), (6, 6, 6)), np.arange(1000, dtype=np.int32), xyz)
    cells = int(ref["exists"].sum())
    vsum = 0.0
    for v, e in zip(ref["vol"], ref["exists"]):
        if e:
            vsum += float(v)
    chk = 0
    for t in ref["nbr"]:
        chk = (chk * 31 + (int(t) + 7)) & ((1 << 64) - 1)
    assert int(out[0]) == cells
    assert float(out[1]) == vsum
    assert int(out[2]) == int(ref["nfaces"].sum())
    assert int(out[3]) == chk
