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


FULL_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_full.cc")
FULL_EXE = os.path.join(ROOT, "tests", "cpp", "_build", "dropin_full")


def build_full():
    os.makedirs(os.path.dirname(FULL_EXE), exist_ok=True)
    libdir = os.path.dirname(voro_amd.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), FULL_SRC,
                    "-L" + libdir, "-lvoro_gpu", "-Wl,-rpath," + libdir, "-o", FULL_EXE], check=True)


def test_full_dropin_compiles_and_links():
    build_full()
    assert os.path.exists(FULL_EXE)


def _mix_stream(seed, count, scale, shift=0.0):
    g = np.uint64(0x9E3779B97F4A7C15)
    m1, m2 = np.uint64(0xBF58476D1CE4E5B9), np.uint64(0x94D049BB133111EB)

    def mix(z):
        with np.errstate(over="ignore"):
            z = (z ^ (z >> np.uint64(30))) * m1
            z = (z ^ (z >> np.uint64(27))) * m2
        return z ^ (z >> np.uint64(31))

    s0 = mix(np.array([seed], dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        k = np.arange(1, 3 * count + 1, dtype=np.uint64)
        u = mix(s0 + k * g)
    return (scale * ((u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - shift).reshape(count, 3)


@pytest.mark.gpu
def test_full_dropin_program_matches_reference(oracle_mod, tmp_path):
    """Walls of all four kinds, particle_order, the cell outputs and
    polyhedron statistics over c_loop_all, c_loop_subset, c_loop_order,
    print_custom, draw_cells_gnuplot and pre_container, driven through the
    C++ header, against the reference library."""
    O = oracle_mod
    if O.reference() is None:
        pytest.skip("reference library not built")
    build_full()
    out = subprocess.run([FULL_EXE, str(tmp_path)], check=True, capture_output=True, text=True).stdout.splitlines()
    lines = {l.split()[0]: l.split()[1:] for l in out}
    xyz = _mix_stream(777, 3000, 10.0)
    ids = np.arange(3000, dtype=np.int32)
    walls = [(0, -5, 5, 5, 5, 4.9), (1, -6, 0.2, 0.1, 1, 8.7), (2, -7, 5, 5, 5, 1, 0.3, 0.2, 4.4),
             (3, -8, 5, 5, -1, 0.1, 0.2, 1, 0.55)]
    spec = O.make_spec(0, (0, 10, 0, 10, 0, 10), (6, 6, 6), walls=walls)
    ref = O.run("reference", spec, ids, xyz)
    geo = O.geometry_reference(spec, ids, xyz)
    e = ref["exists"].astype(bool)
    vsum = area = area_f = edges = 0.0
    fa = geo["face_area"]
    for m in np.nonzero(e)[0]:
        vsum += float(ref["vol"][m])
        area += float(geo["surface"][m])
        for t in range(geo["face_off"][m], geo["face_off"][m + 1]):
            area_f += float(fa[t])
        edges += float(geo["edge_dist"][m])
    chk = 0
    for t in ref["nbr"]:
        chk = (chk * 31 + (int(t) + 101)) & ((1 << 64) - 1)
    a = lines["all"]
    assert int(a[0]) == int(e.sum()) and float(a[1]) == vsum and int(a[2]) == int(ref["nfaces"].sum())
    assert int(a[3]) == chk and float(a[4]) == area and float(a[5]) == area_f
    assert int(a[6]) == int(geo["face_verts"].sum()) and int(a[7]) == int(geo["orders"].sum()) and float(a[8]) == edges
    sub = O.loop_subset_reference(spec, ids, xyz, None, 0, (4.0, 6.0, 5.0, 2.2), None, True)
    assert [int(t) for t in lines["subset"]] == [len(sub), int(sub[:, 0].sum()),
                                                 int((sub[:, 1] + 7 * sub[:, 2] + 49 * sub[:, 3]).sum())]
    ordered = [i for i in range(3000) if i % 3 == 0]
    c = 0
    for i in ordered:
        c = (c * 131 + i) & ((1 << 64) - 1)
    assert [int(t) for t in lines["order"]] == [len(ordered), c]
    want = tmp_path / "custom_ref.txt"
    assert O.print_custom_reference(spec, ids, xyz, None, "%i %q %v %s %n %f %t %l %P %A", str(want))
    assert (tmp_path / "custom.txt").read_text() == want.read_text()
    want = tmp_path / "cells_ref.gnu"
    assert O.draw_reference(spec, ids, xyz, None, "cells_gnuplot", str(want))
    assert (tmp_path / "cells.gnu").read_text() == want.read_text()
    pre = _mix_stream(4242, 1000, 12.0, 1.0)
    inside = np.all((pre >= 0) & (pre <= 10), axis=1)
    p = lines["pre"]
    n = int(inside.sum())
    s = (n / (5.6 * 1000)) ** (1 / 3.0)
    assert [int(t) for t in p[:4]] == [n, int(10 * s + 1), int(10 * s + 1), int(10 * s + 1)]
    g = int(p[1])
    r2 = O.run("reference", O.make_spec(0, (0, 10, 0, 10, 0, 10), (g, g, g)), np.arange(1000, dtype=np.int32)[inside],
               pre[inside])
    assert float(p[4]) == float(sum(float(v) for v in r2["vol"][r2["exists"].astype(bool)]))
