"""Builds the HIP engine library voro_amd/_lib/libvoro_gpu.so for gfx950.

Each kernel tier is its own translation unit; they compile in parallel and
are linked into one shared library exporting the C ABI of
include/voro_gpu.h.  Device code is compiled with -ffp-contract=off so that
no multiply-add is fused (the cell arithmetic must round exactly as the
reference's does) and without packed-FP32 instructions.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
OBJDIR = os.path.join(HERE, "_lib", "obj")
LIB = os.path.join(LIBDIR, "libvoro_gpu.so")
ARCH = os.environ.get("VGPU_ARCH", "gfx950")

UNITS = ["vgpu_api.hip", "cells_fast3.hip", "cells_wide.hip", "cells_huge.hip", "cells_prd.hip", "cells_list.hip",
         "cells_batch.hip", "cells_wave.hip", "find_cell.hip", "learn.hip", "geometry.hip", "text_io.hip"]
HEADERS = ["vgpu_common.h", "cell_dev.h", "compute_dev.h", "vgpu_cells.h", "vgpu_bin.h", "grid_host.h",
           "worklist_table.h", "cell3_dev.h", "list_dev.h", "learn.h", "prd_host.h", "batch_dev.h", "wcell_dev.h", "geom_dev.h",
           "wave_cell_dev.h"]

FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fno-unroll-loops", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-I" + os.path.join(HERE, "..", "include")]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _stale(obj, src):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, h) for h in HEADERS if os.path.exists(os.path.join(CSRC, h))]
    deps.append(os.path.join(HERE, "..", "include", "voro_gpu.h"))
    return any(os.path.getmtime(d) > t for d in deps)


def build(force=True, verbose=False, jobs=None):
    """Compiles every unit (all of them when force) and links the library."""
    os.makedirs(OBJDIR, exist_ok=True)
    cc = hipcc()
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u))]

    def one(u):
        src = os.path.join(CSRC, u)
        obj = os.path.join(OBJDIR, u.replace(".hip", ".o"))
        if not force and not _stale(obj, src):
            return None
        cmd = [cc] + FLAGS + ["-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32" not in l)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed for %s:\n%s" % (u, msg))
        if verbose and msg.strip():
            print(msg)
        return obj

    with ThreadPoolExecutor(max_workers=jobs or min(8, len(units))) as ex:
        list(ex.map(one, units))
    objs = [os.path.join(OBJDIR, u.replace(".hip", ".o")) for u in units]
    tmp = LIB + ".tmp"
    r = subprocess.run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    import time
    t = time.time()
    build(force="--force" in sys.argv, verbose=True)
    print("built", LIB, "in %.1fs" % (time.time() - t))
