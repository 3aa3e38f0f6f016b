"""The wavefront-cell kernels keep the whole cell in registers: compiled for
gfx950 at two waves per SIMD they must use no scratch memory (a conditional
over the cell's member registers, or a dynamically indexed local array, makes
the compiler place the cell in per-lane scratch memory, several times
slower).  Compiles voro_amd/csrc/cells_wave.hip with the build's flags (the
128-vertex tier, which runs at four waves per SIMD, at two for this check)
and reads the compiler's resource report."""
import os
import re
import subprocess

import pytest

from voro_amd import build as B


def _resources(src, extra=()):
    cmd = [B.hipcc()] + B.FLAGS + list(extra) + ["-c", src, "-o", os.devnull, "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out, cur = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            out[cur] = {}
            continue
        m = re.search(r"remark:\s+([A-Za-z ]+?)(?: \[[^\]]*\])?: (\d+)", line)
        if m and cur:
            out[cur][m.group(1).strip()] = int(m.group(2))
    return out


def test_wave_cell_kernels_use_no_scratch_at_two_waves_per_simd():
    try:
        B.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")
    res = _resources(os.path.join(B.CSRC, "cells_wave.hip"), ["-DVG_K2_OCC=2"])
    # k_wcells<K, POLY, OCC, PRD>: the instantiations at OCC = 2
    two = {k: v for k, v in res.items() if re.search(r"k_wcellsILi[124]ELb[01]ELi2ELb[01]E", k)}
    assert len(two) >= 10, sorted(res)
    for k, v in two.items():
        assert "ScratchSize" in v and v["ScratchSize"] == 0, (k, v)
