"""Block worklist generator.

Produces the 64 worklists (one per sub-region of the 4x4x4 lower octant of an
8x8x8 subdivision of a computational block) that order the blocks visited
while a Voronoi cell is built, together with the per-entry "outside
neighbour" flags.  The algorithm follows the reference generator
``src/worklist_gen.pl:15-236`` (parameters ``hr=4``, ``ls=63``; greedy
nearest-block expansion weighted by ``adis + 0.02*|step|``; mark/last-access
bookkeeping for blocks outside the list) so that the resulting table is the
same data as ``src/v_base_wl.cc:14-78``.  Entry encoding (as in the reference):
``(di+64) | (dj+64)<<7 | (dk+64)<<14 | flags<<21`` where the six flag bits say
which face neighbours of the block lie outside the worklist and must be queued
when the block is visited; element 0 of each list is the number of leading
entries whose neighbours are all on the list.

Run as a script to (re)write ``csrc/worklist_table.h``.
"""
import math
import os

HR = 4                 # half the number of sub-regions per block side
FG = 2 * HR            # sub-regions per block side
LS = 63                # entries per worklist (plus the leading count)
DIS = 8                # mask displacement
D = 2 * DIS + 1
DD = D * D
D0 = (1 + D + DD) * DIS


def _adis(x, y, z, xt, yt, zt):
    # Distance from a sub-region centre to block (xt,yt,zt), worklist_gen.pl:216-225
    xco = yco = zco = 0.0
    if xt > 0:
        xco = x - xt
    if xt < 0:
        xco = x - xt - 1
    if yt > 0:
        yco = y - yt
    if yt < 0:
        yco = y - yt - 1
    if zt > 0:
        zco = z - zt
    if zt < 0:
        zco = z - zt - 1
    return math.sqrt(xco * xco + yco * yco + zco * zco)


def _one_list(ii, jj, kk, state):
    """One worklist for sub-region (ii,jj,kk); worklist_gen.pl:108-196."""
    m = state["m"]
    la = state["la"]
    state["v"] += 2
    v = state["v"]
    xp = yp = zp = 0
    x = (ii + 0.5) / FG
    y = (jj + 0.5) / FG
    z = (kk + 0.5) / FG
    cand = []          # frontier blocks (a list in the reference)
    chosen = []        # blocks on the worklist, in order

    def add(a, b, c):
        k = D0 + a + D * b + DD * c
        if m.get(k) != v:
            cand.append((a, b, c))
            m[k] = v

    m[D0] = v
    add(1, 0, 0); add(0, 1, 0); add(0, 0, 1)
    add(-1, 0, 0); add(0, -1, 0); add(0, 0, -1)
    for _ in range(LS):
        minwei = 1e9
        best = 0
        for idx, (xt, yt, zt) in enumerate(cand):
            wei = _adis(x, y, z, xt, yt, zt) + 0.02 * math.sqrt(
                float((xt - xp) ** 2 + (yt - yp) ** 2 + (zt - zp) ** 2))
            if wei < minwei:
                minwei = wei
                best = idx
        xp, yp, zp = cand[best]
        add(xp + 1, yp, zp); add(xp, yp + 1, zp); add(xp, yp, zp + 1)
        add(xp - 1, yp, zp); add(xp, yp - 1, zp); add(xp, yp, zp - 1)
        chosen.append(cand.pop(best))

    # Mark every block on this worklist.
    v += 1
    state["v"] = v
    m[D0] = v
    for (xt, yt, zt) in chosen:
        m[D0 + xt + D * yt + DD * zt] = v

    # Last worklist position from which each outside neighbour is reachable.
    for j, (xt, yt, zt) in enumerate(chosen):
        k = D0 + xt + yt * D + zt * DD
        for off, cond in ((1, xt >= 0), (D, yt >= 0), (DD, zt >= 0),
                          (-1, xt <= 0), (-D, yt <= 0), (-DD, zt <= 0)):
            if cond and m.get(k + off) != v:
                la[k + off] = j
                m[k + off] = v + 1

    for (xt, yt, zt) in chosen:
        k = D0 + xt + yt * D + zt * DD
        for off in (1, -1, D, -D, DD, -DD):
            if m.get(k + off, 0) < v:
                raise RuntimeError("worklist construction failed")

    # Number of leading entries whose six neighbours are all on the list.
    nlead = 0
    for (xt, yt, zt) in chosen:
        k = D0 + xt + yt * D + zt * DD
        if any(m.get(k + off) != v for off in (1, D, DD, -1, -D, -DD)):
            break
        nlead += 1

    out = [nlead]
    for j, (xt, yt, zt) in enumerate(chosen):
        k = D0 + xt + yt * D + zt * DD
        o = 0
        if m.get(k + 1) != v and la.get(k + 1) == j:
            o |= 1
        if m.get(k - 1) != v and la.get(k - 1) == j:
            o ^= 3
        if m.get(k + D) != v and la.get(k + D) == j:
            o |= 8
        if m.get(k - D) != v and la.get(k - D) == j:
            o ^= 24
        if m.get(k + DD) != v and la.get(k + DD) == j:
            o |= 64
        if m.get(k - DD) != v and la.get(k - DD) == j:
            o ^= 192
        out.append((xt + 64) | ((yt + 64) << 7) | ((zt + 64) << 14) | (o << 21))
    return out


def generate():
    """Returns the 64*64 worklist table as a flat list of ints."""
    state = {"m": {}, "la": {}, "v": 0}
    table = []
    for kk in range(HR):
        for jj in range(HR):
            for ii in range(HR):
                table.extend(_one_list(ii, jj, kk, state))
    assert len(table) == 64 * (LS + 1)
    return table


def header_text(table):
    lines = [
        "// Generated by voro_amd/worklist_gen.py -- do not edit.",
        "// 64 block worklists x 64 entries; see the generator for the encoding.",
        "#pragma once",
        "#define VGPU_WL_HGRID %d" % HR,
        "#define VGPU_WL_FGRID %d" % FG,
        "#define VGPU_WL_SEQ %d" % (LS + 1),
        "static const unsigned int vgpu_worklist_table[%d] = {" % len(table),
    ]
    for w in range(64):
        row = table[w * 64:(w + 1) * 64]
        lines.append("  " + ",".join("0x%x" % t for t in row) + ",")
    lines.append("};")
    return "\n".join(lines) + "\n"


HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "worklist_table.h")


def main():
    table = generate()
    with open(HEADER, "w") as f:
        f.write(header_text(table))
    print("wrote", HEADER)


if __name__ == "__main__":
    main()
