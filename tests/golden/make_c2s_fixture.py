"""Write tests/golden/libcint_cart2sph.json: the numeric values of libcint's
Cartesian -> real-spherical table g_trans_cart2sph (src/cart2sph.c:21-806,
offsets at :3562-3568) for l = 0..6, read from the reference source as data.

The default build of libcint orders p as px, py, pz (the #else branch of
PYPZPX, src/cart2sph.c:23-51); that branch is the one taken here.  Rows are
m = -l..l, columns the Cartesian components in libcint order (lx descending,
then ly descending).  Run in a checkout that has the reference at
/root/reference:

    python tests/golden/make_c2s_fixture.py
"""
import json
import os
import re

REF = "/root/reference/src/cart2sph.c"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcint_cart2sph.json")
LMAX = 6


def table_numbers(path):
    with open(path) as fh:
        lines = fh.read().split("\n")
    start = next(i for i, s in enumerate(lines) if "g_trans_cart2sph[] =" in s)
    vals = []
    skip = False
    for s in lines[start + 1:]:
        t = s.strip()
        if t.startswith("};"):
            break
        if t.startswith("#ifdef PYPZPX"):
            skip = True
            continue
        if t.startswith("#else"):
            skip = False
            continue
        if t.startswith("#endif"):
            skip = False
            continue
        if skip:
            continue
        t = re.sub(r"/\*.*?\*/", "", t)
        t = t.split("//")[0]
        for tok in t.split(","):
            tok = tok.strip()
            if tok:
                vals.append(float(tok))
    return vals


def main():
    v = table_numbers(REF)
    out = {"source": "src/cart2sph.c:21-806 (g_trans_cart2sph), p order px,py,pz", "l": {}}
    off = 0
    for l in range(LMAX + 1):
        nc = (l + 1) * (l + 2) // 2
        ns = 1 if l == 0 else (3 if l == 1 else 2 * l + 1)
        n = nc * ns if l > 0 else 1
        blk = v[off:off + n]
        out["l"][str(l)] = [blk[r * nc:(r + 1) * nc] for r in range(ns)] if l > 0 else [[blk[0]]]
        off += n
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", OUT, "entries", off)


if __name__ == "__main__":
    main()
