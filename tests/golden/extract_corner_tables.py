"""Writes corner_tables.json: the region/offset tables of the reference's
corner-fill stencils, read as text from fv3core/utils/corners.py.

Each statement ``with horizontal(region[A +/- a, B +/- b], ...): dst =
[mult *] src[di, dj, 0]`` becomes one entry per region:
``[dst_name, anchor_i, a, anchor_j, b, src_name, di, dj, mult_name]``, in
statement order.  The oracle builds the same tables from rotation rules and
tests/test_corner_tables.py checks them against this file.

Run from the repository root:  python tests/golden/extract_corner_tables.py
"""
import json
import pathlib
import re
import sys

REF = pathlib.Path("/root/reference/fv3core/utils/corners.py")
FUNCS = (
    "fill_corners_2cells_mult_x", "fill_corners_3cells_mult_x",
    "fill_corners_2cells_mult_y", "fill_corners_3cells_mult_y",
    "fill_corners_bgrid_x_defn", "fill_corners_bgrid_y_defn", "fill_corners_dgrid_defn",
)


def _anchor(expr):
    m = re.fullmatch(r"\s*([ij]_(?:start|end))\s*(?:([+-])\s*(\d+))?\s*", expr)
    if not m:
        raise ValueError(expr)
    off = int(m.group(3) or 0) * (-1 if m.group(2) == "-" else 1)
    return m.group(1), off


def extract(text):
    out = {}
    for name in FUNCS:
        start = text.index(f"def {name}(")
        nxt = re.search(r"\n(?:def |class |@gtscript)", text[start + 10:])
        body = text[start: start + 10 + (nxt.start() if nxt else len(text))]
        entries = []
        for block in re.finditer(r"with horizontal\((.*?)\):\s*\n\s*(\w+)\s*=\s*(.*?)\n",
                                 body, re.S):
            regions = re.findall(r"region\[([^,\]]+),([^\]]+)\]", block.group(1))
            dst = block.group(2)
            rhs = block.group(3).strip()
            m = re.fullmatch(r"(?:(\w+)\s*\*\s*)?(\w+)\[\s*(-?\d+),\s*(-?\d+),\s*0\s*\]", rhs)
            if not m:
                raise ValueError(f"{name}: cannot parse {rhs!r}")
            mult, src, di, dj = m.group(1), m.group(2), int(m.group(3)), int(m.group(4))
            for ri, rj in regions:
                ai, a = _anchor(ri)
                aj, b = _anchor(rj)
                entries.append([dst, ai, a, aj, b, src, di, dj, mult or "1"])
        out[name] = entries
    return out


if __name__ == "__main__":
    tables = extract(REF.read_text())
    dest = pathlib.Path(__file__).with_name("corner_tables.json")
    lines = ["{"]
    for n, (k, v) in enumerate(tables.items()):
        rows = ",\n".join("  " + json.dumps(e) for e in v)
        lines.append(f" {json.dumps(k)}: [\n{rows}\n ]" + ("," if n < len(tables) - 1 else ""))
    lines.append("}")
    dest.write_text("\n".join(lines) + "\n")
    print({k: len(v) for k, v in tables.items()}, file=sys.stderr)
