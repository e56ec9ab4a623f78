"""The oracle's corner-fill tables against the reference's own (CPU).

tests/golden/corner_tables.json holds the region/offset tables of
fv3core/utils/corners.py as extracted from its text by
tests/golden/extract_corner_tables.py.  The oracle builds the tables from
rules (quarter turns about the tile corner); both must give the same
destination -> source pairs.
"""
import json
import pathlib

import pytest

from oracle import csw as ocs
from oracle import dsw as ods
from oracle.fvtp2d import Tile

FIXTURE = pathlib.Path(__file__).parent / "golden" / "corner_tables.json"
T = Tile(12, 12, 3, True, True, True, True)
ANCHOR = {"i_start": T.i_start, "i_end": T.i_end, "j_start": T.j_start, "j_end": T.j_end}


def _reference(name):
    rows = json.loads(FIXTURE.read_text())[name]
    out = []
    for dst, ai, a, aj, b, src, di, dj, mult in rows:
        i, j = ANCHOR[ai] + a, ANCHOR[aj] + b
        out.append((dst, i, j, src, i + di, j + dj, mult))
    return out


@pytest.mark.parametrize("kind,name", [("2x", "fill_corners_2cells_mult_x"),
                                       ("2y", "fill_corners_2cells_mult_y")])
def test_cell_fill_tables_match_reference(kind, name):
    ref = [(i, j, si, sj) for _, i, j, _, si, sj, _ in _reference(name)]
    mine = []
    for corner, ai, di, aj, dj, sdi, sdj in ocs.FILL_TABLES[kind]:
        i = {"is": T.i_start, "ie": T.i_end}[ai] + di
        j = {"js": T.j_start, "je": T.j_end}[aj] + dj
        mine.append((i, j, i + sdi, j + sdj))
    assert mine == ref  # same statements in the same order


@pytest.mark.parametrize("kind,name", [("3x", "fill_corners_3cells_mult_x"),
                                       ("3y", "fill_corners_3cells_mult_y")])
def test_three_cell_fill_tables_match_reference(kind, name):
    ref = [(i, j, si, sj) for _, i, j, _, si, sj, _ in _reference(name)]
    mine = []
    for corner, ai, di, aj, dj, sdi, sdj in ocs.FILL_TABLES[kind][8:]:
        i = {"is": T.i_start, "ie": T.i_end}[ai] + di
        j = {"js": T.j_start, "je": T.j_end}[aj] + dj
        mine.append((i, j, i + sdi, j + sdj))
    assert mine == ref


@pytest.mark.parametrize("kind,name", [("bgrid_x", "fill_corners_bgrid_x_defn"),
                                       ("bgrid_y", "fill_corners_bgrid_y_defn"),
                                       ("dgrid", "fill_corners_dgrid_defn")])
def test_rotation_rule_tables_match_reference(kind, name):
    fields = {"q_out": "q", "q_in": "q", "x_out": "x", "x_in": "x", "y_out": "y", "y_in": "y"}
    ref = {(fields[d], i, j): (fields[s], si, sj, m) for d, i, j, s, si, sj, m in _reference(name)}
    assert len(ref) == len(_reference(name))  # every destination is written once
    mine = {}
    for dst, i, j, src, si, sj, corner in ods.corner_fill_entries(kind, T):
        mult = "1"
        if kind == "dgrid" and corner in ("sw", "ne"):
            mult = "mysign"
        mine[(dst, i, j)] = (src, si, sj, mult)
    assert mine == ref
    # no source is also a destination: the statement order cannot matter
    dsts = {(d, i, j) for d, i, j in ref}
    assert not any((s, si, sj) in dsts for s, si, sj, _ in ref.values())
