"""The oracle's setup statistics (src/parcsr_ls/par_stats.c:20) against the
reference's own recorded output and closed forms.

The recorded `ij` run on the 30^3 7-point Laplacian (tests/golden/
reference_runs.json, printed by the reference with print level 1) gives per
level the rows, entries, entries per row and row sums of every operator, and
rows, columns, entries per row and row sums of every interpolation matrix,
to the four significant digits the table prints, plus the grid and operator
complexities.  A 2x2x2 Laplacian stops at one level, where every figure of
the table follows from the stencil.
"""
import re

import pytest


def parse_stats(text):
    """The operator table, the interpolation table and the complexities of a
    setup statistics text."""
    lines = text.splitlines()
    i = lines.index("Operator Matrix Information:")
    ops = []
    k = i + 5
    while lines[k].strip():
        f = lines[k].split()
        ops.append({"lev": int(f[0]), "rows": int(f[1]), "nnz": int(f[2]), "sparse": float(f[3]),
                    "min_row": int(f[4]), "max_row": int(f[5]), "avg": float(f[6]),
                    "rowsum_min": float(f[7]), "rowsum_max": float(f[8])})
        k += 1
    j = lines.index("Interpolation Matrix Information:")
    interp = []
    k = j + 4
    while lines[k].strip():
        f = lines[k].split()
        interp.append({"lev": int(f[0]), "rows": int(f[1]), "cols": int(f[3]),
                       "min_row": int(f[4]), "max_row": int(f[5]), "avg": float(f[6]),
                       "min_weight": float(f[7]), "max_weight": float(f[8]),
                       "rowsum_min": float(f[9]), "rowsum_max": float(f[10])})
        k += 1
    cx = {}
    for key in ("grid", "operator", "memory"):
        m = re.search(r"\b%s = ([0-9.]+)" % key, text)
        cx[key] = float(m.group(1))
    return ops, interp, cx


def test_stats_match_the_recorded_30_7pt_run(oracle, golden):
    g = golden["ij_30_7pt"]
    H = oracle.Hierarchy(oracle.gen_laplace7(30, 30, 30), max_row_sum=g["max_row_sum"])
    ops, interp, cx = parse_stats(H.setup_stats(max_iter=20, tol=1e-7))
    assert len(ops) == len(g["levels"]) and len(interp) == len(g["interp"])
    for got, ref in zip(ops, g["levels"]):
        for key in ("rows", "nnz", "min_row", "max_row"):
            assert got[key] == ref[key], (got["lev"], key)
        for key in ("rowsum_min", "rowsum_max"):
            assert got[key] == pytest.approx(ref[key], rel=1e-12, abs=0), (got["lev"], key)
    for got, ref in zip(interp, g["interp"]):
        for key in ("rows", "cols", "min_row", "max_row"):
            assert got[key] == ref[key], (got["lev"], key)
        for key in ("rowsum_min", "rowsum_max"):
            assert got[key] == pytest.approx(ref[key], rel=1e-12, abs=0), (got["lev"], key)
    assert cx["grid"] == g["grid_complexity"]
    assert cx["operator"] == g["operator_complexity"]
    # each operator once and each P once (par_stats.c:597-615, :883)
    nnz_a = sum(o["nnz"] for o in ops)
    nnz_p = sum(H.P(l).nnz for l in range(len(interp)))
    assert cx["memory"] == pytest.approx((nnz_a + nnz_p) / ops[0]["nnz"], abs=5e-7)


def test_stats_one_level_closed_form(oracle):
    """8 rows, each a cube corner: 4 entries (diagonal 6, three -1), row sum 3;
    with max_levels 1 the hierarchy is the operator alone."""
    H = oracle.Hierarchy(oracle.gen_laplace7(2, 2, 2), max_levels=1)
    text = H.setup_stats(max_iter=1, tol=0.0)
    ops, interp, cx = parse_stats(text)
    assert ops == [{"lev": 0, "rows": 8, "nnz": 32, "sparse": 0.5, "min_row": 4, "max_row": 4,
                    "avg": 4.0, "rowsum_min": 3.0, "rowsum_max": 3.0}]
    assert interp == []
    assert cx == {"grid": 1.0, "operator": 1.0, "memory": 1.0}
    assert " Max levels = 1\n Num levels = 1\n" in text
    assert "  Maximum number of cycles:         1 \n" in text
    assert "   Type 0=Jac, 3=hGS, 6=hSGS, 9=GE:     18   18     9 \n" in text


def test_stats_table_layout(oracle):
    """Column widths of par_stats.c:443-486 and :702-736: at least 7 digits of
    rows and 8 of entries in the operator table, 5 of rows in the
    interpolation table, and rules 49 + widths and 70 + 2 * width long."""
    H = oracle.Hierarchy(oracle.gen_laplace7(12, 12, 12))
    lines = H.setup_stats().splitlines()
    i = lines.index("Operator Matrix Information:")
    assert lines[i + 4] == "=" * (49 + 7 + 8 + 6)
    assert lines[i + 5].startswith("  0    1728    11232  0.004     4    7     6.5   0.000e+00")
    j = lines.index("Interpolation Matrix Information:")
    assert lines[j + 3] == "=" * (70 + 2 * 5)
    assert re.match(r"  0  1728 x \d+ *  1 ", lines[j + 4])
