"""Writes dispersion_rows.json: coefficient rows of the reference's glass
catalogue (raypier/material_data/glass_dispersion_database.db, read as data
with sqlite3) for the dispersion formulas other than Schott's formula 2:

* every formula-1 row (Sellmeier with the pole wavelengths squared at
  evaluation) and every formula-5 row;
* formula-3 (polynomial) rows of other makers' glasses that carry a Schott
  formula-2 counterpart of the same type code (K-BK7 ~ N-BK7, E-SF5 ~ SF5 ...),
  with that counterpart's row.

Run from the repository root with the reference checked out at
/root/reference:  python tests/golden/make_dispersion_rows.py
"""
import json
import os
import re
import sqlite3

DB = "/root/reference/raypier/material_data/glass_dispersion_database.db"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dispersion_rows.json")
COLS = ["C%d" % k for k in range(1, 18)]


def rows(cur, where, args=()):
    cur.execute("select name, book, formula_id, wavelength_min, wavelength_max, n_coefs, %s from dispersion %s"
                % (", ".join(COLS), where), args)
    out = []
    for r in cur.fetchall():
        name, book, fid, wmin, wmax, nc = r[:6]
        out.append({"name": name, "book": book, "formula_id": fid, "wavelength_min": wmin,
                    "wavelength_max": wmax, "coefs": list(r[6:6 + nc])})
    return out


def main():
    con = sqlite3.connect("file:%s?mode=ro" % DB, uri=True)
    cur = con.cursor()
    schott = {r["name"]: r for r in rows(cur, "where book = 'schott' and formula_id = 2")}
    pairs = []
    for r in rows(cur, "where formula_id = 3 order by name, book"):
        base = re.sub(r"^[A-Z]+-", "", r["name"])
        for cand in ("N-" + base, base):
            if cand in schott:
                pairs.append({"poly": r, "schott": schott[cand]})
                break
    data = {"source": "raypier/material_data/glass_dispersion_database.db",
            "formula_1": rows(cur, "where formula_id = 1 order by name"),
            "formula_5": rows(cur, "where formula_id = 5 order by name"),
            "formula_3_pairs": pairs}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=0, sort_keys=True)
    print("%d formula-1 rows, %d formula-5 rows, %d formula-3 pairs -> %s" %
          (len(data["formula_1"]), len(data["formula_5"]), len(pairs), OUT))


if __name__ == "__main__":
    main()
