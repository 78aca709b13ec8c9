"""Writes tests/golden/reference_tables.json: the numeric tables and test
values the reference holds for this path, read from its source text (run once
in a container where the reference is mounted; the JSON is committed as data).

  directions   DIRECTION_BINS_{16,21,32,64,84}   src/sweep/direction/healpix.rs
  bpass        BPASS_AGE_BINS, BPASS_METALLICITY_BINS, BPASS_TABLE and the
               check_is_close values of its test  src/arepo_postprocess/bpass.rs
  cosmology    the time difference test values    src/cosmology.rs

Float literals are converted by Python's float(), which rounds correctly,
as Rust's f64 parsing does, so every value is the f64 the reference uses.

    python tests/golden/make_reference_tables.py [/path/to/reference]
"""

from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FLOAT = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def const_block(text: str, name: str) -> tuple[str, int]:
    """The text between `const NAME ... = [` and its matching `];`, and the
    1-based line of the declaration."""
    m = re.search(r"const\s+" + name + r"\s*:[^=]*=\s*\[", text)
    if not m:
        raise ValueError(f"{name} not found")
    depth, i = 1, m.end()
    while depth:
        depth += {"[": 1, "]": -1}.get(text[i], 0)
        i += 1
    return text[m.end():i - 1], text.count("\n", 0, m.start()) + 1


def floats(block: str) -> list[float]:
    return [float(x) for x in re.findall(FLOAT, block)]


def main(ref: str) -> None:
    out = {"_about": "Tables and test values held by the reference's source (data only), read from the "
                     "files named in each entry by tests/golden/make_reference_tables.py."}
    path = "src/sweep/direction/healpix.rs"
    text = open(os.path.join(ref, path)).read()
    dirs = {}
    for n in (16, 21, 32, 64, 84):
        block, line = const_block(text, f"DIRECTION_BINS_{n}")
        v = floats(block.replace(f"[f64; 3]", ""))
        assert len(v) == 3 * n, (n, len(v))
        dirs[str(n)] = {"source": f"{path}:{line}", "bins": [v[3 * i:3 * i + 3] for i in range(n)]}
    out["directions"] = dirs

    path = "src/arepo_postprocess/bpass.rs"
    text = open(os.path.join(ref, path)).read()
    ages, la = const_block(text, "BPASS_AGE_BINS")
    mets, lm = const_block(text, "BPASS_METALLICITY_BINS")
    table, lt = const_block(text, "BPASS_TABLE")
    ages, mets, table = floats(ages), floats(mets), floats(table)
    assert len(ages) == 51 and len(mets) == 13 and len(table) == 51 * 13
    checks = []
    for m in re.finditer(r"check_is_close\(\s*(" + FLOAT + r")\s*,\s*(" + FLOAT + r")\s*,\s*(" + FLOAT + r")\s*\)",
                         text):
        checks.append({"age_yr": float(m.group(1)), "metallicity": float(m.group(2)),
                       "photons_per_s_per_solar_mass": float(m.group(3)),
                       "source": f"{path}:{text.count(chr(10), 0, m.start()) + 1}"})
    assert len(checks) == 3
    out["bpass"] = {
        "age_bins_yr": ages, "age_bins_source": f"{path}:{la}",
        "metallicity_bins": mets, "metallicity_bins_source": f"{path}:{lm}",
        "photons_per_s_per_solar_mass": [table[13 * i:13 * i + 13] for i in range(51)], "table_source": f"{path}:{lt}",
        "known_answers": checks,
        "known_answer_rel_tol": 1e-5,
    }

    path = "src/cosmology.rs"
    text = open(os.path.join(ref, path)).read()
    ol = float(re.search(r"omega_lambda:\s*(" + FLOAT + ")", text).group(1))
    o0 = float(re.search(r"omega_0:\s*(" + FLOAT + ")", text).group(1))
    h = float(re.search(r"\(cosmology,\s*(" + FLOAT + r")\.into\(\)\)", text).group(1))
    m = re.search(r"diff\((" + FLOAT + r"),\s*(" + FLOAT + r")\)\s*-\s*Time::gigayears\((" + FLOAT +
                  r")\)\)\.abs\(\)\s*<\s*Time::years\((" + FLOAT + r")\)", text)
    out["cosmology"] = {
        "omega_lambda": ol, "omega_0": o0, "little_h": h,
        "time_difference": [{"a0": float(m.group(1)), "a1": float(m.group(2)), "gigayears": float(m.group(3)),
                             "abs_tol_years": float(m.group(4))}, {"a0": 1.0, "a1": 1.0, "gigayears": 0.0,
                                                                   "abs_tol_years": 0.0}],
        "source": f"{path}:{text.count(chr(10), 0, m.start()) + 1}",
    }
    with open(os.path.join(HERE, "reference_tables.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
