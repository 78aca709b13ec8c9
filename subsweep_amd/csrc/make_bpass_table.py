"""Writes bpass_table.h, the BPASS ionising-photon table that
bpass_lookup interpolates (src/arepo_postprocess/bpass.rs:58-880 of the
reference: 51 stellar-population ages, 13 metallicities, photons/s per solar
mass), by reading the numbers out of that file as data.

    python subsweep_amd/csrc/make_bpass_table.py /root/reference/src/arepo_postprocess/bpass.rs
"""

import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def numbers_of(text: str, name: str):
    start = text.index(f"const {name}")
    body = text[text.index("= [", start) + 3:]
    depth, end = 1, 0
    for i, ch in enumerate(body):
        depth += ch == "["
        depth -= ch == "]"
        if depth == 0:
            end = i
            break
    return [float(x) for x in re.findall(r"[-+]?\d+\.\d+(?:e[-+]?\d+)?|[-+]?\d+e[-+]?\d+", body[:end])]


def main(path: str) -> None:
    text = open(path).read()
    ages = numbers_of(text, "BPASS_AGE_BINS")
    mets = numbers_of(text, "BPASS_METALLICITY_BINS")
    table = numbers_of(text, "BPASS_TABLE")
    assert len(ages) == 51 and len(mets) == 13 and len(table) == 51 * 13, (len(ages), len(mets), len(table))
    out = ["// BPASS ionising photon rates per solar mass of a stellar population,",
           "// by age (years) and metallicity (mass fraction): the table that the",
           "// reference's bpass_lookup interpolates (src/arepo_postprocess/bpass.rs:58-880),",
           "// read from that file as data by make_bpass_table.py. Generated; do not edit.",
           "#pragma once", "",
           "constexpr int kBpassAges = 51;", "constexpr int kBpassMetallicities = 13;", "",
           "constexpr double kBpassAgeBins[kBpassAges] = {"]
    out += [f"    {v!r}," for v in ages]
    out += ["};", "", "constexpr double kBpassMetallicityBins[kBpassMetallicities] = {"]
    out += [f"    {v!r}," for v in mets]
    out += ["};", "", "// [age][metallicity], photons/s per solar mass",
            "constexpr double kBpassTable[kBpassAges][kBpassMetallicities] = {"]
    for a in range(51):
        out.append("    {" + ", ".join(repr(v) for v in table[13 * a:13 * a + 13]) + "},")
    out += ["};", ""]
    with open(os.path.join(HERE, "bpass_table.h"), "w") as f:
        f.write("\n".join(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/arepo_postprocess/bpass.rs")
