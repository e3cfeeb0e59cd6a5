"""Extract the TIPS-2024 Q(T) tables (T = 70..400 K, 1 K steps) from the
reference's `src/HITRAN/tips2024_data.jl` into a C header shared by the HIP
library and the CPU oracle.

The tables are data (TIPS-2024 v1.2, Gamache et al. 2025, JQSRT 345, 109568,
doi:10.5281/zenodo.17191976, CC-BY-4.0).  This script only re-formats them;
the numbers are copied verbatim so the partition sums match the reference's
`_tips_lookup` (src/HITRAN/partition.jl:10-18) bit for bit.

Usage:  python extract_tips.py [/root/reference]   (writes tips2024_table.h here)
"""
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "src", "HITRAN", "tips2024_data.jl")
text = open(src, encoding="utf-8").read()

tmin = int(re.search(r"const TIPS_T_MIN\s*=\s*(\d+)", text).group(1))
tmax = int(re.search(r"const TIPS_T_MAX\s*=\s*(\d+)", text).group(1))
n_t = tmax - tmin + 1

entries = []
for m in re.finditer(r"\((\d+),\s*(\d+)\)\s*=>\s*\[([^\]]*)\]", text):
    mol, iso = int(m.group(1)), int(m.group(2))
    vals = [v.strip() for v in m.group(3).replace("\n", " ").split(",") if v.strip()]
    if len(vals) != n_t:
        raise SystemExit(f"({mol},{iso}): {len(vals)} values, expected {n_t}")
    float(vals[0])
    entries.append((mol, iso, vals))

out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tips2024_table.h")
with open(out, "w") as f:
    f.write("/* Generated by extract_tips.py from the reference's src/HITRAN/tips2024_data.jl.\n")
    f.write(" * TIPS-2024 v1.2 total internal partition sums Q(T), Gamache R. R. et al.,\n")
    f.write(" * JQSRT 345, 109568 (2025); erratum JQSRT 346, 109618 (2025);\n")
    f.write(" * doi:10.5281/zenodo.17191976.  License CC-BY-4.0.  Modified from upstream:\n")
    f.write(" * 46 isotopologues of 9 molecules, T = 70..400 K, re-formatted as a C table.\n")
    f.write(" * Do not edit by hand. */\n")
    f.write("#ifndef IRS_TIPS2024_TABLE_H\n#define IRS_TIPS2024_TABLE_H\n\n")
    f.write(f"#define IRS_TIPS_T_MIN {tmin}\n#define IRS_TIPS_T_MAX {tmax}\n")
    f.write(f"#define IRS_TIPS_NT {n_t}\n#define IRS_TIPS_NISO {len(entries)}\n\n")
    f.write("/* (mol_id, iso_id) of each table row, in row order */\n")
    f.write("static const int irs_tips_keys[IRS_TIPS_NISO][2] = {\n")
    f.write(",\n".join(f"    {{{mol}, {iso}}}" for mol, iso, _ in entries))
    f.write("\n};\n\n")
    f.write("static const double irs_tips_q[IRS_TIPS_NISO][IRS_TIPS_NT] = {\n")
    rows = []
    for mol, iso, vals in entries:
        lines = []
        for k in range(0, len(vals), 8):
            lines.append("        " + ", ".join(vals[k:k + 8]))
        rows.append(f"    /* mol {mol} iso {iso} */ {{\n" + ",\n".join(lines) + "\n    }")
    f.write(",\n".join(rows))
    f.write("\n};\n\n#endif\n")
print(f"wrote {out}: {len(entries)} isotopologues x {n_t} temperatures")
