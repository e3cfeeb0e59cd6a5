"""Re-format the MT-CKD continuum coefficient tables that the reference bundles
(data/mt_ckd_h2o/mt_ckd43_h2o_coeffs.csv, data/mt_ckd_co2/mt_ckd_co2_coeffs.csv)
into a C header shared by the HIP library and the CPU oracle.

The tables are third-party data from AER's MT_CKD / LBLRTM distribution; their
license allows redistribution when the AER copyright notice is reproduced, so
the notice is written into the header verbatim.  The numbers are copied as
written (decimal text), so the coefficients match the reference's
`_load_ckd_table` / `_load_ckd_co2_table` (src/CrossSections/continuum.jl:52-105)
exactly.  The CO2 table keeps S and xfac separate; the product S*xfac is formed
in FP64 at load time as the reference does (continuum.jl:90).

Usage:  python extract_mtckd.py [/root/reference]   (writes mtckd_table.h here)
"""
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
h2o = os.path.join(ref, "data", "mt_ckd_h2o", "mt_ckd43_h2o_coeffs.csv")
co2 = os.path.join(ref, "data", "mt_ckd_co2", "mt_ckd_co2_coeffs.csv")


def rows(path, ncol):
    with open(path) as f:
        head = f.readline().strip().split(",")
        if len(head) != ncol:
            raise SystemExit(f"{path}: {len(head)} columns, expected {ncol}")
        out = []
        for line in f:
            p = [v.strip() for v in line.strip().split(",")]
            if len(p) != ncol:
                continue
            [float(v) for v in p]
            out.append(p)
    return head, out


AER = """ *                  The MT_CKD Water Vapor Continuum
 *  Copyright (c), Atmospheric and Environmental Research, Inc., 2022
 *
 *  All rights reserved. This source code was developed as part of the
 *  LBLRTM software and is designed for scientific and research purposes.
 *  Atmospheric and Environmental Research Inc. (AER) grants USER the right
 *  to download, install, use and copy this software for scientific and
 *  research purposes only. This software may be redistributed as long as
 *  this copyright notice is reproduced on any copy made and appropriate
 *  acknowledgment is given to AER. This software or any modified version
 *  of this software may not be incorporated into proprietary software or
 *  commercial software offered for sale without the express written
 *  consent of AER.
 *
 *  This software is provided as is without any express or implied warranties.
"""

hh, hr = rows(h2o, 4)
ch, cr = rows(co2, 4)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mtckd_table.h")
with open(out, "w") as f:
    f.write("/* Generated by extract_mtckd.py from the MT-CKD coefficient tables bundled with\n")
    f.write(" * the reference (data/mt_ckd_h2o/mt_ckd43_h2o_coeffs.csv,\n")
    f.write(" * data/mt_ckd_co2/mt_ckd_co2_coeffs.csv).  Do not edit by hand.\n *\n")
    f.write(" * MT_CKD 4.3 H2O self/foreign continuum and the LBLRTM CO2 line-coupling\n")
    f.write(" * continuum (contnm.f90 BLOCK DATA BFCO2, XFACCO2, bandhead exponents).\n")
    f.write(" * Continuum coefficients from AER's MT_CKD model; cite Mlawer et al. (2012),\n")
    f.write(" * Phil. Trans. R. Soc. A 370, 2520-2556, doi:10.1098/rsta.2011.0295.\n")
    f.write(" * Upstream: AER-RC/MT_CKD and AER-RC/LBLRTM.  AER's notice, reproduced as its\n")
    f.write(" * license requires:\n *\n" + AER + " */\n")
    f.write("#ifndef IRS_MTCKD_TABLE_H\n#define IRS_MTCKD_TABLE_H\n\n")
    f.write(f"#define IRS_CKD_H2O_N {len(hr)}\n")
    f.write(f"/* columns: {', '.join(hh)} */\n")
    f.write("static const double irs_ckd_h2o[IRS_CKD_H2O_N][4] = {\n")
    for p in hr:
        f.write("    {" + ", ".join(p) + "},\n")
    f.write("};\n\n")
    f.write(f"#define IRS_CKD_CO2_N {len(cr)}\n")
    f.write(f"/* columns: {', '.join(ch)} */\n")
    f.write("static const double irs_ckd_co2[IRS_CKD_CO2_N][4] = {\n")
    for p in cr:
        f.write("    {" + ", ".join(p) + "},\n")
    f.write("};\n\n#endif\n")
print("wrote", out, len(hr), len(cr))
