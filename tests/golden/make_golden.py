"""Write the golden fixtures that pin the oracle to the reference's own data.

Reads (only where the reference tree exists) the MRG32k3a jump matrices the
reference ships as constants (random/MRG.hpp:50-69: A1p76, A2p76, A1p127,
A2p127) and a few values of the reference-held Geant4 tables
(proxy/test/BasicCpuTransport/data/table/*.asc) that fix identities between
tables.  Output: tests/golden/reference_constants.json (numbers only).

Usage: python tests/golden/make_golden.py /root/reference
"""
import json
import os
import re
import sys


def matrix(text, name):
    m = re.search(name + r"\[ndim\]\[ndim\]\s*=\s*\{(.*?)\};", text, re.S)
    vals = [int(float(v)) for v in re.findall(r"[-+]?\d+\.?\d*", m.group(1))]
    assert len(vals) == 9, name
    return [vals[0:3], vals[3:6], vals[6:9]]


def table(path, couple):
    with open(path) as f:
        tok = f.read().split()
    pos = 1
    for v in range(int(tok[0])):
        size = int(tok[pos + 4])
        pairs = tok[pos + 5: pos + 5 + 2 * size]
        if v == couple:
            return [float(p) for p in pairs[0::2]], [float(p) for p in pairs[1::2]]
        pos += 5 + 2 * size
    raise KeyError(couple)


def main(ref):
    out = {}
    mrg = open(os.path.join(ref, "source/Geant/random/MRG.hpp")).read()
    for name in ("A1p76", "A2p76", "A1p127", "A2p127"):
        out[name] = matrix(mrg, name)
    tdir = os.path.join(ref, "source/Geant/proxy/test/BasicCpuTransport/data/table")
    ident = {}
    for couple, mat in ((1, "Pb"), (2, "lAr")):
        x1, y1 = table(os.path.join(tdir, "Lambda.compt.gamma.asc"), couple)
        x2, y2 = table(os.path.join(tdir, "LambdaPrim.compt.gamma.asc"), couple)
        ident["compt_lambda_at_1MeV_" + mat] = [x1[-1], y1[-1], x2[0], y2[0]]
        xd, yd = table(os.path.join(tdir, "DEDX.eIoni.e-.asc"), couple)
        xi, yi = table(os.path.join(tdir, "Ionisation.eIoni.e-.asc"), couple)
        xb, yb = table(os.path.join(tdir, "DEDX.eBrem.e-.asc"), couple)
        ident["dedx_sum_" + mat] = {"x": xd, "dedx": yd, "ioni": yi, "brem_x": xb, "brem": yb}
        xr, yr = table(os.path.join(tdir, "Range.eIoni.e-.asc"), couple)
        xv, yv = table(os.path.join(tdir, "InverseRange.eIoni.e-.asc"), couple)
        ident["range_inverse_" + mat] = {"range_x": xr, "range_y": yr, "inv_x": xv, "inv_y": yv}
        xp, yp = table(os.path.join(tdir, "LambdaPrim.phot.gamma.asc"), couple)
        ident["lambdaprim_phot_" + mat] = {"x": xp, "y": yp}
    out["table_identities"] = ident
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "reference_constants.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
