"""Collect the physics data files of the EM step loop from the reference tree.

The reference keeps its Geant4-dumped physics tables as ASCII data under
``source/Geant/proxy/test/BasicCpuTransport/data`` (read by
``proxy/src/ProxyDataManager.cu:49-122`` and ``ProxySeltzerBerger.hpp:134-152``)
and its Sandia photo-absorption coefficients and atomic-shell binding energies
as numeric arrays in ``proxy/ProxySandiaData.hpp`` (``fSandiaTable``,
``fNbOfIntervals``, ``fZtoAratio``, ``fNumberOfShells``, ``fIndexOfShells``,
``fBindingEnergies``); the CMS 2015 solenoid field map is
``examples/magneticfield/simplifiedCMS/cmsmagfield2015.txt``.  This script
copies the table files and the field map verbatim and writes
the numeric arrays, for the elements the configured materials use, into plain
text files beside it.  It only runs where the reference tree is present; the
files it writes are committed so the library never needs the reference.

Usage: python extract_reference_data.py /root/reference
"""
import os
import re
import shutil
import sys

ELEMENTS = (8, 18, 74, 82)  # O, Ar, W, Pb: PbWO4, lAr, Pb


def _array_body(text, name):
    m = re.search(r"\b" + name + r"\s*\[[^\]]*\](?:\s*\[[^\]]*\])?\s*=\s*\{", text)
    if not m:
        raise RuntimeError("array %s not found" % name)
    i = m.end()
    depth = 1
    j = i
    while depth:
        c = text[j]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
        j += 1
    body = text[i:j - 1]
    body = re.sub(r"//[^\n]*", "", body)
    return body


def _numbers(body):
    return [float(x) for x in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", body)]


def main(ref):
    src = os.path.join(ref, "source/Geant/proxy")
    data = os.path.join(src, "test/BasicCpuTransport/data")
    here = os.path.dirname(os.path.abspath(__file__))

    # 1. Geant4-dumped tables (3 couples: Galactic, G4_Pb, G4_lAr) and cuts.
    out = os.path.join(here, "g4_pb_lar")
    os.makedirs(out, exist_ok=True)
    tdir = os.path.join(data, "table")
    for f in sorted(os.listdir(tdir)):
        shutil.copyfile(os.path.join(tdir, f), os.path.join(out, f))

    # 2. Seltzer-Berger scaled DCS grids for the elements in use.
    sbout = os.path.join(here, "brem_SB")
    os.makedirs(sbout, exist_ok=True)
    for z in ELEMENTS:
        shutil.copyfile(os.path.join(data, "brem_SB", "br%d" % z), os.path.join(sbout, "br%d" % z))

    # 3. Sandia coefficients and shell binding energies.
    text = open(os.path.join(src, "ProxySandiaData.hpp")).read()
    table = _numbers(_array_body(text, "fSandiaTable"))
    rows = [table[k:k + 5] for k in range(0, len(table), 5)]
    nint = [int(v) for v in _numbers(_array_body(text, "fNbOfIntervals"))]
    ztoa = _numbers(_array_body(text, "fZtoAratio"))
    nsh = [int(v) for v in _numbers(_array_body(text, "fNumberOfShells"))]
    ish = [int(v) for v in _numbers(_array_body(text, "fIndexOfShells"))]
    be = _numbers(_array_body(text, "fBindingEnergies"))
    assert len(nint) == 101 and len(ztoa) == 101 and len(rows) == 1 + sum(nint)
    assert len(nsh) == 101 and len(ish) == 101 and len(be) == 1540
    cumul = [1] * 101
    for z in range(1, 101):
        cumul[z] = cumul[z - 1] + nint[z]
    with open(os.path.join(here, "sandia.dat"), "w") as f:
        f.write("# Sandia photo-absorption parametrisation per element\n")
        f.write("# line 'Z nintervals Z/A'; then rows 'E_edge[keV] a1 a2 a3 a4'\n")
        f.write("# a_n in cm2 keV^n / g (sigma/rho = sum_n a_n / E^n)\n")
        for z in ELEMENTS:
            f.write("%d %d %.6g\n" % (z, nint[z], ztoa[z]))
            for r in rows[cumul[z - 1]:cumul[z - 1] + nint[z]]:
                f.write(" ".join("%.6g" % v for v in r) + "\n")
    with open(os.path.join(here, "shells.dat"), "w") as f:
        f.write("# atomic shell binding energies in eV, innermost first\n")
        f.write("# line 'Z nshells B_1 ... B_n'\n")
        for z in ELEMENTS:
            vals = be[ish[z]:ish[z] + nsh[z]]
            f.write("%d %d %s\n" % (z, nsh[z], " ".join("%.6g" % v for v in vals)))

    # 4. CMS (r, z) field map (181 x 161 nodes, tesla).
    fout = os.path.join(here, "field")
    os.makedirs(fout, exist_ok=True)
    shutil.copyfile(os.path.join(ref, "examples/magneticfield/simplifiedCMS/cmsmagfield2015.txt"),
                    os.path.join(fout, "cmsmagfield2015.txt"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
