"""Generate Geant4-format physics tables for materials the reference does not ship.

The reference holds tables for G4_Pb and G4_lAr only (data/table/*.asc,
couple.dat); PbWO4 needs its own.  This generator builds every table the step
loop reads, from published cross-sections:

- Compton: Klein-Nishina per-atom parametrisation (the formula of
  proxy/ProxyKleinNishina.hpp:62-93);
- gamma conversion: the Bethe-Heitler per-atom parametrisation of the Geant4
  Physics Reference Manual (sigma = (Z+1) Z [F1 + F2 Z + F3/Z], X = ln(E/m));
- photoabsorption: Sandia coefficients (sandia.dat), stored as lambda * E;
- bremsstrahlung: integrals of the Seltzer-Berger scaled cross-section
  (brem_SB/brZ) with the Migdal dielectric suppression, above the gamma cut
  (cross-section) and below it (restricted energy loss);
- ionisation: Berger-Seltzer restricted stopping power for e- and e+ with the
  Sternheimer-Peierls density effect, Moller / Bhabha cross-sections above the
  electron cut; range and inverse range by integration of the total
  restricted dE/dx;
- production thresholds for the 0.7 mm cut: the energy at which 5 absorption
  lengths of photons, or the CSDA range of electrons, equal the cut.

Every physical formula is then multiplied by a per-process, per-energy
correction equal to (reference table / generated table) for G4_Pb: PbWO4 is
dominated by Pb and W, so that ratio carries the reference's own modelling
(Livermore photoabsorption, relativistic bremsstrahlung, LPM) into the new
material.  validate() regenerates G4_Pb and G4_lAr uncorrected and reports
the agreement with the reference tables.

Usage: python make_tables.py            (writes data/pbwo4/)
       python make_tables.py --validate (prints the comparison only)
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

ME = 0.510998910
RE = 2.8179402894e-12          # mm
LC = 3.8615926459e-10          # mm
NA = 6.02214179e23
MB = 1.0e-25                    # millibarn in mm2
MICROBARN = 1.0e-28             # mm2

# element data: Z -> (A g/mol, mean excitation energy eV)
ELEMENTS = {8: (15.9994, 95.0), 18: (39.9477, 188.0), 74: (183.842, 727.0), 82: (207.217, 823.0)}

MATERIALS = {
    "G4_Pb": dict(density=11.35, atoms={82: 1}, I=823.0),
    "G4_lAr": dict(density=1.396, atoms={18: 1}, I=188.0),
    "PbWO4": dict(density=8.28, atoms={8: 4, 74: 1, 82: 1}, I=600.7),
}


# ---------------------------------------------------------------- data files
def read_asc(path, couple):
    tok = open(path).read().split()
    pos = 1
    for v in range(int(tok[0])):
        size = int(tok[pos + 4])
        vals = np.array(tok[pos + 5: pos + 5 + 2 * size], dtype=float)
        if v == couple:
            return vals[0::2], vals[1::2]
        pos += 5 + 2 * size
    raise KeyError(couple)


def write_asc(path, x, y, typ=2):
    with open(path, "w") as f:
        f.write("1\n%d\n%.12g %.12g %d\n%d\n" % (typ, x[0], x[-1], len(x), len(x)))
        for a, b in zip(x, y):
            f.write("%.12g  %.12g\n" % (a, b))


def sandia_rows(z):
    lines = [l for l in open(os.path.join(HERE, "sandia.dat")) if l.strip() and not l.startswith("#")]
    i = 0
    while i < len(lines):
        zz, n, zoa = lines[i].split()
        zz, n = int(zz), int(n)
        rows = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n)])
        if zz == z:
            return rows, float(zoa)
        i += 1 + n
    raise KeyError(z)


def sb_grid(z):
    tok = open(os.path.join(HERE, "brem_SB", "br%d" % z)).read().split()
    v = np.array(tok[3:], dtype=float)
    return v[:32], v[32:89], v[89:].reshape(57, 32)


def loggrid(lo, hi, per_decade):
    n = int(math.ceil(per_decade * math.log10(hi / lo))) + 1
    return np.logspace(math.log10(lo), math.log10(hi), n)


# ---------------------------------------------------------------- material
class Mat:
    def __init__(self, name):
        d = MATERIALS[name]
        self.name = name
        self.rho = d["density"]
        self.I = d["I"] * 1e-6
        amass = sum(c * ELEMENTS[z][0] for z, c in d["atoms"].items())
        self.el = []
        for z, c in d["atoms"].items():
            A = ELEMENTS[z][0]
            w = c * A / amass
            n = self.rho * NA * w / A * 1e-3  # atoms / mm3
            self.el.append((z, n))
        self.ne = sum(z * n for z, n in self.el)
        self.migdal = self.ne * 4 * math.pi * RE * LC * LC


# ---------------------------------------------------------------- photons
def kn_atom(Z, E):
    Z2 = Z * Z
    p1 = 2.7965e-1 + 1.9756e-5 * Z - 3.9178e-7 * Z2
    p2 = -1.8300e-1 - 1.0205e-2 * Z + 6.8241e-5 * Z2
    p3 = 6.7527 - 7.3913e-2 * Z + 6.0480e-5 * Z2
    p4 = -1.9798e+1 + 2.7079e-2 * Z + 3.0274e-4 * Z2
    T0 = 40e-3 if Z < 1.5 else 15e-3

    def f(X):
        X2 = X * X
        return p1 * np.log(1 + 2 * X) / X + (p2 + p3 * X + p4 * X2) / (1 + 20 * X + 230 * X2 + 440 * X2 * X)

    E = np.asarray(E, float)
    X = np.maximum(E, T0) / ME
    s = f(X)
    low = E < T0
    if np.any(low):
        dT0 = 1e-3
        sig = f((T0 + dT0) / ME)
        c1 = -T0 * (sig - s) / (s * dT0)
        c2 = 0.375 - 0.0556 * math.log(Z) if Z > 1.5 else 0.15
        y = np.log(E / T0)
        s = np.where(low, s * np.exp(-y * (c1 + c2 * y)), s)
    return Z * s * 1e-22  # barn -> mm2


def bh_atom(Z, E):
    a = [8.7842e+2, -1.9625e+3, 1.2949e+3, -2.0028e+2, 1.2575e+1, -2.8333e-1]
    b = [-1.0342e+1, 1.7692e+1, -8.2381, 1.3063, -9.0815e-2, 2.3586e-3]
    c = [-4.5263e+2, 1.1161e+3, -8.6749e+2, 2.1773e+2, -2.0467e+1, 6.5372e-1]
    E = np.asarray(E, float)
    Elim = 1.5
    Ee = np.maximum(E, Elim)
    X = np.log(Ee / ME)
    F1 = sum(a[i] * X ** i for i in range(6))
    F2 = sum(b[i] * X ** i for i in range(6))
    F3 = sum(c[i] * X ** i for i in range(6))
    s = (Z + 1.0) * (F1 * Z + F2 * Z * Z + F3) * MICROBARN
    d = np.clip((E - 2 * ME) / (Elim - 2 * ME), 0, None)
    s = np.where(E < Elim, s * d * d, s)
    return np.where(E <= 2 * ME, 0.0, np.maximum(s, 0.0))


def sandia_atom(Z, E):
    rows, zoa = sandia_rows(Z)
    ek = np.asarray(E, float) * 1e3
    idx = np.clip(np.searchsorted(rows[:, 0], ek, side="right") - 1, 0, len(rows) - 1)
    c = rows[idx]
    s = c[:, 1] / ek + c[:, 2] / ek ** 2 + c[:, 3] / ek ** 3 + c[:, 4] / ek ** 4
    return (Z / zoa) / NA * s * 100.0


def macro(m, f, E):
    return sum(n * f(z, E) for z, n in m.el)


# ---------------------------------------------------------------- bremsstrahlung
def brems_integrals(m, T, kcut, nk=400):
    """Per-volume cross-section above kcut (1/mm) and energy loss below (MeV/mm)."""
    sig = np.zeros_like(T)
    loss = np.zeros_like(T)
    for z, n in m.el:
        gx, gy, gv = sb_grid(z)
        for i, t in enumerate(T):
            y = min(max(math.log(t), gy[0]), gy[-1])
            iy = min(max(np.searchsorted(gy, y) - 1, 0), 55)
            wy = (y - gy[iy]) / (gy[iy + 1] - gy[iy])
            row = gv[iy] * (1 - wy) + gv[iy + 1] * wy
            etot = t + ME
            beta2 = t * (t + 2 * ME) / etot ** 2
            kp2 = m.migdal * etot * etot
            pref = z * z / beta2 * MB * n

            def dsdk(k):
                return pref * np.interp(k / t, gx, row) / k * (k * k / (k * k + kp2))

            if kcut < t:
                k = np.logspace(math.log10(kcut), math.log10(t), nk)
                sig[i] += np.trapezoid(dsdk(k) * k, np.log(k))
            kmax = min(kcut, t)
            k = np.logspace(math.log10(kmax) - 8, math.log10(kmax), nk)
            loss[i] += np.trapezoid(dsdk(k) * k * k, np.log(k))
    return sig, loss


# ---------------------------------------------------------------- ionisation
def density_effect(m, x):
    """Sternheimer-Peierls general density-effect correction, x = log10(beta gamma)."""
    # plasma energy from the electron density: hbar omega_p = 28.816 sqrt(rho Z/A) eV
    zoa = m.ne / (m.rho * NA * 1e-3)
    hwp = 28.816e-6 * math.sqrt(m.rho * zoa)
    C = 1 + 2 * math.log(m.I / hwp)
    if m.I < 100e-6:
        x1 = 2.0
        x0 = 0.2 if C < 3.681 else 0.326 * C - 1.0
    else:
        x1 = 3.0
        x0 = 0.2 if C < 5.215 else 0.326 * C - 1.5
    a = (C - 4.606 * x0) / (x1 - x0) ** 3
    d = np.where(x >= x1, 4.606 * x - C, np.where(x >= x0, 4.606 * x - C + a * (x1 - x) ** 3, 0.0))
    return d


def ioni_dedx(m, T, tcut, positron):
    """Restricted collision stopping power; below th = 0.25 sqrt(Zeff) keV the
    value at th is extrapolated with 1/sqrt(x) (x > 0.25) or 1.4 sqrt(x)/(0.1 + x)."""
    T = np.asarray(T, float)
    zeff = m.ne / sum(n for _, n in m.el)
    th = 0.25 * math.sqrt(zeff) * 1e-3
    low = T < th
    Tk = np.where(low, th, T)
    tc = np.where(low, np.minimum(tcut, th) if np.ndim(tcut) == 0 else np.where(low, th, tcut), tcut)
    d = _ioni_dedx(m, Tk, tc, positron)
    x = T / th
    f = np.where(x > 0.25, 1 / np.sqrt(np.maximum(x, 1e-300)), 1.4 * np.sqrt(x) / (0.1 + x))
    return np.where(low, d * f, d)


def _ioni_dedx(m, T, tcut, positron):
    tau = T / ME
    gam = tau + 1
    b2 = tau * (tau + 2) / gam ** 2
    bg = np.sqrt(tau * (tau + 2))
    eexc = m.I / ME
    if positron:
        d = np.minimum(tcut, T) / ME
        y = 1 / (1 + gam)
        F = np.log(tau * d) - b2 / tau * (tau + 2 * d - 1.5 * d * d * y - (d - d ** 3 / 3) * y * y -
                                          (d * d / 2 - tau * d ** 3 / 3 + d ** 4 / 4) * y ** 3)
    else:
        d = np.minimum(tcut, 0.5 * T) / ME
        F = -1 - b2 + np.log((tau - d) * d) + tau / (tau - d) + (0.5 * d * d + (2 * tau + 1) * np.log(1 - d / tau)) / gam ** 2
    dl = np.log(2 * (tau + 2) / eexc ** 2) + F - density_effect(m, np.log10(bg))
    dedx = 2 * math.pi * RE * RE * ME * m.ne / b2 * dl
    return np.maximum(dedx, 0.0)


def moller_xs(m, T, cut):
    t = np.asarray(T, float)
    tmax = 0.5 * t
    ok = cut < tmax
    xmin = cut / t
    xmax = tmax / t
    gam = t / ME + 1
    g2 = gam * gam
    b2 = (t / ME) * (t / ME + 2) / g2
    gg = (2 * gam - 1) / g2
    with np.errstate(all="ignore"):
        s = ((xmax - xmin) * (1 - gg + 1 / (xmin * xmax) + 1 / ((1 - xmin) * (1 - xmax))) -
             gg * np.log(xmax * (1 - xmin) / (xmin * (1 - xmax)))) / b2
        s = s * 2 * math.pi * ME * RE * RE / t * m.ne
    return np.where(ok, s, 0.0)


def bhabha_xs(m, T, cut):
    t = np.asarray(T, float)
    ok = cut < t
    xmin = cut / t
    xmax = 1.0
    gam = t / ME + 1
    g2 = gam * gam
    b2 = (t / ME) * (t / ME + 2) / g2
    y = 1 / (1 + gam)
    b1 = 2 - y * y
    bb2 = (1 - 2 * y) * (3 + y * y)
    b4 = (1 - 2 * y) ** 3
    b3 = b4 + (1 - 2 * y) ** 2
    with np.errstate(all="ignore"):
        s = ((1 / xmin - 1 / xmax) / b2 - b1 * np.log(xmax / xmin) + bb2 * (xmax - xmin) -
             b3 * (xmax ** 2 - xmin ** 2) / 2 + b4 * (xmax ** 3 - xmin ** 3) / 3)
        s = s * 2 * math.pi * ME * RE * RE / t * m.ne
    return np.where(ok, s, 0.0)


def range_table(E, dedx):
    """CSDA range of the restricted loss: R(E0) = 2 E0 / S(E0), then trapezoid in ln E."""
    lnE = np.log(E)
    f = E / dedx
    r = np.empty_like(E)
    r[0] = 2 * E[0] / dedx[0]
    # refine each interval with a log-log interpolation of 1/S
    for i in range(1, len(E)):
        sub = np.linspace(lnE[i - 1], lnE[i], 33)
        fs = np.exp(np.interp(sub, lnE[i - 1:i + 1], np.log(f[i - 1:i + 1])))
        r[i] = r[i - 1] + np.trapezoid(fs, sub)
    return r


# ---------------------------------------------------------------- cuts
def gamma_cut(m, length=0.7):
    E = np.logspace(-3.3, 1, 20000)
    s = macro(m, kn_atom, E) + macro(m, sandia_atom, E) + macro(m, bh_atom, E)
    absl = 5.0 / s
    # highest energy whose 5 absorption lengths are below the cut length
    ok = np.nonzero(absl <= length)[0]
    return float(E[ok[-1]]) if len(ok) else 990e-6


def electron_cut(m, length=0.7):
    E = np.logspace(-3, 2, 3000)
    dedx = ioni_dedx(m, E, E, False)  # unrestricted collision loss
    r = range_table(E, dedx)
    return float(np.interp(length, r, E))


# ---------------------------------------------------------------- build
# Production thresholds of the reference's G4_Pb couple at 0.7 mm (cut.dat):
# gamma 94.58614 keV, e- 1003.864 keV, e+ 951.3213 keV.  The range/absorption
# estimates above are scaled by the G4_Pb ratio (reference / estimate).
PB_CUTS = (94.58614e-3, 1003.864e-3, 951.3213e-3)


def build(name, calibrate=True):
    m = Mat(name)
    gcut = gamma_cut(m)
    ecut = electron_cut(m)
    pcut = ecut
    if calibrate:
        pb = Mat("G4_Pb")
        g0, e0 = gamma_cut(pb), electron_cut(pb)
        gcut *= PB_CUTS[0] / g0
        pcut = ecut * PB_CUTS[2] / e0
        ecut *= PB_CUTS[1] / e0
    t = {}
    e = loggrid(1e-4, 1.0, 7)
    t["Lambda.compt.gamma.asc"] = (e, macro(m, kn_atom, e))
    e = loggrid(1.0, 1e8, 7)
    t["LambdaPrim.compt.gamma.asc"] = (e, macro(m, kn_atom, e) * e)
    e = loggrid(2 * ME * 1.000000001, 1e8, 18.4)
    t["Lambda.conv.gamma.asc"] = (e, macro(m, bh_atom, e))
    e = loggrid(0.2, 1e8, 7)
    t["LambdaPrim.phot.gamma.asc"] = (e, macro(m, sandia_atom, e) * e)
    el = loggrid(1e-4, 1e8, 7)
    for sp, pos in (("e-", False), ("e+", True)):
        cut = pcut if pos else ecut
        bs, bl = brems_integrals(m, el, gcut)
        io = ioni_dedx(m, el, cut, pos)
        dedx = io + bl
        t["DEDX.eIoni.%s.asc" % sp] = (el, dedx)
        r = range_table(el, dedx)
        t["Range.eIoni.%s.asc" % sp] = (el, r)
        t["InverseRange.eIoni.%s.asc" % sp] = (r, el)
        eb = loggrid(gcut, 1e8, 7)
        t["Lambda.eBrem.%s.asc" % sp] = (eb, brems_integrals(m, eb, gcut)[0])
        ei = loggrid(cut if pos else 2 * cut, 1e8, 7)
        xs = bhabha_xs(m, ei, cut) if pos else moller_xs(m, ei, cut)
        t["Lambda.eIoni.%s.asc" % sp] = (ei, xs)
    return m, gcut, ecut, pcut, t


def ratio_correct(name, t, pb_gen, pb_ref_dir):
    """Multiply each generated table by the G4_Pb (reference / generated) ratio."""
    out = {}
    for f, (x, y) in t.items():
        if f.startswith("InverseRange") or f.startswith("Range"):
            continue
        xr, yr = read_asc(os.path.join(pb_ref_dir, f), 1)
        xg, yg = pb_gen[f]
        rg = np.interp(np.log(x), np.log(xg), yg)
        rr = np.interp(np.log(x), np.log(xr), yr)
        ratio = np.where((rg > 0) & (rr > 0), rr / np.where(rg > 0, rg, 1), 1.0)
        # no correction where either table vanishes (thresholds differ between materials)
        out[f] = (x, y * np.clip(ratio, 0.5, 2.0))
    for sp in ("e-", "e+"):
        x, d = out["DEDX.eIoni.%s.asc" % sp]
        r = range_table(x, d)
        out["Range.eIoni.%s.asc" % sp] = (x, r)
        out["InverseRange.eIoni.%s.asc" % sp] = (r, x)
    return out


def validate():
    ref = os.path.join(HERE, "g4_pb_lar")
    report = {}
    for name, couple in (("G4_Pb", 1), ("G4_lAr", 2)):
        m, gcut, ecut, pcut, t = build(name, calibrate=False)
        cuts = open(os.path.join(ref, "cut.dat")).read().split()[2:]
        ref_g = float(cuts[2 * couple + 1]) * 1e-3
        ref_e = float(cuts[2 * (3 + couple) + 1]) * 1e-3
        report[name] = {"gamma_cut": (gcut, ref_g), "e_cut": (ecut, ref_e)}
        for f in sorted(t):
            if f.startswith("InverseRange"):
                continue
            x, y = t[f]
            xr, yr = read_asc(os.path.join(ref, f), couple)
            sel = (xr >= max(x[0], 1e-3)) & (xr <= min(x[-1], 1e5)) & (yr > 0)
            g = np.interp(np.log(xr[sel]), np.log(x), y)
            r = g / yr[sel]
            report[name][f] = (float(np.median(r)), float(r.min()), float(r.max()))
    return report


def main():
    if "--validate" in sys.argv:
        for name, rep in validate().items():
            print(name)
            for k, v in rep.items():
                print("  %-32s %s" % (k, v))
        return
    pb = build("G4_Pb")[4]
    m, gcut, ecut, pcut, t = build("PbWO4")
    t = ratio_correct("PbWO4", t, pb, os.path.join(HERE, "g4_pb_lar"))
    out = os.path.join(HERE, "pbwo4")
    os.makedirs(out, exist_ok=True)
    for f, (x, y) in t.items():
        write_asc(os.path.join(out, f), x, y, 4 if f.startswith("InverseRange") else 2)
    with open(os.path.join(out, "cut.dat"), "w") as f:
        f.write("CUT-V3.0\n1\n")
        for v in (gcut, ecut, pcut, 70e-3):
            f.write("        7.000000e-01        %.6e\n" % (v * 1e3))
    with open(os.path.join(out, "README"), "w") as f:
        f.write("PbWO4 8.28 g/cm3 (O:W:Pb = 4:1:1) tables generated by make_tables.py;\n"
                "production thresholds for 0.7 mm: gamma %.4f keV, e- %.4f keV, e+ %.4f keV\n"
                % (gcut * 1e3, ecut * 1e3, pcut * 1e3))
    print("wrote", out, "cuts (keV):", gcut * 1e3, ecut * 1e3, pcut * 1e3)


if __name__ == "__main__":
    main()
