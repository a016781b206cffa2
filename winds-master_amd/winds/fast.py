"""FAST (v7) input and output files.

read_fst     core/input_import.m:26-153: the .fst input file, the AeroDyn
             file it names, the airfoil tables the AeroDyn file names, and
             the platform file.  Values are read from fixed lines of each
             file, as there: line k of a file is the k-th line of text.
read_out     core/output_import.m:24-70: the FAST output file (.out, or
             _SFunc.out for a Simulink run) next to the .fst, its time
             column rebuilt from the .fst's DT, every channel interpolated
             linearly to (t0:1/f:t1).
input_mod    core/input_mod.m:14-59: section nodes extended to hub and tip,
             optional cosine respacing to n nodes, trailing-node radii, chord
             and twist, the azimuth unwrapped and its remaining jumps bridged,
             BldPitch(:, k) = BldPitchk.
fast_case    a case for winds.cases.tables from a FAST file set: read_fst,
             read_out, input_mod, and wind.infty from the WindVxi, WindVyi and
             WindVzi channels (the case files set it the same way,
             modeldata/NRELrotor.m:162), section widths |diff(RTrail)| as
             there (:55).  The reference's own FAST branch (WInDS.m:106-108)
             calls neither input_mod.m nor sets wind.infty, and input_mod.m
             leaves one section width fewer than perform.m needs.

Quirk reproduced: with a zero output frequency, output_import.m:53 takes
1/mean(diff(...)) along the first ROW of the output (the first time step's
channel values), not along the time column; so does read_out.
The output file's channel names are the words of its seventh line (the
reference passes that line through genvarname, output_import.m:45); names
that are not identifiers are made into identifiers the same way
(non-identifier characters dropped, an 'x' prefix before a leading digit or
underscore, repeats numbered).
"""
from __future__ import annotations

import math
import re
from pathlib import Path
from types import SimpleNamespace

import numpy as np

from . import airfoils as af


def _num(line: str, kind=float):
    """sscanf(line, '%f') / '%i': the leading number of a line."""
    tok = line.split()
    if not tok:
        raise ValueError("expected a number on an empty line")
    m = re.match(r"[+-]?(\d+\.?\d*([eEdD][+-]?\d+)?|\.\d+([eEdD][+-]?\d+)?)", tok[0])
    if not m:
        raise ValueError(f"expected a number at the start of {line.strip()!r}")
    v = float(m.group(0).replace("d", "e").replace("D", "e"))
    return int(v) if kind is int else v


def _quoted(line: str) -> str:
    """The first double-quoted string of a line (strread(..., 'delimiter', '"'), item 2)."""
    parts = line.split('"')
    if len(parts) < 3:
        raise ValueError(f"expected a quoted file name in {line.strip()!r}")
    return parts[1]


def _lines(path: Path):
    return path.read_text().splitlines()


def _line(lines, k: int, path) -> str:
    """Line k (1-based), as data(k) of importdata."""
    if k > len(lines):
        raise ValueError(f"{path}: has {len(lines)} lines, line {k} is needed")
    return lines[k - 1]


def _rel(fstpath: Path, name: str) -> Path:
    """[fstpath name]: file names are relative to the .fst's folder; the
    reference's backslash separators are accepted."""
    return fstpath / Path(*name.replace("\\", "/").split("/"))


def genvarname(names):
    """Identifiers for channel names, repeats numbered (genvarname)."""
    out, seen = [], {}
    for n in names:
        v = re.sub(r"[^A-Za-z0-9_]", "", n)
        if not v or not v[0].isalpha():
            v = "x" + v
        if v in seen:
            seen[v] += 1
            v = f"{v}{seen[v]}"
        else:
            seen[v] = 0
        out.append(v)
    return out


def _airfoil(path: Path, name: str) -> af.Profile:
    """input_import.m:72-112: 14 header lines (stall data on lines 5 and
    9-14), rows of AoA, Cl, Cd, Cm; rows repeating the next row's angle
    dropped."""
    lines = _lines(path)
    if len(lines) < 15:
        raise ValueError(f"{path}: fewer than 14 header lines and one data row")
    header = {k: _num(lines[i]) for i, k in ((4, "StallAoA"), (8, "Cn0AoA"), (9, "Lift0Cn"),
                                              (10, "StallAoACn"), (11, "StallAoANCn"),
                                              (12, "CdminAoA"), (13, "Cdmin"))}
    rows = []
    for ln in lines[14:]:
        vals = []
        for tok in ln.split():
            try:
                vals.append(float(tok))
            except ValueError:
                break
        if vals:
            rows.append(vals)
    if not rows or min(len(r) for r in rows) < 4:
        raise ValueError(f"{path}: expected rows of AoA, Cl, Cd, Cm")
    data = np.array([r[:4] for r in rows], dtype=np.float64)
    dup = np.nonzero(np.diff(data[:, 0]) == 0)[0]
    data = np.delete(data, dup, axis=0)
    return af.Profile(name, data[:, 0].copy(), data[:, 1].copy(), data[:, 2].copy(),
                      data[:, 3].copy(), header)


def read_fst(filename):
    """core/input_import.m: (airfoils, blade, turbine, platform, wind)."""
    fst = Path(filename)
    fstpath = fst.parent
    data = _lines(fst)
    L = lambda k: _line(data, k, fst)
    turbine = SimpleNamespace(filename=str(fst))
    platform = SimpleNamespace()
    blade = SimpleNamespace()
    pf = _num(L(131), int)                                            # PtfmModel
    if pf >= 2:
        platform.filename = str(_rel(fstpath, _quoted(L(132))))
    else:
        platform.filename = "No platform model used."
    blade.filename = str(_rel(fstpath, _quoted(L(161))))              # ADFile
    blade.TipRad = _num(L(78))
    blade.HubRad = _num(L(79))
    turbine.NumBl = _num(L(9), int)
    turbine.OverHang = _num(L(83))
    turbine.TowerHt = _num(L(87))
    turbine.Twr2Shft = _num(L(88))
    turbine.ShftTilt = _num(L(90))
    turbine.PreCone = np.array([_num(L(92)), _num(L(93)), _num(L(94))])

    ad = Path(blade.filename)
    data = _lines(ad)
    L = lambda k: _line(data, k, ad)
    wind = SimpleNamespace(filename=str(_rel(fstpath, _quoted(L(10)))))
    nfoil = _num(L(18), int)
    names, profiles = [], []
    for c1 in range(1, nfoil + 1):
        rel = _quoted(L(18 + c1))
        stem = Path(rel.replace("\\", "/")).name[:-4]
        name = genvarname([stem])[0]
        names.append(name)
        profiles.append(_airfoil(_rel(fstpath, rel), name))
    airfoils = SimpleNamespace(Names=names, profiles=profiles)
    dm = 19 + nfoil
    ivnames = genvarname(L(dm + 1).split())
    rows = []
    for k in range(dm + 2, len(data) + 1):
        ln = data[k - 1]
        if not ln.strip():
            continue
        tok = ln.split()
        if len(tok) < 5:
            raise ValueError(f"{ad}: line {k} holds fewer than five blade-table values")
        rows.append([float(t) for t in tok[:4]] + [int(float(tok[4]))])
    tab = np.array(rows, dtype=np.float64).reshape(-1, 5)
    if len(ivnames) < 5:
        raise ValueError(f"{ad}: line {dm + 1} must name the five blade-table columns")
    for c3 in range(5):
        setattr(blade, ivnames[c3], tab[:, c3].copy())

    if pf in (0, 1):
        platform.Type = "onshore"
        platform.TwrDraft = platform.PtfmCM = platform.PtfmRef = 0.0
        platform.PtfmDraft = platform.PtfmDiam = 0.0
    elif pf in (2, 3):
        pth = Path(platform.filename)
        data = _lines(pth)
        L = lambda k: _line(data, k, pth)
        if pf == 2:
            platform.Type = "fixedoffshore"
            platform.TwrDraft, platform.PtfmCM, platform.PtfmRef = (_num(L(19)), _num(L(20)),
                                                                    _num(L(21)))
            platform.PtfmDraft = _num(L(36))                          # water depth
            platform.PtfmDiam = _num(L(31))
        else:
            platform.Type = _quoted(L(29)).replace("\\", "/").split("/")[-1]
            platform.TwrDraft, platform.PtfmCM, platform.PtfmRef = (_num(L(19)), _num(L(20)),
                                                                    _num(L(21)))
            platform.PtfmDraft = _num(L(32))
            platform.PtfmDiam = _num(L(33))
    return airfoils, blade, turbine, platform, wind


def _colon(a: float, step: float, b: float) -> np.ndarray:
    """MATLAB a:step:b (empty for a zero or NaN step, or one pointing away
    from b)."""
    if not (step != 0) or not np.isfinite(step) or (b - a) * step < 0:
        return np.zeros(0) if not (a == b and np.isfinite(step) and step != 0) else np.array([a])
    n = int(math.floor((b - a) / step * (1 + 4 * np.finfo(float).eps) + 1e-10)) + 1
    return a + np.arange(max(n, 0)) * step


def read_out(filename, t):
    """core/output_import.m: fastout with one attribute per channel, every
    channel at the times (t0:1/f:t1); t = (t0, t1, f)."""
    fst = Path(filename)
    data = _lines(fst)
    simq = _num(_line(data, 13, fst), int)                            # YCMode
    dt = _num(_line(data, 11, fst))                                   # DT
    s = str(fst)
    out = Path(s[:-4] + "_SFunc.out") if simq == 2 else Path(s[:-3] + "out")
    if not out.exists():
        out = Path(str(out)[:-4] + "_Sfunc.out")
        simq = 2
    lines = _lines(out)
    names = genvarname(_line(lines, 7, out).split())
    rows = [[float(v) for v in ln.split()] for ln in lines[8:] if ln.strip()]
    od = np.array(rows, dtype=np.float64)
    if od.ndim != 2 or od.shape[1] != len(names):
        raise ValueError(f"{out}: data rows must hold one value per channel ({len(names)})")
    if simq != 2:
        tt = _colon(od[0, 0], dt, od[-1, 0])
        if tt.size != od.shape[0]:
            raise ValueError(f"{out}: {od.shape[0]} rows but {tt.size} steps of DT = {dt} "
                             "between the first and last time")
        od[:, 0] = tt
    t0, t1, f = (float(v) for v in t)
    if f == 0:
        with np.errstate(divide="ignore"):
            f = 1 / np.mean(np.diff(od[0, :]))
    if t0 < od[0, 0]:
        t0 = od[0, 0]
        print(f"User selected initial time out-of-range, reset to {t0:g} seconds.\n")
    if t1 > od[-1, 0]:
        t1 = od[-1, 0]
        print(f"User selected final time out-of-range, reset to {t1:g} seconds.\n")
    with np.errstate(divide="ignore"):
        ti = _colon(t0, 1 / f, t1)
    fo = SimpleNamespace(filename=str(out))
    cols = [ti] + [np.interp(ti, od[:, 0], od[:, c], left=np.nan, right=np.nan)
                   for c in range(1, od.shape[1])]
    for name, col in zip(names, cols):
        setattr(fo, name, col)
    return fo


def pchip(x, y, xi):
    """Shape-preserving piecewise cubic Hermite interpolation (Fritsch and
    Carlson's monotone slopes with the Fritsch-Butland weighted harmonic
    mean and the one-sided three-point end slopes, as in C. Moler, Numerical
    Computing with MATLAB, sec. 3.4, pchiptx), extrapolating with the end
    pieces: MATLAB's interp1(x, y, xi, 'cubic').  y may have columns."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xi = np.asarray(xi, dtype=np.float64)
    one = y.ndim == 1
    Y = y[:, None] if one else y
    n = x.size
    h = np.diff(x)
    delta = np.diff(Y, axis=0) / h[:, None]
    d = np.zeros_like(Y)
    if n == 2:
        d[:] = delta[0]
    else:
        for c in range(Y.shape[1]):
            dl = delta[:, c]
            k = np.nonzero(np.sign(dl[:-1]) * np.sign(dl[1:]) > 0)[0] + 1
            w1 = 2 * h[k] + h[k - 1]
            w2 = h[k] + 2 * h[k - 1]
            d[k, c] = (w1 + w2) / (w1 / dl[k - 1] + w2 / dl[k])
            d[0, c] = _end_slope(h[0], h[1], dl[0], dl[1])
            d[-1, c] = _end_slope(h[-1], h[-2], dl[-1], dl[-2])
    k = np.clip(np.searchsorted(x, xi, side="right") - 1, 0, n - 2)
    s = (xi - x[k])[:, None]
    hk = h[k][:, None]
    dk, dk1, dl = d[k], d[k + 1], delta[k]
    c = (3 * dl - 2 * dk - dk1) / hk
    b = (dk - 2 * dl + dk1) / hk ** 2
    out = Y[k] + s * (dk + s * (c + s * b))
    return out[:, 0] if one else out


def _end_slope(h1, h2, del1, del2):
    d = ((2 * h1 + h2) * del1 - h1 * del2) / (h1 + h2)
    if np.sign(d) != np.sign(del1):
        d = 0.0
    elif np.sign(del1) != np.sign(del2) and abs(d) > abs(3 * del1):
        d = 3 * del1
    return d


def _nearest(x, y, xi):
    """interp1(x, y, xi, 'nearest'): a point halfway takes the upper node."""
    k = np.clip(np.searchsorted(x, xi, side="right") - 1, 0, x.size - 2)
    up = xi >= (x[k] + x[k + 1]) / 2
    return np.asarray(y)[k + up.astype(int)]


def input_mod(blade, fastout, turbine, n: int = 0):
    """core/input_mod.m on the imported blade and fastout (in place)."""
    blade.RNodes = np.concatenate([[blade.HubRad], blade.RNodes, [blade.TipRad]])
    blade.DRNodes = np.diff(blade.RNodes)
    blade.AeroTwst = -np.concatenate([blade.AeroTwst[:1], blade.AeroTwst, blade.AeroTwst[-1:]])
    blade.Chord = np.concatenate([blade.Chord[:1], blade.Chord, blade.Chord[-1:] / 2])
    blade.NFoil = np.concatenate([blade.NFoil[:1], blade.NFoil, blade.NFoil[-1:]])
    if n != 0:
        theta = _colon(math.pi / 2, -math.pi / (2 * (n - 1)), 0.0) if n > 1 else \
            np.array([math.pi / 2])
        rn = (blade.RNodes[-1] - blade.RNodes[0]) * np.cos(theta) + blade.RNodes[0]
        itp1 = pchip(blade.RNodes, np.column_stack([blade.AeroTwst, blade.Chord]), rn)
        itp2 = _nearest(blade.RNodes, blade.NFoil, rn)
        blade.RNodes = rn
        blade.DRNodes = np.diff(rn)
        blade.AeroTwst = itp1[:, 0]
        blade.Chord = itp1[:, 1]
        blade.NFoil = itp2
    blade.RTrail = np.concatenate([[blade.HubRad],
                                   blade.RNodes[1:-2] + blade.DRNodes[1:-1] / 2,
                                   [blade.TipRad]])
    blade.AeroTwstTrail = pchip(blade.RNodes, blade.AeroTwst, blade.RTrail)
    blade.ChordTrail = pchip(blade.RNodes, blade.Chord, blade.RTrail)
    blade.RNodes = blade.RNodes[1:-1]
    blade.DRNodes = blade.DRNodes[1:-1]
    blade.AeroTwst = blade.AeroTwst[1:-1]
    blade.Chord = blade.Chord[1:-1]
    blade.NFoil = np.asarray(blade.NFoil[1:-1]).astype(np.int32)

    az = np.unwrap(np.asarray(fastout.Azimuth, dtype=np.float64) * math.pi / 180) * 180 / math.pi
    jump = np.nonzero(np.diff(az) > 25)[0]
    az[np.concatenate([jump, jump + 1])] = np.nan
    bd = np.isnan(az)
    gd = np.nonzero(~bd)[0]
    if gd.size:
        bd[: gd.min()] = False
        bd[gd.max() + 1:] = False
        if bd.any():
            az[bd] = pchip(gd.astype(float), az[gd], np.nonzero(bd)[0].astype(float))
    fastout.Azimuth = az
    nb = int(turbine.NumBl)
    fastout.BldPitch = np.column_stack([getattr(fastout, f"BldPitch{k}") for k in
                                        range(1, nb + 1)])
    return blade, fastout


def fast_case(filename, t, n: int = 0):
    """A case for winds.cases.tables from a FAST file set (see module
    docstring).  Channels the kinematics need: Time, Azimuth, RotSpeed,
    TipSpdRat, BldPitch1..NumBl, NacYaw, PtfmSurge/Sway/Heave/Roll/Pitch/Yaw,
    WindVxi/Vyi/Vzi."""
    airfoils, blade, turbine, platform, wind = read_fst(filename)
    fo = read_out(filename, t)
    need = ["Time", "Azimuth", "RotSpeed", "TipSpdRat", "NacYaw", "PtfmSurge", "PtfmSway",
            "PtfmHeave", "PtfmRoll", "PtfmPitch", "PtfmYaw", "WindVxi", "WindVyi", "WindVzi"]
    need += [f"BldPitch{k}" for k in range(1, int(turbine.NumBl) + 1)]
    missing = [c for c in need if not hasattr(fo, c)]
    if missing:
        raise ValueError(f"{fo.filename}: missing output channels {missing}")
    input_mod(blade, fo, turbine, n)
    if np.any(np.isnan(fo.Time)):
        raise ValueError("output times outside the file")
    wind.infty = np.stack([fo.WindVxi, fo.WindVyi, fo.WindVzi], axis=1)
    wind.time = fo.Time
    wind.inftyM = np.sqrt(np.sum(wind.infty ** 2, axis=1))
    # input_mod.m:42 leaves ns-1 section widths; perform.m:49-50 needs one
    # per station, which the case files give as |diff(RTrail)| (NRELrotor.m:55)
    blade.DRNodes = np.abs(np.diff(blade.RTrail))
    blade.AR = blade.TipRad ** 2 / np.trapezoid(blade.ChordTrail, blade.RTrail)
    platform.PtfmCm = getattr(platform, "PtfmCM", 0.0)
    return SimpleNamespace(blade=blade, turbine=turbine, platform=platform, fastout=fo,
                           airfoils=airfoils, wind=wind, name="FAST")
