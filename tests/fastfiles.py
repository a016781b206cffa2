"""Writers of small FAST v7 file sets for the import tests.

No FAST files ship with the reference, so the tests write their own, laid
out on the lines core/input_import.m and core/output_import.m read: the
.fst (NumBl on line 9, DT 11, YCMode 13, TipRad 78, HubRad 79, OverHang 83,
TowerHt 87, Twr2Shft 88, ShftTilt 90, PreCone 92-94, PtfmModel 131,
PtfmFile 132, ADFile 161), the AeroDyn file (WindFile on line 10, NumFoil
18, airfoil files from 19, then BldNodes, the column names and the blade
table), AeroDyn airfoil files (14 header lines, rows AoA Cl Cd Cm), the
platform file (lines 19-21, 29, 31-33, 36) and the output file (channel
names on line 7, units on line 8, rows from line 9).
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

# NREL 5-MW AeroDyn blade table: RNodes, AeroTwst, DRNodes, Chord, NFoil
NREL_NODES = (2.8667, 5.6000, 8.3333, 11.7500, 15.8500, 19.9500, 24.0500, 28.1500, 32.2500,
              36.3500, 40.4500, 44.5500, 48.6500, 52.7500, 56.1667, 58.9000, 61.6333)
NREL_TWIST = (13.308, 13.308, 13.308, 13.308, 11.480, 10.162, 9.011, 7.795, 6.544, 5.361, 4.188,
              3.125, 2.319, 1.526, 0.863, 0.370, 0.106)
NREL_DR = (2.7333, 2.7500, 2.7333, 4.1000, 4.1000, 4.1000, 4.1000, 4.1000, 4.1000, 4.1000,
           4.1000, 4.1000, 4.1000, 4.1000, 2.7333, 2.7333, 2.7333)
NREL_CHORD = (3.542, 3.854, 4.167, 4.557, 4.652, 4.458, 4.249, 4.007, 3.748, 3.502, 3.256, 3.010,
              2.764, 2.518, 2.313, 2.086, 1.419)
NREL_FOIL = (1, 1, 2, 3, 4, 4, 5, 6, 6, 7, 7, 8, 8, 8, 8, 8, 8)


def _set(lines, k, text):
    lines[k - 1] = text


def write_fst(folder: Path, name="Test", numbl=3, dt=0.05, ycmode=0, tiprad=63.0, hubrad=1.5,
              overhang=5.0, towerht=90.0, twr2shft=0.0, shfttilt=5.0, precone=(-2.5,) * 3,
              ptfm_model=0, ptfm_file="Platform.dat", ad_file="AeroDyn.ipt"):
    lines = [f"---- line {k}: unused by the importer" for k in range(1, 171)]
    _set(lines, 1, "------- FAST INPUT FILE --------------------------------------")
    _set(lines, 9, f"{numbl:<12d}NumBl       - Number of blades (-)")
    _set(lines, 11, f"{dt:<12g}DT          - Integration time step (s)")
    _set(lines, 13, f"{ycmode:<12d}YCMode      - Yaw control mode")
    _set(lines, 78, f"{tiprad:<12g}TipRad      - The tip radius (m)")
    _set(lines, 79, f"{hubrad:<12g}HubRad      - The hub radius (m)")
    _set(lines, 83, f"{overhang:<12g}OverHang    - Distance from yaw axis to rotor apex (m)")
    _set(lines, 87, f"{towerht:<12g}TowerHt     - Height of tower above ground level (m)")
    _set(lines, 88, f"{twr2shft:<12g}Twr2Shft    - Vertical distance to the shaft (m)")
    _set(lines, 90, f"{shfttilt:<12g}ShftTilt    - Rotor shaft tilt angle (deg)")
    for k in range(3):
        _set(lines, 92 + k, f"{precone[k]:<12g}PreCone({k + 1})  - Blade {k + 1} cone angle (deg)")
    _set(lines, 131, f"{ptfm_model:<12d}PtfmModel   - Platform model {{0: none, 1: onshore, "
                     "2: fixed bottom offshore, 3: floating offshore}")
    _set(lines, 132, f'"{ptfm_file}"   PtfmFile    - Name of file containing platform properties')
    _set(lines, 161, f'"{ad_file}"   ADFile      - Name of file containing AeroDyn input parameters')
    p = folder / f"{name}.fst"
    p.write_text("\n".join(lines) + "\n")
    return p


def write_airfoil(path: Path, aoa, cl, cd, cm, stall=15.0):
    head = ["AeroDyn airfoil file", "written for the FAST import tests",
            "1           Number of airfoil tables in this file",
            "0           Table ID parameter",
            f"{stall:<12g}Stall angle (deg)",
            "0           No longer used", "0           No longer used", "0           No longer used",
            "-2.0        Zero lift angle of attack (deg)",
            "6.28        Cn slope for zero lift (dimensionless)",
            "1.5         Cn at stall value for positive angle of attack",
            "-0.8        Cn at stall value for negative angle of attack",
            "0.0         Angle of attack for minimum CD (deg)",
            "0.006       Minimum CD value"]
    rows = [f"{a:10.4f}\t{l:10.6f}\t{d:10.6f}\t{m:10.6f}" for a, l, d, m in zip(aoa, cl, cd, cm)]
    path.write_text("\n".join(head + rows) + "\n")


def write_aerodyn(folder: Path, foils, nodes=NREL_NODES, twist=NREL_TWIST, dr=NREL_DR,
                  chord=NREL_CHORD, nfoil=NREL_FOIL, name="AeroDyn.ipt",
                  wind_file="Wind/steady.wnd"):
    lines = ["AeroDyn input file for the FAST import tests",
             "SI          SysUnits", "BEDDOES     StallMod", "USE_CM      UseCm",
             "DYNIN       InfModel", "SWIRL       IndModel", "0.005       AToler",
             "PRANDtl     TLModel", "PRANDtl     HLModel",
             f'"{wind_file}"  WindFile - Name of file containing wind data',
             "90.0        HH", "0.0         TwrShad", "9999.9      ShadHWid",
             "9999.9      T_Shad_Refpt", "1.225       AirDens", "1.464E-5    KinVisc",
             "0.02479     DTAero", f"{len(foils):<12d}NumFoil"]
    lines += [f'"{f}"' for f in foils]
    lines.append(f"{len(nodes):<12d}BldNodes")
    lines.append("RNodes   AeroTwst  DRNodes  Chord  NFoil  PrnElm")
    for r, t, d, c, f in zip(nodes, twist, dr, chord, nfoil):
        lines.append(f"{r:.4f}\t{t:.3f}\t{d:.4f}\t{c:.3f}\t{f:d}\tNOPRINT")
    p = folder / name
    p.write_text("\n".join(lines) + "\n")
    return p


def write_platform(folder: Path, name="Platform.dat", twrdraft=-10.0, ptfmcm=-89.9, ptfmref=0.5,
                   draft31=6.0, d32=120.0, d33=9.4, depth36=20.0, type_file="HydroData\\Spar"):
    lines = [f"---- platform line {k}" for k in range(1, 41)]
    _set(lines, 19, f"{twrdraft:<12g}TwrDraft")
    _set(lines, 20, f"{ptfmcm:<12g}PtfmCM")
    _set(lines, 21, f"{ptfmref:<12g}PtfmRef")
    _set(lines, 29, f'"{type_file}"   WAMITFile')
    _set(lines, 31, f"{draft31:<12g}line 31")
    _set(lines, 32, f"{d32:<12g}line 32")
    _set(lines, 33, f"{d33:<12g}line 33")
    _set(lines, 36, f"{depth36:<12g}WtrDpth")
    p = folder / name
    p.write_text("\n".join(lines) + "\n")
    return p


def write_out(path: Path, channels: dict):
    """channels: name -> column (Time first)."""
    names = list(channels)
    cols = np.column_stack([np.asarray(channels[n], dtype=np.float64) for n in names])
    lines = ["", "These predictions were written for the FAST import tests", "",
             "Generated on 04-Dec-2026", "", ""]
    lines.append("\t".join(f"{n:<10s}" for n in names))
    lines.append("\t".join(f"{'(-)':<10s}" for _ in names))
    for row in cols:
        lines.append("\t".join(repr(float(v)) for v in row))
    path.write_text("\n".join(lines) + "\n")


def rigid_channels(time, omega, nb=3, wind=(11.4, 0.0, 0.0), pitch=0.0, yaw=0.0, tsr=7.0,
                   wrap=True):
    az = omega * 180 / math.pi * time
    ch = {"Time": time, "WindVxi": np.full_like(time, wind[0]),
          "WindVyi": np.full_like(time, wind[1]), "WindVzi": np.full_like(time, wind[2]),
          "Azimuth": np.mod(az, 360.0) if wrap else az,
          "RotSpeed": np.full_like(time, omega * 30 / math.pi),
          "TipSpdRat": np.full_like(time, tsr), "NacYaw": np.full_like(time, yaw)}
    for k in range(1, nb + 1):
        ch[f"BldPitch{k}"] = np.full_like(time, pitch)
    for k in ("PtfmSurge", "PtfmSway", "PtfmHeave", "PtfmRoll", "PtfmPitch", "PtfmYaw"):
        ch[k] = np.zeros_like(time)
    return ch


def nrel_file_set(folder: Path, t_end=10.0, dt=0.05, ptfm_model=0, ycmode=0):
    """.fst + AeroDyn + eight airfoil files + .out of the NREL 5-MW rotor
    turning rigidly at TSR 7 in 11.4 m/s (the reference's default case)."""
    from winds import airfoils as af
    folder.mkdir(parents=True, exist_ok=True)
    (folder / "AeroData").mkdir(exist_ok=True)
    foils = []
    for p in af.nrel5mw_polars():
        rel = f"AeroData/{p.name}.dat"
        write_airfoil(folder / rel, p.AoA, p.Cl, p.Cd, p.Cm)
        foils.append(rel)
    write_aerodyn(folder, foils)
    write_platform(folder)
    fst = write_fst(folder, dt=dt, ptfm_model=ptfm_model, ycmode=ycmode)
    time = np.arange(int(round(t_end / dt)) + 1) * dt
    omega = 7 * 11.4 / 63.0
    out = fst.with_name(fst.stem + ("_SFunc.out" if ycmode == 2 else ".out"))
    write_out(out, rigid_channels(time, omega))
    return fst
