"""Airfoil polar tables.

read_dat restates the polar reader of modeldata/NRELrotor.m:129-157: an
AeroDyn airfoil file (.dat) has 14 header lines (two lines of text, then the
number of tables, table ID, stall angle, three unused values, zero-Cn angle
of attack, Cn slope, Cn at positive and negative stall, angle of attack of
minimum Cd, minimum Cd) followed by rows of angle of attack (deg), Cl, Cd,
Cm.  Rows whose angle of attack repeats the next row's are dropped (the
first of each equal pair, NRELrotor.m:135-136).

The NREL 5-MW tables (Cylinder1, Cylinder2, DU21/25/30/35/40_A17,
NACA64_A17) are packaged as data (winds/data/nrel5mw_polars.json, written by
winds/data/make_nrel_polars.py from the eight .dat files with read_dat).
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

DATA = Path(__file__).resolve().parent / "data"
NREL5MW_NAMES = ("Cylinder1", "Cylinder2", "DU21_A17", "DU25_A17", "DU30_A17", "DU35_A17",
                 "DU40_A17", "NACA64_A17")


@dataclass
class Profile:
    """airfoils.profiles(k) of the reference (NRELrotor.m:138-155)."""
    name: str
    AoA: np.ndarray
    Cl: np.ndarray
    Cd: np.ndarray
    Cm: np.ndarray
    header: dict = field(default_factory=dict)


_HEADER = {4: "StallAoA", 8: "Cn0AoA", 9: "Lift0Cn", 10: "StallAoACn", 11: "StallAoANCn",
           12: "CdminAoA", 13: "Cdmin"}


def read_dat(path, name: str | None = None) -> Profile:
    """One AeroDyn .dat polar (NRELrotor.m:132-155)."""
    path = Path(path)
    lines = path.read_text().splitlines()
    if len(lines) < 15:
        raise ValueError(f"{path}: fewer than 14 header lines and one data row")
    header = {}
    for i, key in _HEADER.items():
        header[key] = float(lines[i].split()[0])
    rows = []
    for ln in lines[14:]:
        parts = ln.split()
        if not parts:
            continue
        rows.append([float(v) for v in parts[:4]])
    data = np.array(rows, dtype=np.float64)
    if data.ndim != 2 or data.shape[1] < 4:
        raise ValueError(f"{path}: expected rows of AoA, Cl, Cd, Cm")
    dup = np.nonzero(np.diff(data[:, 0]) == 0)[0]
    data = np.delete(data, dup, axis=0)
    if np.any(np.diff(data[:, 0]) <= 0):
        raise ValueError(f"{path}: angles of attack must increase")
    return Profile(name or path.stem, data[:, 0].copy(), data[:, 1].copy(), data[:, 2].copy(),
                   data[:, 3].copy(), header)


def to_json(profiles) -> str:
    return json.dumps([{"name": p.name, "AoA": p.AoA.tolist(), "Cl": p.Cl.tolist(),
                        "Cd": p.Cd.tolist(), "Cm": p.Cm.tolist(), "header": p.header}
                       for p in profiles], indent=0)


def from_json(text: str):
    return [Profile(d["name"], np.array(d["AoA"]), np.array(d["Cl"]), np.array(d["Cd"]),
                    np.array(d["Cm"]), d.get("header", {})) for d in json.loads(text)]


def nrel5mw_polars():
    """The eight NREL 5-MW tables in the order of airfoils.Names
    (NRELrotor.m:130)."""
    return from_json((DATA / "nrel5mw_polars.json").read_text())


def interp1(AoA, tab, x):
    """MATLAB interp1(AoA, tab, x): linear, NaN outside the table."""
    return np.interp(x, AoA, tab, left=np.nan, right=np.nan)
