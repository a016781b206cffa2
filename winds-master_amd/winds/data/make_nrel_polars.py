"""Writes nrel5mw_polars.json beside this script from the NREL 5-MW airfoil
files of a WInDS checkout (modeldata/NREL_files/*.dat), read with
winds.airfoils.read_dat.

    python make_nrel_polars.py /path/to/WInDS-Master
"""
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from winds import airfoils  # noqa: E402


def main(root: str) -> None:
    d = Path(root) / "modeldata" / "NREL_files"
    profs = [airfoils.read_dat(d / f"{n}.dat", n) for n in airfoils.NREL5MW_NAMES]
    (HERE / "nrel5mw_polars.json").write_text(airfoils.to_json(profs))
    print(f"wrote {len(profs)} tables: " + ", ".join(f"{p.name} ({p.AoA.size})" for p in profs))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
