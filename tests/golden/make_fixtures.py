"""Writes the golden fixtures under tests/golden/ from SVL's own validation data.

Run once in a container that has the SVL source tree mounted:

    python tests/golden/make_fixtures.py /path/to/SVL

It opens the debugging-case archives in ``03-Validations/01-Debugging`` and copies
*data only* -- the OpenSees reference histories the cases compare against and the
load time series the cases apply -- into compressed .npz files (no pickle). The model
definitions of the cases are rebuilt in ``tests/test_oracle_fixtures.py`` with this
package's own builder; no SVL source text is stored.

Cases (SVL case name, what it pins):
  F02  1 lin2DQuad4, Elastic2DPlaneStrain, lumped mass, Newmark dt=0.025, 200 steps
  J02  1 lin3DHexa8, Elastic3DLinear, lumped mass, Newmark dt=0.025, 200 steps
  F06  100 lin2DQuad4 soil column, Rayleigh, 2 Viscous1DLinear base dashpots, Newmark
  J05  100 lin3DHexa8 soil column, Rayleigh, 8 Viscous1DLinear base dashpots, Newmark
  F03  1 lin2DQuad4, PlasticPlaneStrainJ2, massless, Newton + Newmark dt=0.001, 999 steps
  F07  100 lin2DQuad4 J2 soil column, Rayleigh, 2 base dashpots, Newton + Newmark
       (first 4000 of 9999 steps kept)
  F11  10 PML2DQuad4 + 90 lin2DQuad4 soil column, EQUAL constraints, Newmark dt=0.01,
       1000 steps (the case ships no reference history: only its load series is kept)

Cases can be named after the path to rewrite only those (e.g. ``... /path/to/SVL F11``).
"""
from __future__ import annotations

import io
import sys
import zipfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

CASES = {
    "F02": ("F02-DY_Lin_2DPointLoad_ElasticPStrain_Quad4", "StressLoad.txt"),
    "J02": ("J02-DY_Lin_3DPointLoad_Elastic_Hexa8", "StressLoad.txt"),
    "F06": ("F06-DY_Lin_2DSoilColumn_ElasticPStrain_Quad4", "ricker.in"),
    "J05": ("J05-DY_Lin_3DSoilColumn_Elastic_Hexa8", "ricker.in"),
    "F03": ("F03-DY_Lin_2DPointLoad_J2PStrain_Quad4", "ricker.in"),
    "F07": ("F07-DY_Lin_2DSoilColumn_J2PStrain_Quad4", "ricker-SoilColumnPlasticPlaneStrainJ2.in"),
    "F11": ("F11-DY_Lin_2DPMLSoilColumn_ElasticPStrain_Quad4", "ricker-SoilColumnElasticPlaneStrainPML.txt"),
}
KEEP_ROWS = {"F07": 4000}


def read_member(z: zipfile.ZipFile, suffix: str) -> str:
    names = [n for n in z.namelist() if n.endswith(suffix)]
    if len(names) != 1:
        raise FileNotFoundError(f"{suffix}: {names}")
    return z.read(names[0]).decode()


def series_file(text: str) -> np.ndarray:
    """SVL TIMESERIES file: first value nt, then nt samples (Driver.hpp:1512-1535)."""
    v = np.loadtxt(io.StringIO(text))
    nt = int(v[0])
    return v[1:1 + nt]


def main(svl_root: str, only=()):
    root = Path(svl_root) / "03-Validations" / "01-Debugging"
    for key, (case, load_file) in CASES.items():
        if only and key not in only:
            continue
        with zipfile.ZipFile(root / f"{case}.zip") as z:
            load = series_file(read_member(z, f"{case}/{load_file}"))
            out = {"load": load}
            for resp in ("displacement", "velocity", "acceleration"):
                for prefix in ("OpenSees", "OpenSEES"):
                    try:
                        out[resp] = np.loadtxt(io.StringIO(read_member(z, f"{case}/{prefix}/{resp}.out")))
                        break
                    except FileNotFoundError:
                        continue
        if key in KEEP_ROWS:
            nk = KEEP_ROWS[key]
            out = {k: (v[:nk + 1] if k == "load" else v[:nk]) for k, v in out.items()}
        np.savez_compressed(HERE / f"{key}.npz", **out)
        print(key, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", tuple(sys.argv[2:]))
