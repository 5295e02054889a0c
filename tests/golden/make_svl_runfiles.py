"""Writes tests/golden/svl_runfiles/: the run files SVL's own pre-processor writes for its
validation cases (test infrastructure; run in a container that has the SVL tree):

    python tests/golden/make_svl_runfiles.py /path/to/SVL

For each case it unpacks the case archive from 03-Validations/01-Debugging into a temporary
folder and runs the case script there with SVL's 01-Pre_Process on PYTHONPATH, which calls
SeismoVLAB.CreateRunAnalysisFiles (01-Pre_Process/Core/SeismoVLAB.py:801-849). The JSON it
writes to Partition/, and the load files the JSON names, are copied here as data; no SVL
source text is stored. The pre-processor reads the case's data files as shipped.

Cases: F02 J02 F06 J05 F03 F07 F11 (see tests/golden/make_fixtures.py for what each pins).
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile
import zipfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "svl_runfiles"

CASES = {
    "F02": "F02-DY_Lin_2DPointLoad_ElasticPStrain_Quad4",
    "J02": "J02-DY_Lin_3DPointLoad_Elastic_Hexa8",
    "F06": "F06-DY_Lin_2DSoilColumn_ElasticPStrain_Quad4",
    "J05": "J05-DY_Lin_3DSoilColumn_Elastic_Hexa8",
    "F03": "F03-DY_Lin_2DPointLoad_J2PStrain_Quad4",
    "F07": "F07-DY_Lin_2DSoilColumn_J2PStrain_Quad4",
    "F11": "F11-DY_Lin_2DPMLSoilColumn_ElasticPStrain_Quad4",
}


def run_case(svl: Path, tag: str, name: str):
    with tempfile.TemporaryDirectory() as tmp:
        with zipfile.ZipFile(svl / "03-Validations" / "01-Debugging" / f"{name}.zip") as z:
            z.extractall(tmp)
        case = Path(tmp) / name
        env = dict(os.environ, PYTHONPATH=str(svl / "01-Pre_Process"))
        r = subprocess.run([sys.executable, f"{name}.py"], cwd=case, env=env, capture_output=True, text=True,
                           timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"{tag}: the case script failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        dest = OUT / tag
        if dest.exists():
            shutil.rmtree(dest)
        (dest / "Partition").mkdir(parents=True)
        files = set()
        for js in sorted((case / "Partition").glob("*.json")):
            shutil.copy(js, dest / "Partition" / js.name)
            d = json.load(open(js))
            for ld in d.get("Loads", {}).values():
                f = ld.get("attributes", {}).get("file")
                if f:
                    files.add(f)
        for f in sorted(files):
            src = case / f
            if not src.exists():
                raise RuntimeError(f"{tag}: load file {f} named by the run file is missing")
            (dest / f).parent.mkdir(parents=True, exist_ok=True)
            shutil.copy(src, dest / f)
        print(f"{tag}: {[p.name for p in (dest / 'Partition').iterdir()]} + {sorted(files)}")


def main():
    if len(sys.argv) < 2:
        raise SystemExit(__doc__)
    svl = Path(sys.argv[1])
    tags = sys.argv[2:] or list(CASES)
    for tag in tags:
        run_case(svl, tag, CASES[tag])


if __name__ == "__main__":
    main()
