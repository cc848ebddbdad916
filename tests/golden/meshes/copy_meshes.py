"""Copies the mesh files of four of the reference's test cases into this folder as fixtures (data
files, used unchanged as inputs of the mesh-file tests).

    python tests/golden/meshes/copy_meshes.py /path/to/reference

  tests/TaylorGreenVortexTest/mesh.xyz  -> tgv.xyz          128^3 periodic box (type 2 on every axis)
  tests/NeutralABLTest/mesh.xyz         -> neutral_abl.xyz  50^3 ABL box, i and k periodic (type 2)
  tests/ChannelFlowRe395Test/mesh.xyz   -> channel395.xyz   64^3 channel, i and k periodic (type 2)
  tests/SphereFlowTest/mesh.xyz         -> sphere.xyz       90 x 60 x 60, no periodic axis
"""
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
FILES = {"TaylorGreenVortexTest": "tgv.xyz", "NeutralABLTest": "neutral_abl.xyz",
         "ChannelFlowRe395Test": "channel395.xyz", "SphereFlowTest": "sphere.xyz"}

if __name__ == "__main__":
    ref = Path(sys.argv[1])
    for case, name in FILES.items():
        shutil.copyfile(ref / "tests" / case / "mesh.xyz", HERE / name)
        print(name, (HERE / name).stat().st_size, "bytes")
