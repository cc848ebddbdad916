"""Writes the golden fixtures of tests/golden/: seeded small instances of each workload type, run
through the CPU oracle's rhs_les_step (UpdateCs -> UpdateNut -> Buoyancy -> FormU).

    python tests/golden/make_golden.py

Each .npz holds the case builder's name and keyword arguments (the inputs are rebuilt from them,
deterministically, by tosca_amd.case), input checksums that detect a drift of the builder, and the
four outputs.  The outputs start from zero arrays; rhs is accumulated with scale 1.
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from tosca_amd import case as C  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {
    "tgv_wrap_smag_central": ("tgv_case", dict(n=8, ptype=C.PERIODIC_WRAP)),
    "tgv_copy_dsm_quick": ("tgv_case", dict(n=8, ptype=C.PERIODIC_COPY, les_model=C.LES_DSM, div_scheme=C.DIV_QUICK)),
    "hill_dsm_quick": ("hill_case", dict(nx=12, ny=8, nz=8, les_model=C.LES_DSM)),
    "hill_dpasd_copy": ("hill_case", dict(nx=12, ny=8, nz=8, les_model=C.LES_DPASD, iperiod=C.PERIODIC_COPY)),
    "abl_smag_fourier": ("abl_case", dict(nx=12, ny=8, nz=8, n_modes=32)),
    "abl_dsm_16": ("abl_case", dict(nx=16, ny=16, nz=16, les_model=C.LES_DSM, n_modes=32)),
}


def build(name):
    builder, kw = CASES[name]
    return getattr(C, builder)(**kw)


def checksums(case):
    out = {}
    for k, t in (("ucat", case.ucat), ("ucont", case.ucont), ("theta", case.theta), ("aj", case.mesh["aj"]),
                 ("icsi", case.mesh["icsi"])):
        a = t.numpy()
        out[k] = [float(np.sum(a)), float(np.sqrt(np.sum(a * a)))]
    return out


def run_oracle(case):
    lay = case.lay
    mesh = {k: v.numpy().copy() for k, v in case.mesh.items() if k != "cent"}
    out = dict(rhs=np.zeros((lay.mz, lay.my, lay.mx, 3)), cs=np.zeros(lay.gshape), nut=np.zeros(lay.gshape),
               btheta=np.zeros((lay.mz, lay.my, lay.mx, 3)))
    O.rhs_les_step(lay.info(), mesh, case.params, case.ucont.numpy().copy(), case.ucat.numpy().copy(),
                   case.theta.numpy().copy(), out["rhs"], 1.0, out["cs"], out["nut"], out["btheta"], 1.0)
    return out


def main():
    for name in CASES:
        case = build(name)
        out = run_oracle(case)
        builder, kw = CASES[name]
        np.savez_compressed(HERE / f"{name}.npz", meta=np.array(json.dumps(dict(builder=builder, kwargs=kw))),
                            checksums=np.array(json.dumps(checksums(case))), **out)
        print(name, case.lay, {k: float(np.abs(v).max()) for k, v in out.items()})


if __name__ == "__main__":
    main()
