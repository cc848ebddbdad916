"""The RHS+LES step on the reference test cases' own meshes (tests/golden/meshes, read with
tosca_amd.io), GPU against the CPU oracle on identical inputs, both kernel paths; the outputs then
go through a PETSc-binary restart round trip.  Bar: relative L2 <= 1e-12 per output field."""
import os

import numpy as np
import pytest
import torch

from tosca_amd import case as C
from tosca_amd import io
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr

pytestmark = pytest.mark.gpu
MESHES = "tests/golden/meshes"

CASES = [
    ("neutral_abl", "abl", C.LES_DSM, C.DIV_CENTRAL),
    ("channel395", "abl", C.LES_DSM, C.DIV_QUICK),
    ("sphere", "abl", C.LES_SMAGORINSKY, C.DIV_CENTRAL),
]


@pytest.fixture(scope="module", params=["fused", "ref"])
def rhs_mod(request):
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    old = os.environ.get("TOSCA_PATH")
    os.environ["TOSCA_PATH"] = request.param
    yield rhs
    if old is None:
        os.environ.pop("TOSCA_PATH", None)
    else:
        os.environ["TOSCA_PATH"] = old


@pytest.mark.parametrize("name,flow,model,scheme", CASES, ids=[c[0] for c in CASES])
def test_step_on_reference_mesh(rhs_mod, tmp_path, name, flow, model, scheme):
    case = C.mesh_file_case(f"{MESHES}/{name}.xyz", flow=flow, les_model=model, div_scheme=scheme)
    lay = case.lay
    print(f"\nstep on {name}.xyz: {lay} periodic={case.periodic} bc_u={case.params['bc_u']}", flush=True)
    nm = np_mesh(case.mesh)
    gl = (lay.mz, lay.my, lay.mx, 3)
    ref = dict(rhs=np.zeros(gl), cs=np.zeros(lay.gshape), nut=np.zeros(lay.gshape), btheta=np.zeros(gl))
    O.rhs_les_step(lay.info(), nm, case.params, to_np(case.ucont), to_np(case.ucat), to_np(case.theta), ref["rhs"],
                   1.0, ref["cs"], ref["nut"], ref["btheta"], 1.0)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    got = {k: to_dev(np.zeros_like(v)) for k, v in ref.items()}
    ctx.rhs_les_step(dm, case.params, to_dev(to_np(case.ucont)), to_dev(to_np(case.ucat)), to_dev(to_np(case.theta)),
                     got["rhs"], 1.0, got["cs"], got["nut"], got["btheta"], 1.0)
    torch.cuda.synchronize()
    for k in ("rhs", "cs", "nut", "btheta"):
        e = relerr(to_np(got[k]), ref[k])
        print(f"  {k:7s} relL2={e:.2e}", flush=True)
        assert e <= 1e-12, k
    # restart files of the outputs (global arrays), read back unchanged
    own = lambda a: a[3:3 + lay.mz, 3:3 + lay.my, 3:3 + lay.mx]
    fields = {"nut": own(to_np(got["nut"])), "cs": own(to_np(got["cs"])), "U": own(to_np(case.ucat))}
    io.write_fields(tmp_path / "0", fields)
    back = io.read_fields(tmp_path / "0", lay)
    for k, v in fields.items():
        assert np.array_equal(back[k], v), k
    ctx.close()
