"""FormU's face dot with the cell-centred csi, eta, zet formed from the face bases icsi, jeta, kzet
(0.5 f + 0.5 f' of the cell's two faces, how SetMeshMetrics forms them, mesh.c:925-960) on meshes
tosca_mesh_check found to satisfy that identity bit for bit, and loaded on any other mesh.

On meshes of two i tiles and three j tiles of the fused kernel, for three periodicities:
  - a builder mesh (the identity holds) and the same mesh with one interior cell's csi, eta and zet
    scaled by 1 + 1e-6 (it fails; the bases are loaded) both match the oracle on their own mesh to
    the FP64 bar, and their Rhs differ;
  - which path ran: after the check, the same perturbation written into the checked arrays in place
    leaves the formed-basis Rhs unchanged bit for bit (the check does not see later writes, as
    documented), and a new check of the perturbed arrays gives the loaded-basis Rhs bit for bit.
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, random_case, out_arrays

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def rhs_mod():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    return rhs


@pytest.mark.parametrize("periodic", [(0, 0, 0), (2, 0, 1), (1, 1, 1)])
def test_formu_cell_bases_formed_or_loaded(rhs_mod, periodic):
    seed = 17 + sum(periodic)
    lay, mesh, f, params = random_case(n=(6, 70, 20), periodic=periodic, seed=seed, div_scheme=C.DIV_QUICK,
                                       bcu_nonper=(C.BCU_GENERIC, C.BCU_SLIP))
    print(f"\nFormU cell bases periodic={periodic} mx={lay.mx} my={lay.my} mz={lay.mz}", flush=True)
    nm = np_mesh(mesh)
    # one interior cell away from every boundary: its three cell bases scaled
    kc, jc, ic = 3 + lay.mz // 2, 3 + lay.my // 2, 3 + lay.mx // 2
    pert = {k: v.copy() for k, v in nm.items()}
    for n in ("csi", "eta", "zet"):
        pert[n][kc, jc, ic] *= 1.0 + 1e-6
    tau_i, tau_j = to_np(f["tau_i"]), to_np(f["tau_j"])
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    fields = [to_np(f[n]) for n in ("ucont", "ucat", "nut")]
    rhs0 = out_arrays(lay)["rhs"]
    ref = {}
    for name, m in (("builder", nm), ("perturbed", pert)):
        r = rhs0.copy()
        O.formu(lay.info(), m, po, *fields, r, 0.7, 0)
        ref[name] = r
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    dev_fields = [to_dev(a) for a in fields]

    def run(ctx, dm):
        r = to_dev(rhs0)
        ctx.FormU(dm, pg, *dev_fields, r, 0.7, 0)
        torch.cuda.synchronize()
        return to_np(r)

    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    formed = run(ctx, dm)
    ctx2 = rhs_mod.RhsContext(lay.info())
    loaded = run(ctx2, {k: to_dev(v) for k, v in pert.items()})
    e1, e2 = relerr(formed, ref["builder"]), relerr(loaded, ref["perturbed"])
    print(f"  formed relL2={e1:.2e}  loaded (perturbed mesh) relL2={e2:.2e}", flush=True)
    assert e1 <= TOL and e2 <= TOL
    assert not np.array_equal(formed, loaded)
    # the perturbation written into the checked arrays: the formed bases ignore it
    for n in ("csi", "eta", "zet"):
        dm[n].copy_(to_dev(pert[n]))
    assert np.array_equal(run(ctx, dm), formed)
    # checked again: the identity fails, the bases are loaded
    ctx.mesh_check(dm)
    assert np.array_equal(run(ctx, dm), loaded)
    ctx.close()
    ctx2.close()
