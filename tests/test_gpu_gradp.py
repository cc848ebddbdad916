"""GPU parity of GradP (tosca_grad_p, src/peqn.c:3437-3773) against the CPU oracle's restatement on
seeded random curvilinear meshes with every periodic type per axis.  The output vector starts from
random values, so the zeros GradP must write everywhere else are compared too.  Both sides evaluate
the same expressions without FP contraction: bar 1e-14 relative L2 (bit-identical in practice)."""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, random_nodes

pytestmark = pytest.mark.gpu

PERS = [(0, 0, 0), (1, 1, 1), (2, 2, 2), (2, 0, 1), (0, 1, 2), (1, 2, 0)]


@pytest.fixture(scope="module")
def rhs_mod():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    return rhs


@pytest.mark.parametrize("n", [(7, 6, 5), (40, 33, 21)])
@pytest.mark.parametrize("periodic", PERS)
def test_grad_p_parity(rhs_mod, periodic, n):
    seed = PERS.index(periodic) + (0 if n[0] == 7 else 10)
    lay, mesh = C.mesh_metrics(random_nodes(n, seed), periodic)
    nm = np_mesh(mesh)
    rng = np.random.Generator(np.random.PCG64(600 + seed))
    p = rng.standard_normal(lay.gshape)
    bcs = [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)]
    params = C.default_params(periodic, bcs, bcs, les_active=0)
    start = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    print(f"\nGradP {n} periodic={periodic}", flush=True)
    dp_o = start.copy()
    O.grad_p(lay.info(), nm, params, p, dp_o)
    ctx = rhs_mod.RhsContext(lay.info())
    dp_g = to_dev(start)
    ctx.GradP({k: to_dev(v) for k, v in nm.items()}, params, to_dev(p), dp_g)
    torch.cuda.synchronize()
    g = to_np(dp_g)
    e = relerr(g, dp_o)
    print(f"  dp relL2={e:.2e} identical={np.array_equal(g, dp_o)}", flush=True)
    assert e <= 1e-14
    ctx.close()
