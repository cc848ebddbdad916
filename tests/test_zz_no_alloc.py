"""tosca_formu_noalloc (FormU with its scratch in the library's fixed per-device array) against
tosca_formu (the context's workspace), through the C ABI: identical results, bit for bit, on two
calls in a row, the second on a smaller layout (so it runs on scratch the first call left
behind), for every form mode and both convective schemes; and the rejection of a layout whose
scratch would exceed the array."""
import ctypes as C

import numpy as np
import pytest
import torch

from tosca_amd import abi
from tosca_amd import case as C_
from tests._util import np_mesh, to_dev, random_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rhs_mod():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    return rhs


def _inputs(n, periodic, seed, scheme):
    lay, mesh, f, params = random_case(n=n, periodic=periodic, seed=seed, div_scheme=scheme,
                                       bcu_nonper=(C_.BCU_GENERIC, C_.BCU_SLIP, C_.BCU_WALLFUNCTION))
    dm = {k: to_dev(v) for k, v in np_mesh(mesh).items()}
    x = {k: f[k].to("cuda").contiguous() for k in ("ucont", "ucat", "nut", "tau_i", "tau_j")}
    pg = dict(params, tau_wall_i=x["tau_i"].data_ptr(), tau_wall_j=x["tau_j"].data_ptr())
    rng = np.random.Generator(np.random.PCG64(500 + seed))
    rhs0 = to_dev(rng.standard_normal((lay.mz, lay.my, lay.mx, 3)))
    return lay, dm, x, pg, rhs0


# the first layout spans several FormU tiles along i and j (62 x 7 cells each); the second is smaller
LAYOUTS = [((40, 70, 20), (2, 0, 1), 11), ((9, 12, 7), (0, 1, 0), 12)]


@pytest.mark.parametrize("scheme", [C_.DIV_CENTRAL, C_.DIV_QUICK])
@pytest.mark.parametrize("form_mode", [0, 1, 2])
def test_noalloc_matches_formu(rhs_mod, scheme, form_mode):
    for n, periodic, seed in LAYOUTS:
        lay, dm, x, pg, rhs0 = _inputs(n, periodic, seed, scheme)
        print(f"\nFormU noalloc vs FormU: cells {n} periodic {periodic} scheme {scheme} form_mode {form_mode}",
              flush=True)
        ctx = rhs_mod.RhsContext(lay.info())
        want, got = rhs0.clone(), rhs0.clone()
        ctx.FormU(dm, pg, x["ucont"], x["ucat"], x["nut"], want, 0.7, form_mode)
        ctx.FormU_noalloc(dm, pg, x["ucont"], x["ucat"], x["nut"], got, 0.7, form_mode)
        torch.cuda.synchronize()
        diff = float((got - want).abs().max())
        print(f"  max |noalloc - formu| = {diff:.3e}", flush=True)
        assert torch.equal(got, want)
        assert not torch.equal(got, rhs0)
        ctx.close()


def test_noalloc_rejects_oversized_layout(rhs_mod):
    # 998 x 298 x 598 interior cells: FormU would need about 68 M doubles of scratch
    info = dict(xs=0, ys=0, zs=0, xm=1000, ym=300, zm=600, gxs=-3, gys=-3, gzs=-3, gxm=1006, gym=306, gzm=606,
                mx=1000, my=300, mz=600)
    ctx = rhs_mod.RhsContext(info)
    params = abi.make_params(C_.default_params((0, 0, 0), (C_.BC_SLIP,) * 6, (C_.BC_ZEROGRAD,) * 6))
    mesh = abi.Mesh()
    rc = ctx.lib.tosca_formu_noalloc(ctx.h, C.byref(mesh), C.byref(params), None, None, None, None, 1.0, 0)
    msg = ctx.lib.tosca_last_error(ctx.h).decode()
    print(f"\n  oversized layout: rc={rc} ({msg})", flush=True)
    assert rc == abi.TOSCA_ERR_LAYOUT
    assert "scratch" in msg
    ctx.close()
