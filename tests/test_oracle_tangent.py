"""The oracle's tangent (oracle/tosca_oracle_tangent.cpp) against the double oracle.

The tangent is the FP64 reference the hand-written adjoint of the library is checked against, so
it is pinned here on the CPU: its value channel must reproduce the double oracle bit for bit, and
its tangent must match central finite differences of the double oracle in every input block
(ucont, ucat, theta, Cs), on random curvilinear meshes with every boundary kind.
"""
import numpy as np
import pytest

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, relerr, random_case


def _inputs(periodic, seed, scheme):
    lay, mesh, f, params = random_case(n=(7, 6, 8), periodic=periodic, seed=seed, div_scheme=scheme,
                                       bcu_nonper=(C.BCU_GENERIC, C.BCU_SLIP, C.BCU_WALLFUNCTION))
    rng = np.random.Generator(np.random.PCG64(50 + seed))
    cs = rng.uniform(0.01, 0.2, lay.gshape)
    x = dict(ucont=to_np(f["ucont"]), ucat=to_np(f["ucat"]), theta=to_np(f["theta"]), cs=cs)
    params = dict(params, tau_wall_i=to_np(f["tau_i"]).ctypes.data, tau_wall_j=to_np(f["tau_j"]).ctypes.data)
    keep = (f["tau_i"], f["tau_j"])
    return lay, np_mesh(mesh), params, x, keep


@pytest.mark.parametrize("periodic", [(0, 0, 0), (2, 0, 1), (1, 1, 1), (2, 2, 2)])
@pytest.mark.parametrize("scheme", [C.DIV_CENTRAL, C.DIV_QUICK])
def test_tangent_matches_finite_differences(periodic, scheme):
    lay, nm, params, x, keep = _inputs(periodic, 3, scheme)
    tau_i, tau_j = (to_np(t) for t in keep)
    params = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    info = lay.info()
    R0 = O.rhs_forward(info, nm, params, x["ucont"], x["ucat"], x["theta"], x["cs"], 0.7, 1.3)
    rng = np.random.Generator(np.random.PCG64(8))
    for name in ("ucont", "ucat", "theta", "cs"):
        d = rng.standard_normal(x[name].shape)
        kw = {"d_" + name: d}
        R, dR = O.rhs_jvp(info, nm, params, x["ucont"], x["ucat"], x["theta"], x["cs"], scale=0.7, buoy_scale=1.3, **kw)
        assert np.array_equal(R, R0)                     # value channel = double oracle, bit for bit
        eps = 1e-6 * max(1.0, float(np.abs(x[name]).max()))
        xp, xm = dict(x), dict(x)
        xp[name] = x[name] + eps * d
        xm[name] = x[name] - eps * d
        Rp = O.rhs_forward(info, nm, params, xp["ucont"], xp["ucat"], xp["theta"], xp["cs"], 0.7, 1.3)
        Rm = O.rhs_forward(info, nm, params, xm["ucont"], xm["ucat"], xm["theta"], xm["cs"], 0.7, 1.3)
        fd = (Rp - Rm) / (2 * eps)
        err = relerr(dR, fd)
        print(f"  {name:6s} tangent vs FD relL2={err:.2e}")
        # QUICK switches branch where ucont changes sign; a +-eps step across zero is rare but allowed
        assert err <= (1e-6 if name != "theta" else 1e-5), (name, err)
        assert np.abs(dR).max() > 0
