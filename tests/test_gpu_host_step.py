"""The staged residency mode (tosca_rhs_les_step_host): flow fields in host memory, copied through
the context's device buffers around the fused step.  Its outputs must equal the device-resident
step's bit for bit, for pinned and for pageable host memory, and nothing must be left un-copied
(ghost entries the step does not write keep their host values)."""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from tests._util import np_mesh, to_np, to_dev, random_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pinned", [True, False])
@pytest.mark.parametrize("periodic", [(2, 0, 1), (0, 0, 0)])
def test_host_step_equals_device_step(periodic, pinned):
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    lay, mesh, f, params = random_case(n=(9, 8, 7), periodic=periodic, seed=11, bcu_nonper=(C.BCU_GENERIC, C.BCU_SLIP))
    print(f"\nhost-staged step periodic={periodic} pinned={pinned}", flush=True)
    rng = np.random.Generator(np.random.PCG64(5))
    gl = (lay.mz, lay.my, lay.mx, 3)
    start = dict(rhs=rng.standard_normal(gl), cs=rng.standard_normal(lay.gshape), nut=rng.standard_normal(lay.gshape),
                 btheta=rng.standard_normal(gl))
    ctx = rhs.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in np_mesh(mesh).items()}
    dev = {k: to_dev(v) for k, v in start.items()}
    ctx.rhs_les_step(dm, params, to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(to_np(f["theta"])),
                     dev["rhs"], 0.7, dev["cs"], dev["nut"], dev["btheta"], 1.3)
    torch.cuda.synchronize()

    def host(a):
        t = torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64)
        return t.pin_memory() if pinned else t.clone()
    h = {k: host(v) for k, v in start.items()}
    copy_ms, step_ms = ctx.rhs_les_step_host(dm, params, host(to_np(f["ucont"])), host(to_np(f["ucat"])),
                                             host(to_np(f["theta"])), h["rhs"], 0.7, h["cs"], h["nut"], h["btheta"], 1.3)
    print(f"  copies {copy_ms:.3f} ms, step {step_ms:.3f} ms", flush=True)
    for k in ("rhs", "cs", "nut", "btheta"):
        assert np.array_equal(h[k].numpy(), to_np(dev[k])), k
    assert copy_ms > 0 and step_ms > 0
    ctx.close()
