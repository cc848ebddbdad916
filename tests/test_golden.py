"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the CPU oracle).

CPU: the case builders still produce the same inputs (checksums) and the oracle still produces the
stored outputs (a regression lock on both).  GPU: the HIP library reproduces the stored outputs to
the parity bar (normwise relative L2 <= 1e-12 per field), through the C ABI, on both kernel paths.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

from tests._util import relerr, to_dev
from tests.golden.make_golden import CASES, build, checksums, run_oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
FIELDS = ("rhs", "cs", "nut", "btheta")


def _load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_oracle(name):
    g = _load(name)
    meta = json.loads(str(g["meta"]))
    assert CASES[name][0] == meta["builder"]
    case = build(name)
    ref = json.loads(str(g["checksums"]))
    for k, (s, n) in checksums(case).items():
        assert abs(s - ref[k][0]) <= 1e-12 * max(1.0, abs(ref[k][0])) and abs(n - ref[k][1]) <= 1e-12 * ref[k][1], k
    out = run_oracle(case)
    for k in FIELDS:
        assert relerr(out[k], g[k]) <= 1e-14, k


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["fused", "ref"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_gpu(name, path, monkeypatch):
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    monkeypatch.setenv("TOSCA_PATH", path)
    from tosca_amd import rhs
    g = _load(name)
    case = build(name)
    lay = case.lay
    ctx = rhs.RhsContext(lay.info())
    mesh = {k: to_dev(v.numpy()) for k, v in case.mesh.items() if k != "cent"}
    out = dict(rhs=torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda"),
               cs=torch.zeros(lay.gshape, dtype=torch.float64, device="cuda"),
               nut=torch.zeros(lay.gshape, dtype=torch.float64, device="cuda"),
               btheta=torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda"))
    ctx.rhs_les_step(mesh, case.params, to_dev(case.ucont.numpy()), to_dev(case.ucat.numpy()),
                     to_dev(case.theta.numpy()), out["rhs"], 1.0, out["cs"], out["nut"], out["btheta"], 1.0)
    torch.cuda.synchronize()
    print(f"\ngolden {name} [{path}]", flush=True)
    for k in FIELDS:
        e = relerr(out[k].cpu().numpy(), g[k])
        print(f"  {k:7s} relL2={e:.2e}", flush=True)
        assert e <= 1e-12, k
