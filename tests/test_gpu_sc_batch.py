"""Batched Shan-Chen lattices (taichi_lbm.batch.ShanChenBatch, one launch per
step over the tiles of every lattice) against each lattice run alone with
its own Stepper: fields, psi, forces and the pressure observable equal bit
for bit over runs of several lengths with a readback of one lattice in
between; and one lattice of the batch against the FP32 oracle.

The lattices differ in size, collision model, pseudopotential (SC_psi and
the Carnahan-Starling EOS), walls with a fixed psi, and (two components)
adhesion at solids.
"""
import contextlib
import io

import numpy as np
import pytest

from tolerances import check

pytestmark = pytest.mark.gpu

ONE = ["shan_chen_c1_32x32", "shan_chen_c1_wall_psi_40x32", "shan_chen_c1_cs_eos_32x32"]
TWO = ["shan_chen_c2_32x24", "shan_chen_c2_adhesion_40x32"]
FIELDS = ("f", "f2", "rho", "vel", "total_rho", "SCforce", "pressure")


def _cases(names):
    from golden.make_golden import build
    from taichi_lbm import cases
    out = []
    for name in names:
        case, _ = build(name)
        with contextlib.redirect_stdout(io.StringIO()):
            cases.initialise(case)
        out.append(case)
    return out


def _psi(case):
    return case["lbf"].sc_field.psi_field.to_numpy()


@pytest.mark.parametrize("names", [ONE, TWO], ids=["one-component", "two-component"])
def test_batch_equals_each_lattice_alone(names):
    from taichi_lbm import cases
    from taichi_lbm.batch import ShanChenBatch
    print(f"Shan-Chen batch of {names}", flush=True)
    batched, alone = _cases(names), _cases(names)
    sb = ShanChenBatch([cases.stepper(c) for c in batched])
    solo = [cases.stepper(c) for c in alone]
    for n, read in ((7, True), (6, False), (1, False), (4, False)):
        sb.run(n)
        for st in solo:
            st.run(n)
        if read:
            _ = batched[1]["lbf"].vel.to_numpy()
            _ = alone[1]["lbf"].vel.to_numpy()
    print(f"  tiles per launch: {sb.tiles()}", flush=True)
    for name, a, b in zip(names, batched, alone):
        for field in FIELDS:
            x, y = getattr(a["lbf"], field).to_numpy(), getattr(b["lbf"], field).to_numpy()
            assert np.array_equal(x, y), (name, field)
        assert np.array_equal(_psi(a), _psi(b)), (name, "psi")
        assert np.isfinite(a["lbf"].vel.to_numpy()).all()
    sb.close()


def test_batch_lattice_against_oracle():
    from oracle_bridge import oracle_solver
    from taichi_lbm import cases
    from taichi_lbm.batch import ShanChenBatch
    names = ["shan_chen_c2_adhesion_40x32", "shan_chen_c2_32x24"]
    print(f"Shan-Chen batch against the oracle: {names[0]}", flush=True)
    cs = _cases(names)
    sb = ShanChenBatch([cases.stepper(c) for c in cs])
    sb.run(20)
    sol = oracle_solver(cs[0], np.float32)
    sol.run(20)
    for field in ("f", "rho", "vel", "SCforce"):
        check(field, getattr(cs[0]["lbf"], field).to_numpy(), getattr(sol.lat, field))
    sb.close()


def test_batch_refuses_mixed_components():
    from taichi_lbm import cases
    from taichi_lbm.batch import ShanChenBatch
    a, b = _cases([ONE[0], TWO[0]])
    with pytest.raises(ValueError):
        ShanChenBatch([cases.stepper(a), cases.stepper(b)])
