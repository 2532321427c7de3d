"""IPIBatchDriver: four i-PI client connections (the beads of a path-integral
run) served with one packed engine launch per step, against the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H
from tests.ipi_fake import FakeIPIServer

pytestmark = pytest.mark.gpu


def test_batched_beads_match_oracle(cuda):
    from tensormol_amd.engine import BPEEEngine
    from tensormol_amd.ipi import IPIBatchDriver
    from tensormol_amd.params import BOHRPERA
    Z, x = H.system("water10")[:2]
    rng = np.random.default_rng(4)
    frames = [[x + rng.normal(0, 0.02, x.shape) for _ in range(4)] for _ in range(3)]
    srv = FakeIPIServer(frames, np.diag([9.3215] * 3), nconn=4).start()
    model = H.model_for("water")
    eng = BPEEEngine(model, precision="fp32")
    drv = IPIBatchDriver(eng, Z, [("127.0.0.1", srv.port)] * 4, wait=2.0)
    drv.run()
    srv.join()
    assert drv.evaluated == 12 and drv.launches == 3      # one launch per step for all beads
    sp = O.SymParams(model.eles)
    for k, fr in enumerate(frames):
        for c, pos in enumerate(fr):
            ref = O.evaluate(pos, Z, np.array([0, len(Z)]), sp, model.energy_nets,
                             model.charge_nets, alpha=model.alpha, dsf_alpha=model.dsf_alpha)
            r = srv.results[k][c]
            assert abs(r["energy"] - ref["Etotal"][0]) < 2e-6 * len(Z)
            np.testing.assert_allclose(r["forces"], -ref["gradient"] / BOHRPERA, atol=2e-5)
