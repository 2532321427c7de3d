"""The drop-in i-PI client (TMIPIinterface.py:7-61) against a local fake
i-PI server, on the CPU: protocol replies, units (positions bohr -> A in,
energy Ha and forces Ha/bohr out) and messages split across many reads."""
import numpy as np

from tensormol_amd.ipi import TMIPIManger
from tensormol_amd.params import BOHRPERA, JOULEPERHARTREE
from tests.ipi_fake import FakeIPIServer
from tests.test_drivers import X0, distorted, enandforce, spring_energy_grad


def test_client_protocol_and_units():
    frames = [[distorted(20 + k)] for k in range(3)]
    srv = FakeIPIServer(frames, np.diag([9.3215] * 3)).start()
    seen = []

    def ff(x):
        seen.append(x.copy())
        return enandforce(x)
    cl = TMIPIManger(ff, "127.0.0.1", srv.port)
    cl.md_run()
    srv.join()
    assert srv.replies == ["READY", "HAVEDATA"] * 3
    for k, fr in enumerate(frames):
        np.testing.assert_allclose(seen[k], fr[0], rtol=0, atol=1e-12)   # bohr -> A round trip
        r = srv.results[k][0]
        e, g = spring_energy_grad(fr[0])
        assert r["head"] == "FORCEREADY" and r["extra"] == b"nothing"
        assert abs(r["energy"] - e) < 1e-14
        np.testing.assert_allclose(r["forces"], -g / BOHRPERA, rtol=1e-12, atol=1e-15)
        assert not r["virial"].any()
    np.testing.assert_allclose(cl.cell, np.diag([9.3215] * 3), rtol=1e-12)
