"""GPU: gauge fields read from and written to ILDG/LIME files through device
fields, against the host reader/writer and the oracle's plaquette."""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu


def test_device_lime_round_trip(q, tmp_path):
    L = (8, 8, 8, 16)
    U = q.gen_gauge(L, 0.4, 1)
    host_file = tmp_path / "host.lime"
    q.write_lime(host_file, L, U, "config 1")
    print("lime round trip", flush=True)
    u = q.Gauge(L, boundaryT=-1)
    u.read_lime(host_file)
    assert np.array_equal(u.download(), U)
    assert abs(u.calculatePlaquette() - orc.plaquette(U, L)) <= 1e-13
    dev_file = tmp_path / "dev.lime"
    u.write_lime(dev_file, "config 1")
    assert dev_file.read_bytes() == host_file.read_bytes()
    with pytest.raises(q.QkxError):
        q.Gauge((8, 8, 8, 8), boundaryT=-1).read_lime(host_file)
