"""Reader of the learned-closure network file (format "CFMLP001" of
include/channelflow.h), for the tests."""
import numpy as np


def read_cfmlp(path):
    with open(path, "rb") as f:
        raw = f.read()
    assert raw[:8] == b"CFMLP001", "not a CFMLP001 network file"
    hdr = np.frombuffer(raw, dtype="<i4", count=4, offset=8)
    assert list(hdr) == [11, 64, 6, 9478]
    vals = np.frombuffer(raw, dtype="<f4", offset=24).astype(np.float64)
    in_mu, in_sd = vals[0:11], vals[11:22]
    out_mu, out_sd = vals[22:28], vals[28:34]
    flat = vals[34:34 + 9478]
    assert flat.size == 9478
    return flat, in_mu, in_sd, out_mu, out_sd
