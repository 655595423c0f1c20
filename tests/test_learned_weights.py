"""The committed warm-start weights (CPU): loadable without pickle, the
parameter count of the C ABI, finite, unused channels zero."""
import json
import os

import numpy as np

from cuda_sto_pf_amd import learned

W = learned.DEFAULT_WEIGHTS


def test_weights_file():
    assert os.path.exists(W)
    d = np.load(W, allow_pickle=False)
    p = d["params"]
    assert p.dtype == np.float32 and p.size == 34640 and np.isfinite(p).all()
    per = learned.P_PER_LAYER
    W0 = p[:27 * 256].reshape(27, 16, 16)
    assert np.all(W0[:, 7:, :] == 0)          # input channels beyond phi + 6 eigenstress
    W4 = p[4 * per:4 * per + 27 * 256].reshape(27, 16, 16)
    assert np.all(W4[:, :, 3:] == 0)          # output channels beyond the 3-vector
    assert float(d["s_out"]) > 0 and float(d["s_tau"]) > 0
    meta = json.loads(str(d["meta"]))
    assert meta["snapshots"] == 4000 and meta["final_val_loss"] < 0.05 * meta["initial_val_loss"]
