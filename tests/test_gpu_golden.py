"""The device path against the committed golden fixtures (tests/golden/*.npz):
encrypt on the device with the fixture's seed, the cloud step on a band built
from the fixture's model, and decryption, each compared bit for bit with the
stored known answers (no oracle at run time). Cohorts at R = 8, 1, 64 and 3
(eval/idash.cpp:409-410), with missing genotypes and dropped zero weights."""
import glob
import hashlib
import os

import numpy as np
import pytest

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_device_path_reproduces_fixture(engine, path):
    f = np.load(path, allow_pickle=False)
    ns = int(f["ns"])
    C, num_out = f["enc_var"].size, f["row_ptr"].size - 1
    key = engine.key(f["key"])
    d_geno = engine.to_device(np.ascontiguousarray(f["geno"]))
    d_ct = engine.empty((C, 2, 1024), np.int32)
    d_var = engine.empty((C,), np.float64)
    assert engine.encrypt(key, ns, d_geno, int(f["enc_seed"]), d_ct, d_var) == C
    ct = d_ct.numpy()
    assert sha(ct) == str(f["enc_sha"])
    assert (d_var.numpy() == f["enc_var"]).all()
    band = engine.band(ns, C, f["row_ptr"], f["in_bidx"], f["weight"])
    d_out = engine.empty((num_out, 2, 1024), np.int32)
    d_ov = engine.empty((num_out,), np.float64)
    engine.cloud(band, d_ct, d_var, d_out, d_ov)
    assert sha(d_out.numpy()) == str(f["cloud_sha"])
    assert (d_ov.numpy() == f["cloud_var"]).all()
    d_ph = engine.empty((num_out, ns), np.float32)
    engine.decrypt(key, ns, d_out, num_out, d_ph)
    ph = d_ph.numpy()
    assert sha(ph) == str(f["phase_sha"])
    band.free()
    key.free()
