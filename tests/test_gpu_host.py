"""End to end through the idash-compatible executables on the GPU:
keygen -> encrypt -> cloud -> decrypt (eval/keygen.cpp, encrypt.cpp,
cloud.cpp, decrypt.cpp) in a scratch directory, each file checked against
the oracle:
  encrypted_data.bin        phase(b - a*s) = one-hot message + noise, |noise| small
  encrypted_prediction.bin  bit-exact against the oracle's cloud step on the
                            ciphertexts and model just written
  result.csv                byte-exact against the oracle's decryption, printed
                            with the default ostream float format (%g)
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from idash2019_2_amd import formats, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# IHE_HOST_DIR runs the same cases on another build of the executables
# (make -C idash2019_2_amd/host asan: AddressSanitizer + UBSan)
HOST = os.path.join(ROOT, os.environ.get("IHE_HOST_DIR", os.path.join("idash2019_2_amd", "host")))
# the sanitizer build's leak check, less the objects the API leaves to its callers
LSAN = "suppressions=" + os.path.join(ROOT, "idash2019_2_amd", "host", "lsan.supp")
ONE, NAN = 1 << 18, (131072, 87381, 43690)


def run(exe, *args, cwd):
    r = subprocess.run([os.path.join(HOST, exe), *args], cwd=cwd, capture_output=True, text=True,
                       timeout=300, env=dict(os.environ, IHE_SEED="9", LSAN_OPTIONS=LSAN))
    assert r.returncode == 0, f"{exe}: {r.stdout}{r.stderr}"
    return r.stdout


def expected_messages(geno, ns, R, RS, num_cts):
    """b-polynomial messages of each input ciphertext (setScore, eval/idash.cpp:642-679)."""
    m = np.zeros((num_cts, 1024), np.int64)
    tags = geno.shape[0]
    for q in range(tags):
        for v in range(3):
            bidx = 3 * q + v
            f, r = divmod(bidx, R)
            g = geno[q].astype(np.int64)
            val = np.where(g == v, ONE, 0)
            val = np.where(g < 0, NAN[v], val)
            m[f, r * RS:r * RS + ns] += val
    return m


@pytest.mark.parametrize("ns,nan", [(128, 0.05), (1024, 0.0), (100, 0.02)])
def test_cli_pipeline_matches_oracle(tmp_path, ns, nan):
    c = synth.generate(60, 150, ns, 5, seed=ns)
    files = synth.write_reference_files(c, str(tmp_path), nan_fraction=nan, seed=1)
    d = str(tmp_path)
    run("keygen", "targets.txt", "tags.txt", "1", cwd=d)
    run("encrypt", "tags.txt", cwd=d)
    run("cloud", "model", cwd=d)
    first = open(os.path.join(d, "encrypted_prediction.bin"), "rb").read()
    assert os.path.exists(os.path.join(d, "model", ".ihe_model_cache"))
    run("cloud", "model", cwd=d)                 # the second run reads the model cache
    assert open(os.path.join(d, "encrypted_prediction.bin"), "rb").read() == first
    out = run("decrypt", cwd=d)
    assert "decrypt wall time" in out
    p, key, _ = formats.read_key(os.path.join(d, "keys.bin"))
    NS, R, RS = p["NUM_SAMPLES"], p["NUM_REGIONS"], p["REGION_SIZE"]
    assert (NS, R, RS) == c.layout

    # encrypted data: one record per FeatIndex of the tags, phase = message + noise
    num_cts = (3 * c.num_tags + R - 1) // R
    idx, ct, var = formats.read_records(os.path.join(d, "encrypted_data.bin"), num_cts)
    assert sorted(idx.tolist()) == list(range(num_cts))
    assert (var == 2.0 ** -50).all()
    msg = expected_messages(files["geno"], NS, R, RS, num_cts)
    for f in range(num_cts):
        noise = oracle.phase(ct[f], key).astype(np.int64) - msg[f]
        noise = (noise + (1 << 31)) % (1 << 32) - (1 << 31)
        assert np.abs(noise).max() < 1200, f
    assert np.abs(ct[:, 0, :].astype(np.float64)).mean() > 2.0 ** 29    # a is uniform

    # predictions: bit-exact against the oracle's cloud on the same inputs
    pidx, pred, pvar = formats.read_records(os.path.join(d, "encrypted_prediction.bin"),
                                            c.num_outputs)
    assert sorted(pidx.tolist()) == list(range(c.num_outputs))
    rp, ib, w = c.model_csr()
    ref, rv = oracle.cloud(ct, var, rp, ib, w, NS, R, RS)
    assert (pred == ref).all()
    assert (pvar == rv).all()

    # result.csv: header, then for each sample each target in file order
    dec = oracle.decrypt(ref, key, NS)
    lines = open(os.path.join(d, "result.csv")).read().split("\n")
    assert lines[0] == "Subject ID,target SNP,0,1,2" and lines[-1] == ""
    body = lines[1:-1]
    assert len(body) == NS * c.num_targets
    names = [str(x) for x in c.target_pos.tolist()]
    for k in range(0, len(body), max(1, len(body) // 500)):
        s, t = divmod(k, c.num_targets)
        want = "%d,%s,%g,%g,%g" % (s, names[t], dec[3 * t, s], dec[3 * t + 1, s], dec[3 * t + 2, s])
        assert body[k] == want, (k, body[k], want)

    # the decrypted scores approximate the plaintext model (compute_score,
    # eval/idash.cpp:888-927) within the encryption noise: 6 sigma of 128 torus
    # units times sqrt(sum w^2), i.e. well below 2e-3 here
    for o in range(0, c.num_outputs, 7):
        plain = oracle.plaintext_score(files["geno"], ib[rp[o]:rp[o + 1]], w[rp[o]:rp[o + 1]], NS)
        assert np.abs(dec[o] - plain).max() < 2e-3, o


def test_cli_decrypt_by_position(tmp_path):
    c = synth.generate(30, 40, 64, 3, seed=5)
    synth.write_reference_files(c, str(tmp_path))
    d = str(tmp_path)
    run("keygen", "targets.txt", "tags.txt", "1", cwd=d)
    run("encrypt", "tags.txt", cwd=d)
    run("cloud", "model", cwd=d)
    run("decrypt", "bypos", cwd=d)
    lines = open(os.path.join(d, "result_bypos.csv")).read().split("\n")
    assert lines[0] == "Subject ID,target SNP,0,1,2"
    assert lines[1].split(",")[1] == str(int(c.target_pos[0]))
    assert len(lines) == 2 + 64 * 40


def test_cloud_passes_the_framing_word_through(tmp_path):
    """Records read with a framing word other than this build's 0 (a file from
    another TLWE implementation) keep it: the cloud step writes the word its
    inputs carried on every prediction record, and the predictions decrypt as
    they do with the word 0."""
    c = synth.generate(40, 80, 128, 5, seed=21)
    synth.write_reference_files(c, str(tmp_path))
    d = str(tmp_path)
    run("keygen", "targets.txt", "tags.txt", "1", cwd=d)
    run("encrypt", "tags.txt", cwd=d)
    run("cloud", "model", cwd=d)
    run("decrypt", cwd=d)
    plain = open(os.path.join(d, "result.csv")).read()
    pred0 = formats.read_records(os.path.join(d, "encrypted_prediction.bin"))
    assert (formats.read_stream_words(os.path.join(d, "encrypted_prediction.bin")) == 0).all()
    path = os.path.join(d, "encrypted_data.bin")
    idx, ct, var = formats.read_records(path)
    formats.write_records(path, idx, ct, var, stream_word=-0x5A17F00D)
    run("cloud", "model", cwd=d)
    words = formats.read_stream_words(os.path.join(d, "encrypted_prediction.bin"))
    assert words.size == c.num_outputs and (words == -0x5A17F00D).all()
    pred1 = formats.read_records(os.path.join(d, "encrypted_prediction.bin"))
    assert (pred0[1] == pred1[1]).all() and (pred0[2] == pred1[2]).all()
    run("decrypt", cwd=d)
    assert open(os.path.join(d, "result.csv")).read() == plain
