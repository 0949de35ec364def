"""The idash-compatible C++ host on the CPU: keygen (eval/idash.cpp:315-434)
and the parameter/key files it writes (eval/idash.cpp:128-246, 476-502),
checked against the index rules of the reference; error behaviour
(REQUIRE_DRAMATICALLY: "ERROR: ..." on stdout, then abort)."""
import os
import subprocess

import numpy as np
import pytest

from idash2019_2_amd import formats, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# IHE_HOST_DIR runs the same cases on another build of the executables
# (make -C idash2019_2_amd/host asan: AddressSanitizer + UBSan)
HOST = os.path.join(ROOT, os.environ.get("IHE_HOST_DIR", os.path.join("idash2019_2_amd", "host")))
# the sanitizer build's leak check, less the objects the API leaves to its callers
LSAN = "suppressions=" + os.path.join(ROOT, "idash2019_2_amd", "host", "lsan.supp")


def run(exe, *args, cwd, env=None):
    path = os.path.join(HOST, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (run __graft_entry__.build())")
    e = dict(os.environ, IHE_SEED="5", LSAN_OPTIONS=LSAN, **(env or {}))
    return subprocess.run([path, *args], cwd=cwd, env=e, capture_output=True, text=True, timeout=120)


def test_keygen_writes_reference_layout(tmp_path):
    c = synth.generate(60, 200, 128, 5, seed=11)
    synth.write_reference_files(c, str(tmp_path))
    r = run("keygen", "targets.txt", "tags.txt", "1", cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    for line in ("NUM_SAMPLES: 128", "NUM_REGIONS: 8", "REGION_SIZE: 128",
                 "NUM_TARGET_POSITIONS: 200", "NUM_TAG_POSITIONS: 60"):
        assert line in r.stdout
    p = formats.read_params(str(tmp_path / "params.bin"))
    assert (p["NUM_SAMPLES"], p["NUM_REGIONS"], p["REGION_SIZE"]) == (128, 8, 128)
    assert p["NUM_INPUT_POSITIONS"] == 60 and p["NUM_OUTPUT_POSITIONS"] == 200
    assert p["NUM_INPUT_FEATURES"] == 3 * 60 + 1          # + the constant
    assert p["NUM_OUTPUT_FEATURES"] == 600
    # tag q (file order) owns big indices 3q..3q+2; target t owns 3t..3t+2
    for q, pos in enumerate(c.tag_pos.tolist()):
        assert p["in_features_index"][pos] == (3 * q, 3 * q + 1, 3 * q + 2)
    assert [pos for pos, _ in p["out_position_names"]] == c.target_pos.tolist()
    assert all(name == str(pos) for pos, name in p["out_position_names"])
    for t, pos in enumerate(c.target_pos.tolist()):
        assert p["out_features_index"][pos] == (3 * t, 3 * t + 1, 3 * t + 2)
    pk, key, tp = formats.read_key(str(tmp_path / "keys.bin"))
    assert pk["NUM_SAMPLES"] == 128 and pk["in_features_index"] == p["in_features_index"]
    assert tp == dict(N=1024, k=1, alpha_min=2.0 ** -25, alpha_max=0.25)
    assert set(np.unique(key)) <= {0, 1} and 400 < key.sum() < 624


def test_keygen_layout_for_odd_sample_counts(tmp_path):
    c = synth.generate(20, 30, 100, 3, seed=2)        # R = 10, RS = 102
    synth.write_reference_files(c, str(tmp_path))
    assert run("keygen", "targets.txt", "tags.txt", "1", cwd=tmp_path).returncode == 0
    p = formats.read_params(str(tmp_path / "params.bin"))
    assert (p["NUM_SAMPLES"], p["NUM_REGIONS"], p["REGION_SIZE"]) == (100, 10, 102)


def test_keygen_keeps_fifty_tags_past_the_last_target(tmp_path):
    # 30 tags, 5 targets, then 70 more tags: keygen reads until the 51st tag past
    # the last target (eval/idash.cpp:383-386)
    tags = [1000 + 10 * i for i in range(100)]
    targets = [1003 + 10 * i for i in range(0, 30, 6)]
    with open(tmp_path / "tags.txt", "w") as f:
        for p in tags:
            f.write(f"22 {p} {p + 1} rs{p} 0 1 2 NA\n")
    with open(tmp_path / "targets.txt", "w") as f:
        f.write("".join(f"{t}\n" for t in targets))
    assert run("keygen", "targets.txt", "tags.txt", "1", cwd=tmp_path).returncode == 0
    p = formats.read_params(str(tmp_path / "params.bin"))
    past = [t for t in tags if t > targets[-1]]
    assert p["NUM_INPUT_POSITIONS"] == len(tags) - len(past) + 50
    assert p["NUM_SAMPLES"] == 4 and p["NUM_REGIONS"] == 256 and p["REGION_SIZE"] == 4


def test_keygen_full_target_format_keeps_names(tmp_path):
    c = synth.generate(15, 10, 64, 3, seed=4)
    synth.write_reference_files(c, str(tmp_path))
    with open(tmp_path / "targets_full.txt", "w") as f:
        for pos in c.target_pos.tolist():
            f.write(f"22 {pos} {pos + 1} snp{pos} 0 0\n")
    assert run("keygen", "targets_full.txt", "tags.txt", "0", cwd=tmp_path).returncode == 0
    p = formats.read_params(str(tmp_path / "params.bin"))
    assert [n for _, n in p["out_position_names"]] == [f"snp{x}" for x in c.target_pos.tolist()]


def test_bad_tag_file_dies_dramatically(tmp_path):
    (tmp_path / "targets.txt").write_text("100\n")
    (tmp_path / "tags.txt").write_text("21 90 91 rs90 0 1\n")
    r = run("keygen", "targets.txt", "tags.txt", "1", cwd=tmp_path)
    assert r.returncode != 0 and "ERROR: file format error" in r.stdout


def test_missing_files_die_dramatically(tmp_path):
    r = run("cloud", str(tmp_path / "model"), cwd=tmp_path)
    assert r.returncode != 0 and "ERROR: Cannot open parameters file for read" in r.stdout
    r = run("decrypt", cwd=tmp_path)
    assert r.returncode != 0 and "ERROR: Cannot open key file for read" in r.stdout


def read_model_cache(path):
    import struct
    buf = open(path, "rb").read()
    assert buf[:8] == b"IHEMODL1"
    _, n = struct.unpack_from("<QQ", buf, 8)
    off, files = 24, {}
    for _ in range(n):
        pos, snp, bidx, size, mtime, cnt = struct.unpack_from("<QIIQqQ", buf, off)
        off += 40
        ent = np.frombuffer(buf, np.int32, 2 * cnt, off).reshape(cnt, 2)
        off += 8 * cnt
        files[(pos, snp)] = (bidx, {int(c) & 0xFFFFFFFF: int(w) for c, w in ent})
    return files


def test_model_cache_is_written_validated_and_refreshed(tmp_path):
    """read_model parses the .hr files once, keeps their entries (mapped to big
    indices) in <model dir>/.ihe_model_cache, and re-parses when a file changes."""
    import os
    import time
    c = synth.generate(20, 30, 64, 3, seed=9)
    synth.write_reference_files(c, str(tmp_path))
    assert run("keygen", "targets.txt", "tags.txt", "1", cwd=tmp_path).returncode == 0
    # no encrypted data yet: cloud reads params and the model, then stops
    r = run("cloud", "model", cwd=tmp_path)
    assert r.returncode != 0 and "Cannot open encrypted data file" in r.stdout
    cache = tmp_path / "model" / ".ihe_model_cache"
    files = read_model_cache(cache)
    p = formats.read_params(str(tmp_path / "params.bin"))
    assert len(files) == 3 * c.num_targets
    for t, pos in enumerate(c.target_pos.tolist()):
        for k in range(3):
            bidx, ent = files[(pos, k)]
            assert bidx == 3 * t + k
            lo = int(c.lo[t])
            want = {3 * (lo + j // 3) + j % 3: int(c.weights[3 * t + k, j]) for j in range(3 * c.neighbours)}
            want[0xFFFFFFFF] = int(c.bias[3 * t + k])
            assert ent == want
    # an edited file (new content, new mtime) is parsed again
    pos = c.target_pos.tolist()[4]
    hr = tmp_path / "model" / f"{pos}_1.hr"
    time.sleep(0.01)
    hr.write_text("".join(l if not l.startswith("Constant") else "Constant 4321\n"
                          for l in hr.read_text().splitlines(True)))
    os.utime(hr, ns=(time.time_ns(), time.time_ns()))
    run("cloud", "model", cwd=tmp_path)
    assert read_model_cache(cache)[(pos, 1)][1][0xFFFFFFFF] == 4321
    assert p["NUM_OUTPUT_FEATURES"] == 3 * c.num_targets
    # a params file that changes only the middle big index of two tags (the .hr
    # files untouched): the cache no longer matches, and the model is re-read
    # with the new index
    raw = bytearray((tmp_path / "params.bin").read_bytes())
    q0, q1 = int(c.lo[0]), int(c.lo[0]) + 1          # two tags of target 0's window
    tp = c.tag_pos.tolist()
    # entries are (uint64 pos, 3 x uint32) in the file's own order
    entry = {int.from_bytes(raw[28 + 20 * i:36 + 20 * i], "little"): 28 + 20 * i
             for i in range(c.num_tags)}
    off = lambda q, v: entry[tp[q]] + 8 + 4 * v
    b0 = raw[off(q0, 1):off(q0, 1) + 4]
    raw[off(q0, 1):off(q0, 1) + 4] = raw[off(q1, 1):off(q1, 1) + 4]
    raw[off(q1, 1):off(q1, 1) + 4] = b0
    (tmp_path / "params.bin").write_bytes(bytes(raw))
    p2 = formats.read_params(str(tmp_path / "params.bin"))
    assert p2["in_features_index"][tp[q0]][1] == 3 * q1 + 1
    run("cloud", "model", cwd=tmp_path)
    ent = read_model_cache(cache)[(c.target_pos.tolist()[0], 0)][1]
    j0, j1 = 3 * (q0 - int(c.lo[0])) + 1, 3 * (q1 - int(c.lo[0])) + 1
    assert ent[3 * q1 + 1] == int(c.weights[0, j0])  # tag q0's "_1" column now maps there
    assert ent[3 * q0 + 1] == int(c.weights[0, j1])


def write_inputs(tmp_path, c, keep, stream_word=0):
    """encrypted_data.bin with random ciphertexts at the FeatIndex values `keep`."""
    rng = np.random.default_rng(1)
    ct = rng.integers(-(1 << 31), 1 << 31, (c.num_in_cts, 2, 1024), dtype=np.int64).astype(np.int32)
    var = np.full(c.num_in_cts, 2.0 ** -50)
    formats.write_records(str(tmp_path / "encrypted_data.bin"), keep, ct, var, stream_word)


def test_cloud_dies_on_a_missing_input_record(tmp_path):
    """A model coefficient whose ciphertext the encrypted-data file lacks ends
    the cloud step with ERROR before any device work (getTLWE,
    eval/idash.h:162-166), whatever the file's length."""
    c = synth.generate(20, 30, 64, 3, seed=9)        # R = 16: 4 input ciphertexts
    synth.write_reference_files(c, str(tmp_path))
    assert run("keygen", "targets.txt", "tags.txt", "1", cwd=tmp_path).returncode == 0
    R = c.layout[1]
    used = sorted({(3 * int(c.lo[t]) + j) // R for t in range(c.num_targets)
                   for j in range(3 * c.neighbours)})
    drop = used[len(used) // 2]
    write_inputs(tmp_path, c, [i for i in range(c.num_in_cts) if i != drop])
    r = run("cloud", "model", cwd=tmp_path)
    assert r.returncode != 0 and "ERROR: missing input ciphertext" in r.stdout
    assert not (tmp_path / "encrypted_prediction.bin").exists()


def test_input_records_with_different_framing_words_are_rejected(tmp_path):
    """Every record of a file opens its sample stream with one framing word;
    a file mixing two is rejected when read."""
    c = synth.generate(20, 30, 64, 3, seed=9)
    synth.write_reference_files(c, str(tmp_path))
    assert run("keygen", "targets.txt", "tags.txt", "1", cwd=tmp_path).returncode == 0
    write_inputs(tmp_path, c, list(range(c.num_in_cts)), stream_word=0x1234)
    path = tmp_path / "encrypted_data.bin"
    raw = bytearray(path.read_bytes())
    raw[8 + formats.RECORD_BYTES + 4:8 + formats.RECORD_BYTES + 8] = (7).to_bytes(4, "little")
    path.write_bytes(bytes(raw))
    assert formats.read_stream_words(str(path)).tolist()[:2] == [0x1234, 7]
    r = run("cloud", "model", cwd=tmp_path)
    assert r.returncode != 0
    assert "ERROR: encrypted data records open with different framing words" in r.stdout


def test_sanitizer_build_recipe():
    """make asan (a dry run): the host library and executables with AddressSanitizer
    and UBSan, and the engine library with the same on its host code only."""
    src = os.path.join(ROOT, "idash2019_2_amd", "host")
    r = subprocess.run(["make", "-n", "-B", "-C", src, "asan"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    hip = [l for l in lines if "hipcc" in l and " -c " in l]
    assert len(hip) == 6 and all("-Xarch_host -fsanitize=address" in l for l in hip)
    assert not any(" -fsanitize=" in l.replace("-Xarch_host -fsanitize=", "") for l in hip)
    for exe in ("keygen", "encrypt", "cloud", "decrypt"):
        assert any(f"-o asan/{exe} " in l and "-fsanitize=address,undefined" in l for l in lines), exe


DRIVE = os.path.join(ROOT, "tests", "native", "records_drive")


def drive(*args):
    if not os.path.exists(DRIVE):
        pytest.fail(f"{DRIVE} is not built (run __graft_entry__.build())")
    return subprocess.run([DRIVE, *args], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("mode", ["pred", "data"])
def test_record_files_round_trip(tmp_path, mode):
    """The host's ciphertext-file reader (mapped a window of records at a time,
    payloads copied by several threads) and writer (threads writing ranges of
    records at their file offsets), on 9,000 records in shuffled index order:
    more than one reader window and many writer ranges. The file written is
    byte for byte the one a sequential writer makes from the samples read, in
    the order of the collection's map (eval/idash.cpp:540-556, 596-613)."""
    n, slots = 9000, 9501                       # some slots stay empty
    rng = np.random.default_rng(3)
    idx = rng.permutation(slots)[:n]
    ct = rng.integers(-(1 << 31), 1 << 31, (slots, 2, 1024), dtype=np.int64).astype(np.int32)
    var = rng.random(slots)
    src, out = tmp_path / "in.bin", tmp_path / "out.bin"
    formats.write_records(str(src), idx, ct, var, -5)
    count = slots if mode == "pred" else slots // 3   # encrypted data: 3 slots per tag position
    r = drive(mode, str(src), str(out), str(count))
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    assert lines[0] == "word -5"
    order = np.array([int(x) for x in lines[1:] if x], np.uint32)
    assert sorted(order.tolist()) == sorted(idx.tolist())
    want = tmp_path / "want.bin"
    formats.write_records(str(want), order, ct, var, -5)
    assert out.read_bytes() == want.read_bytes()


def test_record_file_errors(tmp_path):
    """Truncated files, duplicate indices and indices past the slots end the
    reader with ERROR (REQUIRE_DRAMATICALLY), checked in file order."""
    rng = np.random.default_rng(4)
    ct = rng.integers(-(1 << 31), 1 << 31, (20, 2, 1024), dtype=np.int64).astype(np.int32)
    var = np.zeros(20)
    path = tmp_path / "in.bin"

    def run_case(idx, slots, cut=0):
        if idx:
            formats.write_records(str(path), idx, ct, var, 0)
        else:
            path.write_bytes(bytes(8))                 # a count of zero
        if cut:
            path.write_bytes(path.read_bytes()[:-cut])
        return drive("pred", str(path), str(tmp_path / "out.bin"), str(slots))

    r = run_case([1, 2, 3], 20, cut=100)
    assert r.returncode != 0 and "ERROR: truncated encrypted predictions file" in r.stdout
    r = run_case([1, 2, 1], 20)
    assert r.returncode != 0 and "ERROR: duplicate encrypted predictions index" in r.stdout
    r = run_case([1, 19, 2], 10)
    assert r.returncode != 0 and "ERROR: encrypted predictions index out of range" in r.stdout
    r = run_case([], 10)
    assert r.returncode == 0 and r.stdout.split("\n")[0] == "word 0"
    assert (tmp_path / "out.bin").read_bytes() == bytes(8)


@pytest.mark.parametrize("names", [0, 1])
def test_result_file_lines(tmp_path, names):
    """write_decrypted_predictions (eval/idash.cpp:436-469): the header, then per
    sample and target position "sample,position or name,s0,s1,s2" with the
    scores in printf's %g; 150 samples (several formatting blocks, written by
    several threads at their offsets) over 300 positions."""
    P, ns = 300, 150
    rng = np.random.default_rng(5)
    v = (rng.standard_normal((P, 3, ns)) * 10.0 ** rng.integers(-9, 9, (P, 3, ns))).astype(np.float32)
    v[0, 0, :4] = [0.0, 1.0, -2.5, 123456789.0]
    src, out = tmp_path / "scores.bin", tmp_path / "result.csv"
    v.tofile(src)
    r = drive("csv", str(src), str(out), str(P), str(ns), str(names))
    assert r.returncode == 0, r.stdout + r.stderr
    want = ["Subject ID,target SNP,0,1,2"]
    for s in range(ns):
        for i in range(P):
            label = f"rs{i}" if names else str(1000 + 7 * i)
            want.append(f"{s},{label}," + ",".join("%g" % float(v[i, k, s]) for k in range(3)))
    assert out.read_text() == "\n".join(want) + "\n"


def test_score_formatter_matches_printf():
    """The result file's score formatter (host/gfmt.h: exact scaling in 128-bit
    integers, ties to even) prints what snprintf's %g prints, on every 997th
    float bit pattern, every multiple of 5 from 10^6 to 2^24 (exact ties at six
    significant digits, both signs) and each power of ten and its neighbours."""
    r = drive("gfmt", "997")
    assert r.returncode == 0 and "mismatches 0" in r.stdout, r.stdout[-2000:]
    assert int(r.stdout.split("checked ")[1].split()[0]) > 10_000_000
