"""Batched cloud step (ihe_batch_*) and the kernel autotuner
(ihe_band_create_tuned): cohorts of uneven sample counts on resident model
bands of several neighbour counts, in one persistent launch (or one launch per
kernel family) and as a captured hipGraph; every cohort bit-exact against its
own per-cohort launch and, on small sizes, against the oracle."""
import numpy as np
import pytest

import oracle
from idash2019_2_amd import _abi, synth
from idash2019_2_amd._abi import IheError

pytestmark = pytest.mark.gpu


def random_cts(rng, num_in):
    return rng.integers(-(1 << 31), 1 << 31, (num_in, 2, 1024), dtype=np.int64).astype(np.int32)


def make_models(tags, targets, sets, neighbours):
    """Model sets sharing one SNP layout (positions and windows) with
    independently drawn weights and biases."""
    models = {}
    for n in neighbours:
        for s in range(sets):
            c = synth.generate(tags, targets, 1024, n, seed=n)
            if s > 0:
                rng = np.random.default_rng(1000 * s + n)
                c.weights = rng.integers(-2048, 2048, c.weights.shape).astype(np.int16)
                c.bias = rng.integers(-2048, 2048, c.bias.shape).astype(np.int32)
            models[(s, n)] = (c, c.model_csr())
    return models


# the second model set is compiled for the per-wave kernels (32- or 64-row
# tiles), the first for the default shape, so that batches mix kernel families
PER_WAVE_SHAPES = {5: "2,1", 10: "2,1", 20: "1,2", 50: "1,5"}


@pytest.fixture(scope="module")
def setup(engine):
    import os
    models = make_models(60, 150, 2, (5, 10, 20, 50))
    bands = {}
    for (s, n), (c, csr) in models.items():
        os.environ.pop("IHE_CLOUD_SHAPE", None)
        if s == 1:
            os.environ["IHE_CLOUD_SHAPE"] = PER_WAVE_SHAPES[n]
        try:
            bands[(s, n)] = engine.band(1024, c.num_in_cts, *csr)
        finally:
            os.environ.pop("IHE_CLOUD_SHAPE", None)
        if s == 1:
            rows, kb = (int(v) for v in PER_WAVE_SHAPES[n].split(","))
            assert (bands[(s, n)].info()["tile_rows"], bands[(s, n)].info()["tile_kblocks"]) == \
                (32 * rows, kb)
    yield models, bands
    for b in bands.values():
        b.free()


def persistent_capable(info):
    """Tilings the persistent batch kernel runs: one K-block over 32 or 64
    rows, or 128 rows over at most 6 K-blocks."""
    return (info["tile_rows"] in (32, 64) and info["tile_kblocks"] == 1) or \
        (info["tile_rows"] == 128 and info["tile_kblocks"] <= 6)


def make_cohorts(engine, models, bands, sample_counts, with_var=True, keys=None):
    rng = np.random.default_rng(7)
    keys = keys or list(models)
    cohorts, host = [], []
    for i, ns in enumerate(sample_counts):
        key = keys[i % len(keys)]
        c, csr = models[key]
        R = 1024 // ns
        num_in = (3 * c.num_tags + R - 1) // R
        ct = random_cts(rng, num_in)
        var = rng.integers(1, 1 << 16, num_in) * 2.0 ** -60
        d_in, d_var = engine.to_device(ct), engine.to_device(var)
        d_out = engine.empty((c.num_outputs, 2, 1024), np.int32)
        d_ov = engine.empty((c.num_outputs,), np.float64) if with_var else None
        cohorts.append((bands[key], ns, num_in, d_in, d_var, d_out, d_ov))
        host.append((key, ns, ct, var))
    return cohorts, host


@pytest.mark.parametrize("mode", ["persistent", "families", "families-one-stream"])
def test_batch_matches_per_cohort_launches_and_oracle(engine, setup, mode, monkeypatch):
    """Persistent: cohorts on every band whose tiling the persistent kernel
    runs (both row-block kinds, rotated and not), one launch. Families: every
    band, including per-wave multi-K-block tilings, one launch per family,
    alternating over two streams (or all on one, IHE_BATCH_STREAMS=1)."""
    models, bands = setup
    counts = [1024, 700, 341, 200, 128, 100, 64, 33, 16, 1024, 512, 77]
    keys = None
    monkeypatch.delenv("IHE_BATCH_STREAMS", raising=False)
    if mode.startswith("families"):
        monkeypatch.setenv("IHE_BATCH_PERSISTENT", "0")
        if mode == "families-one-stream":
            monkeypatch.setenv("IHE_BATCH_STREAMS", "1")
    else:
        monkeypatch.setenv("IHE_BATCH_PERSISTENT", "1")
        keys = [k for k in models if persistent_capable(bands[k].info())]
        rows = {bands[k].info()["tile_rows"] for k in keys}
        assert len(keys) >= 4 and 128 in rows and rows & {32, 64}, keys
    cohorts, host = make_cohorts(engine, models, bands, counts, keys=keys)
    # per-cohort launches
    ref = []
    for band, ns, num_in, d_in, d_var, d_out, d_ov in cohorts:
        engine.cloud_layout(band, ns, num_in, d_in, d_var, d_out, d_ov)
        ref.append((d_out.numpy(), d_ov.numpy()))
        d_out.fill_zero()
        d_ov.fill_zero()
    batch = engine.batch(cohorts)
    info = batch.info()
    assert info["num_cohorts"] == len(counts)
    if mode == "persistent":
        assert info["persistent"] == 1 and info["num_launches"] == 1
    else:
        assert info["persistent"] == 0
        assert 2 <= info["num_launches"] <= 2 * 4 * 3     # one per family
    batch.run()
    batch.run()                                 # the work counters start at zero again
    for (_, _, _, _, _, d_out, d_ov), (o, v) in zip(cohorts, ref):
        assert (d_out.numpy() == o).all() and (d_ov.numpy() == v).all()
    # the oracle on a few cohorts (including rotated layouts)
    for i in (0, 2, 4, 8):
        key, ns, ct, var = host[i]
        c, (rp, ib, w) = models[key]
        R = 1024 // ns
        want, wv = oracle.cloud(ct, var, rp, ib, w, ns, R, 1024 // R)
        assert (ref[i][0] == want).all() and (ref[i][1] == wv).all(), i
    # captured as a graph: every replay equals the eager result
    batch.capture()
    assert batch.info()["captured"] == 1
    for _ in range(3):
        for _, _, _, _, _, d_out, d_ov in cohorts:
            d_out.fill_zero()
            d_ov.fill_zero()
        engine.sync()
        batch.run()
        engine.sync()
        for (_, _, _, _, _, d_out, d_ov), (o, v) in zip(cohorts, ref):
            assert (d_out.numpy() == o).all() and (d_ov.numpy() == v).all()
    batch.free()


def test_batch_without_variances_and_errors(engine, setup):
    models, bands = setup
    cohorts, _ = make_cohorts(engine, models, bands, [256, 1024, 50], with_var=False)
    batch = engine.batch(cohorts)
    batch.run()
    engine.sync()
    for band, ns, num_in, d_in, d_var, d_out, _ in cohorts:
        got = d_out.numpy()
        engine.cloud_layout(band, ns, num_in, d_in, None, d_out, None)
        assert (d_out.numpy() == got).all()
    batch.free()
    band, ns, num_in, d_in, d_var, d_out, d_ov = cohorts[0]
    with pytest.raises(IheError) as e:
        engine.batch([(band, ns, num_in - 1, d_in, d_var, d_out, d_ov)])
    assert e.value.code == _abi.IHE_ERR_RANGE
    c, csr = models[(0, 5)]
    valu = engine.band(1024, c.num_in_cts, *csr, variant=_abi.VARIANT_VALU)
    d_in_1024 = engine.empty((c.num_in_cts, 2, 1024), np.int32)
    with pytest.raises(IheError) as e:
        engine.batch([(valu, 1024, c.num_in_cts, d_in_1024, None, d_out, None)])
    assert e.value.code == _abi.IHE_ERR_UNSUPPORTED
    with pytest.raises(ValueError):                 # buffers below the sizes the step addresses
        engine.batch([(band, ns, num_in, d_in, d_var, d_out, d_ov),
                       (valu, 1024, c.num_in_cts, d_in, None, d_out, None)])
    valu.free()


@pytest.mark.parametrize("n", [5, 20])
def test_tuned_band_matches_default(engine, n):
    c = synth.generate(80, 300, 1024, n, seed=n)
    rp, ib, w = c.model_csr()
    rng = np.random.default_rng(n)
    ct = random_cts(rng, c.num_in_cts)
    var = np.full(c.num_in_cts, 2.0 ** -50)
    d_in, d_var = engine.to_device(ct), engine.to_device(var)
    d_out = engine.empty((c.num_outputs, 2, 1024), np.int32)
    d_ov = engine.empty((c.num_outputs,), np.float64)
    band, table = engine.band_tuned(1024, c.num_in_cts, rp, ib, w, d_in, d_out, reps=2)
    assert len(table) >= 3 and all(t[3] > 0 for t in table)
    assert {4, 8, 16} <= {t[2] for t in table}
    assert {32, 128} <= {t[0] for t in table}
    engine.cloud(band, d_in, d_var, d_out, d_ov)
    want, wv = oracle.cloud(ct, var, rp, ib, w, 1024, 1, 1024)
    assert (d_out.numpy() == want).all() and (d_ov.numpy() == wv).all()
    band.free()


def test_a_band_outlives_its_batch(engine, setup):
    """ihe_band_destroy refuses a band a live batch holds (IHE_ERR_BUSY), and
    frees it once the batch is gone."""
    models, _ = setup
    c, csr = models[(0, 5)]
    band = engine.band(1024, c.num_in_cts, *csr)
    d_in = engine.empty((c.num_in_cts, 2, 1024), np.int32)
    d_out = engine.empty((c.num_outputs, 2, 1024), np.int32)
    batch = engine.batch([(band, 1024, c.num_in_cts, d_in, None, d_out, None),
                          (band, 512, c.num_in_cts, d_in, None, d_out, None)])
    with pytest.raises(IheError) as e:
        band.free()
    assert e.value.code == _abi.IHE_ERR_BUSY
    batch.free()
    band.free()


def test_default_form(engine, setup, monkeypatch):
    """Without IHE_BATCH_PERSISTENT a batch is one persistent launch when all
    its families are one-K-block ones, and one launch per family when a
    128-row family is among them."""
    monkeypatch.delenv("IHE_BATCH_PERSISTENT", raising=False)
    models, bands = setup
    rows = {k: bands[k].info()["tile_rows"] for k in models}
    kb = {k: bands[k].info()["tile_kblocks"] for k in models}
    one = [k for k in models if rows[k] in (32, 64) and kb[k] == 1]
    wide = [k for k in models if rows[k] == 128 and kb[k] <= 6]
    assert one and wide
    for keys, persistent in ((one[:2], 1), ([one[0], wide[0]], 0)):
        cohorts, _ = make_cohorts(engine, models, bands, [1024, 100, 33], keys=keys)
        batch = engine.batch(cohorts)
        assert batch.info()["persistent"] == persistent
        batch.run()
        engine.sync()
        for band, ns, num_in, d_in, d_var, d_out, d_ov in cohorts:
            got, gv = d_out.numpy(), d_ov.numpy()
            engine.cloud_layout(band, ns, num_in, d_in, d_var, d_out, d_ov)
            assert (d_out.numpy() == got).all() and (d_ov.numpy() == gv).all()
        batch.free()
        for c in cohorts:
            for d in c[3:]:
                d.free()


@pytest.mark.parametrize("kinds", ["one-K-block only", "128-row only", "one cohort"])
@pytest.mark.parametrize("with_var", [True, False])
def test_one_launch_edge_mixes(engine, setup, kinds, with_var, monkeypatch):
    """The one-launch batch when one of its two parts is empty: only one-K-block
    families (the resident workgroups get variance blocks or nothing), only
    128-row families (no leading blocks), or a single cohort; with and without
    variances. Two runs in a row (the work counters start at zero again), each
    bit-exact against the per-cohort launches."""
    monkeypatch.setenv("IHE_BATCH_PERSISTENT", "1")
    models, bands = setup
    rows = {k: bands[k].info()["tile_rows"] for k in models}
    kb = {k: bands[k].info()["tile_kblocks"] for k in models}
    if kinds == "one-K-block only":
        keys = [k for k in models if rows[k] in (32, 64) and kb[k] == 1]
        counts = [1024, 341, 100, 16, 1024]
    elif kinds == "128-row only":
        keys = [k for k in models if rows[k] == 128 and kb[k] <= 6]
        counts = [700, 1024, 33, 128]
    else:
        keys = [k for k in models if persistent_capable(bands[k].info())][:1]
        counts = [200]
    assert keys, kinds
    cohorts, _ = make_cohorts(engine, models, bands, counts, with_var=with_var, keys=keys)
    ref = []
    for band, ns, num_in, d_in, d_var, d_out, d_ov in cohorts:
        engine.cloud_layout(band, ns, num_in, d_in, d_var if with_var else None, d_out, d_ov)
        ref.append((d_out.numpy(), d_ov.numpy() if with_var else None))
    batch = engine.batch(cohorts)
    assert batch.info()["persistent"] == 1 and batch.info()["num_launches"] == 1
    for _ in range(2):
        for c in cohorts:
            c[5].fill_zero()
            if with_var:
                c[6].fill_zero()
        engine.sync()
        batch.run()
        engine.sync()
        for c, (o, v) in zip(cohorts, ref):
            assert (c[5].numpy() == o).all()
            if with_var:
                assert (c[6].numpy() == v).all()
    batch.free()
