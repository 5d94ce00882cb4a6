"""The core's counter hash on the GPU (iscg_pcg4d / iscg_hash_units / iscg_hash_stream; core/src/rng.h:18-52): the
spec's golden `rng` and `stream` lines, bit for bit, then large batches (negative cells, k and base past 2^31, the
uint32 wrap of base * 4096 + i) against the oracle's scalar restatement on a sample and the vectorised host pcg4d on
all of them."""
import json
import os

import numpy as np
import pytest

from insiliscope_amd import configs, host

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "rng_vectors.json")


def unit(a):
    return ((np.asarray(a, np.uint32) >> 9).astype(np.float64) + 0.5) * 1.1920928955078125e-7


def test_golden_rng_lines(ctx):
    g = json.load(open(GOLDEN))["rng"]
    w = host.Pcg4d([r["seed"] for r in g], [r["cx"] for r in g], [r["cy"] for r in g], [r["k"] for r in g], ctx=ctx)
    assert w.tolist() == [r["abcd"] for r in g]
    for r in g:
        got = host.HashUnits(r["seed"], [r["cx"]], [r["cy"]], [r["k"]], ctx=ctx)
        assert got[0] == float(r["unit"])


def test_golden_stream_lines(ctx):
    for s in json.load(open(GOLDEN))["stream"]:
        got = host.HashStream(s["seed"], s["cx"], s["cy"], s["base"], len(s["draws"]), ctx=ctx)
        assert got.tolist() == [float(x) for x in s["draws"]]


def test_large_batches(ctx, orc):
    rng = np.random.default_rng(17)
    n = 1 << 20
    seed = int(rng.integers(0, 2 ** 32))
    cx = rng.integers(-2 ** 31, 2 ** 31, n)
    cy = rng.integers(-5000, 5000, n)
    k = rng.integers(0, 2 ** 32, n)
    words = host.Pcg4d(np.full(n, seed), cx, cy, k, ctx=ctx)
    want = np.stack(configs.pcg4d(np.full(n, seed, np.uint32), cx.astype(np.uint32), cy.astype(np.uint32),
                                  k.astype(np.uint32)), axis=1)
    assert np.array_equal(words, want)
    u = host.HashUnits(seed, cx, cy, k, ctx=ctx)
    assert np.array_equal(u, unit(want[:, 0]))
    assert u.min() > 0 and u.max() < 1
    for i in rng.integers(0, n, 200):
        assert orc.hash_unit(seed, int(cx[i]), int(cy[i]), int(k[i])) == u[i]


@pytest.mark.parametrize("base", [0, 7, 2 ** 20 - 1, 2 ** 20, 0xFFFFFFFF])
def test_streams_wrap(ctx, orc, base):
    n = 5000
    got = host.HashStream(99, -7, 12, base, n, ctx=ctx)
    ks = (np.uint64(base) * np.uint64(4096) + np.arange(n, dtype=np.uint64)) & np.uint64(0xFFFFFFFF)
    w = configs.pcg4d(np.full(n, 99, np.uint32), np.full(n, -7, np.int64).astype(np.uint32),
                      np.full(n, 12, np.uint32), ks.astype(np.uint32))
    assert np.array_equal(got, unit(w[0]))
    assert got[:50].tolist() == list(orc.hash_stream(99, -7, 12, base, 50))


def test_empty_and_null(ctx):
    assert host.HashStream(1, 0, 0, 0, 0, ctx=ctx).size == 0
    assert host.HashUnits(1, [], [], [], ctx=ctx).size == 0
    assert ctx.L.iscg_hash_units(ctx.h, 1, None, None, None, 4, None) != 0
