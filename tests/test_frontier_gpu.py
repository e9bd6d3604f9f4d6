"""MLB frontier on the device (frontier/impls/mlb_frontier.hpp) against the
reference's frontier/intersection known answers and brute-force set algebra."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _expect():
    with open(os.path.join(GOLDEN, "reference_cases.json")) as f:
        return json.load(f)["expect"]


def test_mlb_known_answers(sg, device):
    # tests/frontier/mlb_frontier.cpp:10-24 (star_5 has 5 vertices)
    e = _expect()
    f = sg.Frontier(device, 5)
    for v in e["mlb_insert"]:
        f.insert(v)
    assert f.active() == e["mlb_insert"]
    saved = f.save()
    f.remove(3)
    assert f.active() == e["mlb_after_remove_3"]
    f.clear()
    assert f.empty()
    f.load(saved)
    assert f.active() == e["mlb_insert"]
    assert f.check(1) and not f.check(0) and f.size() == 3


def test_intersection_known_answer(sg, device):
    # tests/operators/intersection.cpp:14-20
    e = _expect()
    a, b = sg.Frontier(device, 5), sg.Frontier(device, 5)
    for v in e["intersection_lhs"]:
        a.insert(v)
    for v in e["intersection_rhs"]:
        b.insert(v)
    a.intersect(b)
    assert a.active() == e["intersection_out"]


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4095, 4096, 4097, 300_001])
def test_compaction_and_set_algebra(sg, device, n):
    rng = np.random.default_rng(n)
    for density in (0.0, 0.001, 0.3, 1.0):
        sa = np.nonzero(rng.random(n) < density)[0]
        sb = np.nonzero(rng.random(n) < 0.5)[0]
        fa, fb = sg.Frontier(device, n), sg.Frontier(device, n)
        _fill(device, fa, sa)
        _fill(device, fb, sb)
        assert fa.size() == len(sa) and fa.empty() == (len(sa) == 0)
        assert fa.compact(0).tolist() == sa.tolist()  # ascending element ids
        words = np.unique(sa // 64)
        assert fa.compact(1).tolist() == words.tolist()
        l0, l1 = fa.words()
        not_full = np.nonzero(l0 != np.uint64(0xFFFFFFFFFFFFFFFF))[0]
        assert fa.compact(2).tolist() == not_full.tolist()
        fa.merge(fb)
        union = np.union1d(sa, sb)
        assert fa.compact(0).tolist() == union.tolist()
        # level 1 is rebuilt by merge: exactly the non-empty level-0 words
        l0, l1 = fa.words()
        nz = np.nonzero(l0)[0]
        bits = np.unpackbits(l1.view(np.uint8), bitorder="little")
        assert np.nonzero(bits)[0].tolist() == nz.tolist()
        fa.intersect(fb)
        assert fa.compact(0).tolist() == sb.tolist()


def _fill(device, f, ids):
    """Load a set by words (bulk), level 1 derived on the host."""
    v = f.view()
    l0 = np.zeros(v.words0, dtype=np.uint64)
    np.bitwise_or.at(l0, ids // 64, np.left_shift(np.uint64(1), (ids % 64).astype(np.uint64)))
    l1 = np.zeros(v.words1, dtype=np.uint64)
    nz = np.nonzero(l0)[0]
    np.bitwise_or.at(l1, nz // 64, np.left_shift(np.uint64(1), (nz % 64).astype(np.uint64)))
    f.load(np.concatenate([l0, l1]))


def test_swap(sg, device):
    a, b = sg.Frontier(device, 100), sg.Frontier(device, 100)
    a.insert(7)
    b.insert(42)
    sg.swap(a, b)
    assert a.active() == [42] and b.active() == [7]


def test_out_of_range_raises(sg, device):
    f = sg.Frontier(device, 10)
    with pytest.raises(sg.SygraphError):
        f.insert(10)
    with pytest.raises(sg.SygraphError):
        f.check(-1)
