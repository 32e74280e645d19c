"""Host logic of the serving loop (no GPU): the LRU pool's pinning, eviction
and admission arithmetic, and the synthetic request stream."""

import numpy as np
import pytest

from contract_amd.serving import KINDS, OutKey, PoolFull, PropagatorPool, PropKey, make_requests

SMALL = (4, 4, 4, 8)  # 512 sites, 1,179,648 bytes per FP64 propagator
NB = 512 * 2304


def pool(cap_entries):
    made = []

    def materialise(key, recycled=None):
        made.append(key)
        return ("tensor", key, recycled)

    return PropagatorPool(cap_entries * NB, materialise), made


def test_lru_eviction_order_and_pins():
    p, made = pool(3)
    a, b, c, d = (PropKey(SMALL, i) for i in range(4))
    for k in (a, b, c):
        p.acquire(k)
        p.release(k)
    assert p.used == 3 * NB and made == [a, b, c]
    p.acquire(a)  # hit: a becomes most recently used, pinned
    assert p.hits == 1
    p.acquire(d)  # evicts b, the least recently used unpinned entry
    assert b not in p and a in p and c in p and d in p
    assert p.evictions == 1
    p.release(a)
    p.release(d)
    p.acquire(b)  # c is now the oldest unpinned
    assert c not in p and made[-1] == b


def test_pinned_entries_are_never_evicted():
    p, _ = pool(2)
    a, b, c = (PropKey(SMALL, i) for i in range(3))
    p.acquire(a)
    p.acquire(b)
    assert not p.fits([c])
    with pytest.raises(PoolFull):
        p.acquire(c)
    assert a in p and b in p
    p.release(b)
    assert p.fits([c]) and p.fits([a, c])
    p.acquire(c)
    assert b not in p


def test_fits_counts_resident_inputs_and_outputs():
    p, _ = pool(3)
    a, b, c = (PropKey(SMALL, i) for i in range(3))
    p.acquire(a)
    p.release(a)
    p.acquire(b)  # pinned
    out = OutKey(99, SMALL)
    # a is resident and unpinned: keeping it + one new output fits in 3
    assert p.fits([a], [out])
    # a, a new c and an output would need 4 entries with b pinned
    assert not p.fits([a, c], [out])
    p.insert(out, None)
    assert p.pinned_bytes() == 2 * NB


def test_eviction_hands_back_same_size_memory():
    p, _ = pool(2)
    a, b, c = (PropKey(SMALL, i) for i in range(3))
    va = p.acquire(a)
    p.release(a)
    p.acquire(b)
    vc = p.acquire(c)  # evicts a, whose memory has exactly c's size
    assert vc[2] is va
    big = OutKey(7, (8, 4, 4, 8))  # twice the size: nothing of that size to reuse
    p.release(b)
    p.release(c)
    assert p.insert(big, None) is None


def test_entry_larger_than_cap_is_rejected():
    p, _ = pool(1)
    with pytest.raises(PoolFull):
        p.acquire(PropKey((8, 8, 8, 8), 0))


def test_request_stream_is_poisson_and_mixed():
    reqs = make_requests(400.0, 5.0, seed=3, catalogue=4)
    n = len(reqs)
    assert abs(n - 2000) < 4 * np.sqrt(2000)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 400) < 0.1 / 400 * 4
    frac = {k: sum(r.kind == k for r in reqs) / n for k in KINDS}
    assert abs(frac["meson"] - 0.5) < 0.04 and abs(frac["baryon"] - 0.3) < 0.04 and abs(frac["diquark"] - 0.2) < 0.04
    for r in reqs:
        assert r.dims in [(16, 16, 16, 32), (24, 24, 24, 64), (32, 32, 32, 64), (48, 48, 48, 96)]
        assert all(k.dims == r.dims for k in r.props)
        if r.kind == "baryon":
            assert 1 <= len(r.gammas) <= 16 and len(set(r.gammas)) == len(r.gammas)
            assert r.props[0] == r.props[2]
        if r.kind != "diquark":
            assert 1 <= len(r.momenta) <= 27 and len(set(r.momenta)) == len(r.momenta)
    again = make_requests(400.0, 5.0, seed=3, catalogue=4)
    assert [(r.kind, r.dims, r.props, r.gammas, r.momenta, r.arrival) for r in again] == \
        [(r.kind, r.dims, r.props, r.gammas, r.momenta, r.arrival) for r in reqs]


def test_rates_share_the_request_sequence():
    # one seed: request i is the same request at every rate, its arrival
    # time scales as 1 / rate
    slow = make_requests(10.0, 6.0, seed=7)
    fast = make_requests(25.0, 6.0, seed=7)
    n = min(len(slow), len(fast))
    assert n > 40
    for a, b in zip(slow[:n], fast[:n]):
        assert (a.kind, a.dims, a.props, a.gammas, a.momenta) == (b.kind, b.dims, b.props, b.gammas, b.momenta)
        assert abs(a.arrival * 10.0 - b.arrival * 25.0) < 1e-9
    other = make_requests(10.0, 6.0, seed=8)
    assert [r.kind for r in other[:n]] != [r.kind for r in slow[:n]]


def test_catalogue_seeds_are_stable():
    k = PropKey((16, 16, 16, 32), 2, t0=14)
    assert k.seed == PropKey((16, 16, 16, 32), 2, t0=14).seed
    assert k.seed != PropKey((16, 16, 16, 32), 3, t0=14).seed


def test_meson_momentum_groups_pair_plus_minus_k():
    """The serving loop cuts the union of meson momenta into grid passes of
    at most 16 +-k slots: every momentum lands in exactly one pass, a momentum
    and its negative share a pass, and the 27 momenta of {-1,0,1}^3 (13 pairs
    and the origin) take one pass."""
    from contract_amd.serving import MOM27, Engine
    groups = Engine._momentum_groups(sorted(MOM27))
    assert [len(g) for g in groups] == [27]
    wide = [(x, y, z) for x in range(-2, 3) for y in range(-2, 3) for z in range(-2, 3)]  # 62 pairs + origin
    assert [len({min(m, tuple(-v for v in m)) for m in g}) for g in Engine._momentum_groups(wide)] == [16, 16, 16, 15]
    flat = [m for g in groups for m in g]
    assert sorted(flat) == sorted(MOM27) and len(set(flat)) == 27
    for g in groups:
        slots = {min(m, tuple(-v for v in m)) for m in g}
        assert len(slots) <= 16
        for m in g:
            neg = tuple(-v for v in m)
            assert neg in g  # every momentum of {-1,0,1}^3 has its negative in the union
    rng = np.random.default_rng(3)
    for _ in range(20):
        pick = sorted({MOM27[i] for i in rng.choice(27, size=int(rng.integers(1, 28)), replace=False)})
        groups = Engine._momentum_groups(pick)
        flat = [m for g in groups for m in g]
        assert sorted(flat) == pick
        assert all(len({min(m, tuple(-v for v in m)) for m in g}) <= 16 for g in groups)
        # pairs present in the union are never split across passes
        for g in groups:
            for m in g:
                neg = tuple(-v for v in m)
                assert neg not in pick or neg in g


def test_request_pins_resident_inputs_before_loading_missing_ones():
    """Resident B, C, D in a pool of three, then a request on (A, B): loading
    A must evict C (the least recently used entry the request does not
    need), never the request's own resident input B."""
    from contract_amd.serving import Engine, Request
    p, made = pool(3)
    a, b, c, d = (PropKey(SMALL, i) for i in range(4))
    for k in (b, c, d):
        p.acquire(k)
        p.release(k)
    eng = Engine.__new__(Engine)  # host-side scheduler state only
    eng.pool = p
    r = Request(0, "meson", SMALL, (a, b), [(5, 5)], [(0, 0, 0)])
    assert p.fits(r.props)
    eng._acquire_inputs(r)
    assert a in p and b in p and c not in p and d in p
    assert made == [b, c, d, a] and p.evictions == 1 and p.misses == 4
    assert p.entries[a][2] == 1 and p.entries[b][2] == 1


def test_repeated_input_is_pinned_once_per_listing():
    """A nucleon request lists u twice (u, d, u): it is pinned twice and the
    two releases of completion unpin it."""
    from contract_amd.serving import Engine, Request
    p, made = pool(4)
    u, d = PropKey(SMALL, 0), PropKey(SMALL, 1)
    eng = Engine.__new__(Engine)
    eng.pool = p
    r = Request(0, "baryon", SMALL, (u, d, u), [0], [(0, 0, 0)])
    eng._acquire_inputs(r)
    assert made == [u, d] and p.entries[u][2] == 2 and p.entries[d][2] == 1
    for k in r.props:
        p.release(k)
    assert p.pinned_bytes() == 0


def test_saturation_summary():
    from contract_amd.serving import saturation
    per = {"10.0": {"throughput_rps": 10.1}, "100.0": {"throughput_rps": 99.0},
           "300.0": {"throughput_rps": 240.0}}
    s = saturation(per)
    assert s["max_sustained_offered_rps"] == 100.0 and s["saturated"]
    assert s["peak_throughput_rps"] == 240.0 and s["at_offered_rps"] == 300.0


def test_saturation_uses_realised_arrivals_and_p99():
    from contract_amd.serving import saturation
    # 300 req/s nominal, 320 realised, 272 served with a growing queue: not
    # sustained although 272 >= 0.9 x 300
    per = {"100.0": {"throughput_rps": 113.7, "arrival_rps": 114.0, "p99_ms": 54.0},
           "200.0": {"throughput_rps": 215.7, "arrival_rps": 216.5, "p99_ms": 175.0},
           "300.0": {"throughput_rps": 272.2, "arrival_rps": 320.0, "p99_ms": 1520.0}}
    s = saturation(per)
    assert s["max_sustained_offered_rps"] == 200.0 and s["saturated"]
    # the throughput keeps up but p99 is past the bound: not sustained either
    per["200.0"]["p99_ms"] = 600.0
    assert saturation(per)["max_sustained_offered_rps"] == 100.0


def test_summary_records_realised_arrival_rate():
    from contract_amd import serving

    class R:
        def __init__(self, t, lat=0.01):
            self.arrival, self.done, self.latency, self.kind = t, t + lat, lat, "meson"
            self.dims, self.admitted = (4, 4, 4, 8), t + 0.002

    class P:
        cap = peak = hits = misses = evictions = 0

    reqs = [R(0.1 * i) for i in range(11)]  # 10 gaps over 1 s
    out = serving.summarise(reqs, 1.1, P(), {"meson": [1], "baryon": []}, {})
    assert abs(out["arrival_rps"] - 10.0) < 1e-9
    assert serving.summarise([R(0.0), R(0.0)], 0.1, P(), {"meson": [], "baryon": []}, {})["arrival_rps"] is None
    # the slowest requests, slowest first, with their queueing time
    reqs[3] = R(0.3, lat=0.5)
    slow = serving.summarise(reqs, 1.1, P(), {"meson": [1], "baryon": []}, {})["slowest"]
    assert len(slow) == 5 and slow[0] == ["meson", "4x4x4x8", 500.0, 2.0, 0.3]
    assert all(a[2] >= b[2] for a, b in zip(slow, slow[1:]))
