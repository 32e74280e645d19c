"""Host logic of the autotuner (no GPU): the batch signatures follow the
kernels the library would run, the candidate sets hold the defaults, and the
cache file round-trips."""

import json

import pytest

from contract_amd import autotune

MOM27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
MOM7 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


def brute_slots(moms):
    """Independent count: pair each momentum greedily with the first unused
    later negative (the zero momentum and self-negatives stay alone)."""
    left = list(range(len(moms)))
    n = 0
    while left:
        p = left.pop(0)
        n += 1
        m = tuple(moms[p])
        if any(m):
            for q in left:
                if tuple(moms[q]) == tuple(-v for v in m):
                    left.remove(q)
                    break
    return n


@pytest.mark.parametrize("moms,slots,passes", [
    (MOM27, 14, (2,)),
    (MOM7, 4, (1,)),
    ([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
      (1, 0, 1), (-1, 0, -1)], 5, (1,)),                                    # 10 momenta, 5 +-k pairs
    ([(k, 0, 0) for k in range(1, 21)], 20, (2, 1)),                         # 20 unpaired: two passes
    ([(1, 0, 0), (1, 0, 0), (-1, 0, 0)], 2, (1,)),                           # a duplicate pairs once
    ([(0, 0, 0), (0, 0, 0)], 2, (1,)),
    ([(k, 1, 0) for k in range(-4, 5)] + [(k, -1, 0) for k in range(-4, 5)], 9, (2,)),
])
def test_momentum_slots_and_passes(moms, slots, passes):
    assert autotune.momentum_slots(moms) == slots == brute_slots(moms)
    assert autotune.grid_passes(moms) == passes


def test_signature_separates_kernel_forms():
    class P:  # stands in for a propagator tensor: only dtype is read
        def __init__(self, dtype):
            self.dtype = dtype
    r = [(P("torch.complex128"), P("torch.complex128"), (4, 4, 4, 8), 0)]
    ten_pairs = MOM27[:5] + [tuple(-v for v in m) for m in MOM27[:5]]
    # 10 momenta in 5 pairs run the one-tile kernel, like 4 momenta do
    assert autotune.signature(r, ten_pairs) == autotune.signature(r, MOM7[:4])
    # 20 unpaired momenta need two passes: not the key of 10 paired ones
    assert autotune.signature(r, [(k, 0, 0) for k in range(1, 21)]) != autotune.signature(r, ten_pairs)
    # storage form and lattice order are part of the key
    r32 = [(P("torch.complex64"), P("torch.complex64"), (4, 4, 4, 8), 0)]
    assert autotune.signature(r32, MOM7) != autotune.signature(r, MOM7)
    two = r + [(P("torch.complex128"), P("torch.complex128"), (8, 4, 4, 8), 0)]
    assert autotune.signature(two, MOM7) != autotune.signature(two[::-1], MOM7)


def test_candidates_hold_the_defaults():
    assert autotune.BARYON_DEFAULT in autotune.baryon_candidates()
    assert autotune.BARYON_DEFAULT in autotune.baryon_candidates(items=(4,), sites=(64,))
    c = autotune.candidates()
    assert (1024, 4, 2, 2) in c and (1024, 4, 4, 3) in c
    assert len(c) == len(set(c)) == 5 * 2 * 4 * 2 + 5 * 2
    assert all(u != 3 for *_, u in autotune.candidates(units=(1, 2)))
    # plain FP32: vector-ALU plans and the wave-local form, no shared-tile MFMA form
    c13 = autotune.candidates(units=(1, 3))
    assert {u for *_, u in c13} == {1, 3} and (1024, 4, 8, 3) in c13 and (1024, 8, 2, 1) in c13


def test_cache_file_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("CONTRACT_AMD_TUNE_CACHE", str(path))
    autotune.clear_cache()
    key = ("baryon", "torch.complex128", True, True, True, 1, (((24, 24, 24, 64), 0),))
    assert autotune.load_cached("gfx950", key) is None
    autotune.store_cached("gfx950", key, (8, 128), 1.25)
    autotune.store_cached("other", key, (4, 64), 2.0)
    assert autotune.load_cached("gfx950", key) == ((8, 128), 1.25)
    assert autotune.load_cached("other", key) == ((4, 64), 2.0)
    assert len(json.loads(path.read_text())) == 2
    assert autotune._lookup("gfx950", key, True) == ((8, 128), 1.25)  # file -> memory
    assert autotune._lookup("gfx950", key, False) is None
    path.write_text("{not json")
    assert autotune.load_cached("gfx950", key) is None
    autotune.store_cached("gfx950", key, (16, 256), 0.5)  # a corrupt file is replaced
    assert autotune.load_cached("gfx950", key) == ((16, 256), 0.5)
    autotune.clear_cache(file=True)
    assert not path.exists() and autotune._lookup("gfx950", key, True) is None
