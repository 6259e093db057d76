"""The crowded regime of a large settled pile, bitwise against the reference
(golden fixture `slabstack`, written by the strict reference World): one
island of 35,000-90,000 contacts, 728-925 colours (far past one 64-bit colour
mask), and warm-start matches decided by ties among a pair's cached contacts,
which only the reference's std::sort order settles.  Also the device's exact
std::sort against the host libstdc++.  Needs an MI355X."""
import numpy as np
import pytest

from vxfix import as_bodies, as_events, as_state, bits_equal, body_state, default_options, load_case, oracle, ptr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from velox_amd.device import Context
    c = Context()
    yield c
    c.close()


def _report(got, want):
    a8 = np.ascontiguousarray(got).view(np.uint8).reshape(len(got), -1)
    b8 = np.ascontiguousarray(want).view(np.uint8).reshape(len(want), -1)
    bad = np.nonzero((a8 != b8).any(axis=1))[0]
    return f"{len(bad)} of {len(got)} differ; first {bad[:8].tolist()}"


def _host_std_sort(keys, vals, starts):
    keys, vals = keys.copy(), vals.copy()
    for a, b in zip(starts[:-1], starts[1:]):
        k, v = np.ascontiguousarray(keys[a:b]), np.ascontiguousarray(vals[a:b])
        oracle().vxo_std_sort(ptr(k), ptr(v), len(k))
        keys[a:b], vals[a:b] = k, v
    return keys, vals


@pytest.mark.parametrize("n,kinds,segments,seed", [(2000, 50, 1, 1), (70000, 9000, 1, 2), (131072, 40, 1, 3),
                                                   (90000, 30000, 37, 4), (5000, 2, 1, 5), (17, 3, 1, 6)])
def test_device_std_sort_matches_libstdcxx(ctx, n, kinds, segments, seed):
    """Equal keys come out in exactly the order std::sort leaves them, for
    one range and for many independent ranges (batched worlds)."""
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, kinds, n).astype(np.uint64)
    if seed % 2 == 0:  # two sorted runs: the warm-start cache's usual shape
        h = n // 3
        keys[:h] = np.sort(keys[:h], kind="stable")
        keys[h:] = np.sort(keys[h:], kind="stable")
    vals = np.arange(n, dtype=np.uint32)
    starts = np.array([0, n], np.uint32) if segments == 1 else \
        np.concatenate([[0], np.sort(rng.integers(0, n, segments - 1)), [n]]).astype(np.uint32)
    print(f"[gpu] std::sort n={n} kinds={kinds} segments={segments}", flush=True)
    got_k, got_v = ctx.std_sort(keys, vals, None if segments == 1 else starts)
    want_k, want_v = _host_std_sort(keys, vals, starts)
    assert np.array_equal(got_k, want_k)
    assert np.array_equal(got_v, want_v), f"{np.count_nonzero(got_v != want_v)} values out of std::sort's order"


def test_slab_stack_matches_reference(ctx):
    """Three steps of the crowded fixture through the resident device step:
    body states, contact and colour counts, and the last step's contact
    events bitwise; the big-island kernels and the device-wide colouring
    must all have run, and the cache order must have decided some matches."""
    case = load_case("slabstack")
    meta = case["meta"]
    print("[gpu] slabstack 3 steps", flush=True)
    ctx.load(as_bodies(case["bodies_in"]), case["soup"], gravity=tuple(meta["gravity"]), substeps=meta["substeps"],
             options=default_options())
    done = 0
    ties = 0
    ran = {}
    for k, target in enumerate(meta["steps"]):
        ctx.step(np.float32(meta["dt"]), target - done)
        done = target
        st = ctx.stats()
        print(f"[gpu] step {target}: {st}", flush=True)
        got = body_state(ctx.bodies())
        want = as_state(case[f"state_{target}"])
        assert bits_equal(got, want), f"slabstack step {target}: " + _report(got, want)
        assert st["contacts"] == meta["contacts_at"][k]
        assert st["colors"] == meta["colors"][k]
        assert st["colors"] > 64
        assert st["bigIslandContacts"] > 4096
        assert st["launches"]["greedy_device"] == 1
        for kind, count in st["launches"].items():
            ran[kind] = ran.get(kind, 0) + count
        ties += st["warmTies"]
        if f"events_{target}" in case:
            want_e = as_events(case[f"events_{target}"])
            got_e = ctx.events()
            assert len(got_e) == len(want_e)
            assert np.array_equal(got_e.view(np.uint8), np.ascontiguousarray(want_e).view(np.uint8))
    # Every big-island kernel ran (the medium stack is one island of 2,049 to
    # 4,096 contacts at step 1, kIslandSweeps' range).
    assert all(ran[k] > 0 for k in ("batch", "island", "color_pair", "color_run", "tail")), ran
    assert ties > 0, "no warm-start tie was reached: the fixture no longer exercises the cache order"


@pytest.mark.parametrize("n,bits,spread,seed", [(2, 8, 4, 1), (1000, 8, 256, 2), (4096, 16, 3, 3), (4097, 16, 70000, 4),
                                                (20000, 21, 1 << 21, 5), (70000, 24, 50, 6), (185000, 42, 1 << 42, 7),
                                                (300001, 64, 1 << 63, 8), (100000, 3, 8, 9)])
def test_device_radix_sort_is_a_stable_sort(ctx, n, bits, spread, seed):
    """The step's radix sort (one launch per 8-bit pass, tile offsets by
    look-back) against numpy's stable argsort on the low `bits` of the key:
    single and many tiles, few and many distinct keys, keys with bits above
    `bits` set (ignored, order kept)."""
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, spread, n, dtype=np.uint64) | (rng.integers(0, 2, n, dtype=np.uint64) << np.uint64(63))
    vals = np.arange(n, dtype=np.uint32)
    print(f"[gpu] radix sort n={n} bits={bits}", flush=True)
    got_k, got_v = ctx.radix_sort(keys, vals, bits)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    order = np.argsort(keys & mask, kind="stable")
    assert np.array_equal(got_v, vals[order])
    assert np.array_equal(got_k, keys[order])
