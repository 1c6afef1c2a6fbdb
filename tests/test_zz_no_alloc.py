"""fp_survivor_map_static, the entry point that allocates nothing, against
fp_compact: the same survivor map (ascending active slots) and active count
on two calls in a row, the second on a smaller input, through the C ABI.

The new entry point keeps its tile counts in the library's static device
array, so the second call reuses what the first left there; the test also
covers an empty input, the size limit (rejected on the host before any
device work) and the largest input the scratch covers (against numpy)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def flagged(fp, n, frac, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, fp.PARTICLE_DTYPE)
    p["flags"] = (rng.random(n) < frac).astype(np.uint32)
    p["pos_z"] = rng.random(n).astype(np.float32)
    return p


def compact_reference(fp, parts):
    with fp.GpuTimestep(len(parts), (4, 4, 4), n_species=1) as ctx:
        ctx.upload_store(parts)
        return ctx.compact()


def call_static(fp, flags_dev, n, map_dev, count_dev):
    import torch
    lib = fp.library()
    rc = lib.fp_survivor_map_static(torch.cuda.current_device(), C.c_void_p(flags_dev.data_ptr()), n,
                                    C.c_void_p(map_dev.data_ptr()), C.c_void_p(count_dev.data_ptr()), None)
    assert rc == fp.FP_OK, lib.fp_last_error(None).decode()
    torch.cuda.synchronize()
    a = int(count_dev.cpu().numpy().view(np.uint32)[0])
    return a, map_dev[:a].cpu().numpy().view(np.uint32)


def test_static_survivor_map_equals_fp_compact_on_two_calls(fp):
    import torch
    dev = torch.device("cuda:0")
    n1, n2 = 3_000_777, 1_000_003
    map_dev = torch.full((n1,), -1, dtype=torch.int32, device=dev)
    count_dev = torch.full((1,), 7, dtype=torch.int32, device=dev)
    for n, frac, seed in ((n1, 0.61, 3), (n2, 0.37, 4)):
        parts = flagged(fp, n, frac, seed)
        flags_dev = torch.from_numpy(parts["flags"].view(np.int32).copy()).to(dev)
        got_a, got_map = call_static(fp, flags_dev, n, map_dev, count_dev)
        want_a, want_map = compact_reference(fp, parts)
        assert got_a == want_a, (n, got_a, want_a)
        assert np.array_equal(got_map, want_map), n
        assert np.array_equal(got_map, np.flatnonzero(parts["flags"] & 1).astype(np.uint32))


def test_static_survivor_map_empty_input_and_size_limit(fp):
    import torch
    dev = torch.device("cuda:0")
    count_dev = torch.full((1,), 7, dtype=torch.int32, device=dev)
    dummy = torch.zeros(4, dtype=torch.int32, device=dev)
    a, _ = call_static(fp, dummy, 0, dummy, count_dev)
    assert a == 0
    lib = fp.library()
    rc = lib.fp_survivor_map_static(torch.cuda.current_device(), C.c_void_p(dummy.data_ptr()),
                                    fp.SURVIVOR_MAP_STATIC_MAX + 1, C.c_void_p(dummy.data_ptr()),
                                    C.c_void_p(count_dev.data_ptr()), None)
    assert rc == fp.FP_E_INVALID
    assert b"FP_SURVIVOR_MAP_STATIC_MAX" in lib.fp_last_error(None)
    assert int(count_dev.cpu().numpy()[0]) == 0  # untouched by the rejected call


def test_static_survivor_map_at_the_largest_size_the_scratch_covers(fp):
    import torch
    dev = torch.device("cuda:0")
    n = fp.SURVIVOR_MAP_STATIC_MAX
    rng = np.random.default_rng(5)
    flags = (rng.random(n, dtype=np.float32) < 0.5).astype(np.uint32)
    flags[-1] = 1
    flags_dev = torch.from_numpy(flags.view(np.int32)).to(dev)
    map_dev = torch.empty(n, dtype=torch.int32, device=dev)
    count_dev = torch.zeros(1, dtype=torch.int32, device=dev)
    a, got = call_static(fp, flags_dev, n, map_dev, count_dev)
    want = np.flatnonzero(flags).astype(np.uint32)
    assert a == len(want) and got[-1] == n - 1
    assert np.array_equal(got, want)
