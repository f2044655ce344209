"""Packet-pool compaction (keep_pool) and per-destination binning (comm reg/cpy_d).

CPU: the oracle restatement is pinned against brute force written from the
definitions (the kept multiset, the hole/survivor pairing, a stable sort by
key) and against invariants.  GPU: kp_pool_keep / kp_bin_by_key / kp_gather
through the C ABI must equal the oracle bit for bit on the same inputs,
including empty, all-kept, none-kept, one-bin, many-bin (three radix
passes) and out-of-range-key cases.
"""
import ctypes as C

import numpy as np
import pytest

from kratos_polrad_amd import _abi


def _records(n, unit_words, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**31, size=(n, unit_words), dtype=np.int64).astype(np.uint32)


# ------------------------------------------------------------------ CPU pins
@pytest.mark.parametrize("n,p_keep,seed", [(0, 0.5, 0), (1, 0.0, 1), (1, 1.0, 2), (17, 0.5, 3),
                                           (1000, 0.9, 4), (1000, 0.1, 5), (4096, 0.5, 6)])
def test_oracle_pool_keep_matches_definition(oracle_lib, n, p_keep, seed):
    rng = np.random.default_rng(seed)
    pool = _records(n, 5, seed)
    keep = (rng.random(n) < p_keep).astype(np.uint8)
    got, nk = oracle_lib.pool_keep(pool, keep)
    assert nk == int(keep.sum())
    # survivors as a multiset, and nothing else, in the head
    want = pool[keep.astype(bool)]
    assert sorted(map(tuple, got[:nk])) == sorted(map(tuple, want))
    # kept packets of the head never move
    head_kept = np.nonzero(keep[:nk])[0]
    np.testing.assert_array_equal(got[head_kept], pool[head_kept])
    # pairing: j-th head hole <- j-th tail survivor from the end
    holes = np.nonzero(keep[:nk] == 0)[0]
    movers = np.nonzero(keep[nk:])[0][::-1] + nk
    assert len(holes) == len(movers)
    np.testing.assert_array_equal(got[holes], pool[movers])
    # the tail is untouched
    np.testing.assert_array_equal(got[nk:], pool[nk:])


@pytest.mark.parametrize("n,n_bins,seed", [(0, 3, 0), (1, 1, 1), (100, 1, 2), (1000, 7, 3),
                                           (5000, 300, 4), (3000, 70000, 5)])
def test_oracle_bin_by_key_is_a_stable_sort(oracle_lib, n, n_bins, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(-2, n_bins + 2, size=n).astype(np.int32)
    offsets, order, slot = oracle_lib.bin_by_key(keys, n_bins)
    valid = (keys >= 0) & (keys < n_bins)
    idx = np.nonzero(valid)[0]
    want_order = idx[np.argsort(keys[idx], kind="stable")]
    np.testing.assert_array_equal(order, want_order)
    counts = np.bincount(keys[valid], minlength=n_bins)
    np.testing.assert_array_equal(np.diff(offsets), counts)
    assert offsets[0] == 0 and offsets[-1] == valid.sum()
    # slot is the inverse permutation, -1 for records that stay
    np.testing.assert_array_equal(slot[~valid], -1)
    np.testing.assert_array_equal(slot[order], np.arange(len(order)))
    # gather puts every record in its destination's range
    recs = _records(n, 3, seed)
    out = oracle_lib.gather(recs, order)
    np.testing.assert_array_equal(out, recs[want_order])


def test_workspace_size_is_monotone():
    L = _abi.lib()
    sizes = [L.kp_compact_workspace_bytes(n, 1000) for n in (0, 1, 4096, 10**6, 10**8)]
    assert all(b >= a for a, b in zip(sizes, sizes[1:]))
    assert sizes[-1] >= 16 * 10**8  # four 32-bit arrays of n


# ------------------------------------------------------------------ GPU parity
def _ws(n, n_bins, torch):
    b = int(_abi.lib().kp_compact_workspace_bytes(n, n_bins))
    return torch.empty(max(b, 1), dtype=torch.uint8, device="cuda"), b


def _gpu_pool_keep(pool, keep):
    import torch

    n = pool.shape[0]
    unit = pool.shape[1] * 4
    d_pool = torch.from_numpy(pool.copy()).to("cuda")
    d_keep = torch.from_numpy(keep).to("cuda")
    d_nk = torch.full((1,), -7, dtype=torch.int64, device="cuda")
    ws, wb = _ws(n, 1, torch)
    s = torch.cuda.current_stream()
    _abi.check(_abi.lib().kp_pool_keep(C.c_void_p(d_pool.data_ptr()), n, unit,
                                        C.c_void_p(d_keep.data_ptr()),
                                        C.c_void_p(d_nk.data_ptr()), C.c_void_p(ws.data_ptr()),
                                        wb, C.c_void_p(s.cuda_stream)))
    torch.cuda.synchronize()
    return d_pool.cpu().numpy(), int(d_nk.item())


def _gpu_bin(keys, n_bins, recs=None):
    import torch

    n = len(keys)
    d_keys = torch.from_numpy(keys).to("cuda")
    offs = torch.full((n_bins + 1,), -1, dtype=torch.int64, device="cuda")
    order = torch.full((max(n, 1),), -5, dtype=torch.int32, device="cuda")
    slot = torch.full((max(n, 1),), -5, dtype=torch.int32, device="cuda")
    ws, wb = _ws(n, n_bins, torch)
    s = torch.cuda.current_stream()
    _abi.check(_abi.lib().kp_bin_by_key(C.c_void_p(d_keys.data_ptr()), n, n_bins,
                                         C.c_void_p(offs.data_ptr()),
                                         C.c_void_p(order.data_ptr()),
                                         C.c_void_p(slot.data_ptr()), C.c_void_p(ws.data_ptr()),
                                         wb, C.c_void_p(s.cuda_stream)))
    out = None
    if recs is not None and n > 0:
        d_recs = torch.from_numpy(recs).to("cuda")
        d_out = torch.zeros_like(d_recs)
        _abi.check(_abi.lib().kp_gather(C.c_void_p(d_recs.data_ptr()), recs.shape[1] * 4,
                                         C.c_void_p(order.data_ptr()),
                                         C.c_void_p(offs.data_ptr() + 8 * n_bins), n,
                                         C.c_void_p(d_out.data_ptr()),
                                         C.c_void_p(s.cuda_stream)))
        torch.cuda.synchronize()
        out = d_out.cpu().numpy()
    torch.cuda.synchronize()
    o = offs.cpu().numpy()
    return o, order.cpu().numpy()[:o[-1]], slot.cpu().numpy()[:n], out


@pytest.mark.gpu
@pytest.mark.parametrize("n,p_keep,words,seed", [
    (0, 0.5, 4, 0), (1, 0.0, 4, 1), (1, 1.0, 4, 2), (4095, 0.5, 1, 3), (4097, 0.5, 2, 4),
    (100_003, 0.9, 16, 5), (100_003, 0.1, 3, 6), (2_000_000, 0.7, 16, 7),
    (300_000, 1.0, 5, 8), (300_000, 0.0, 5, 9)])
def test_gpu_pool_keep_equals_oracle(oracle_lib, n, p_keep, words, seed):
    rng = np.random.default_rng(seed)
    pool = _records(n, words, seed)
    keep = (rng.random(n) < p_keep).astype(np.uint8)
    want, nk = oracle_lib.pool_keep(pool, keep)
    got, gnk = _gpu_pool_keep(pool, keep)
    assert gnk == nk
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("n,n_bins,seed", [
    (0, 5, 0), (1, 1, 1), (5000, 1, 2), (4096, 255, 3), (4097, 256, 4), (123_457, 2234, 5),
    (1_000_000, 64, 6), (200_000, 70_000, 7), (50_000, 20_000_000, 8)])
def test_gpu_bin_by_key_equals_oracle(oracle_lib, n, n_bins, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(-3, n_bins + 3, size=n).astype(np.int32)
    recs = _records(n, 4, seed)
    w_off, w_order, w_slot = oracle_lib.bin_by_key(keys, n_bins)
    g_off, g_order, g_slot, g_out = _gpu_bin(keys, n_bins, recs)
    np.testing.assert_array_equal(g_off, w_off)
    np.testing.assert_array_equal(g_order, w_order)
    np.testing.assert_array_equal(g_slot, w_slot)
    if n:
        np.testing.assert_array_equal(g_out[:w_off[-1]], oracle_lib.gather(recs, w_order))


@pytest.mark.gpu
def test_gpu_bin_sorted_and_reversed_keys(oracle_lib):
    """Long runs of one key across tiles (the stable order spans tiles)."""
    n, n_bins = 50_000, 17
    for keys in (np.sort(np.arange(n) % n_bins), np.sort(np.arange(n) % n_bins)[::-1],
                 np.full(n, 3)):
        keys = np.ascontiguousarray(keys, dtype=np.int32)
        w_off, w_order, w_slot = oracle_lib.bin_by_key(keys, n_bins)
        g_off, g_order, g_slot, _ = _gpu_bin(keys, n_bins)
        np.testing.assert_array_equal(g_off, w_off)
        np.testing.assert_array_equal(g_order, w_order)
        np.testing.assert_array_equal(g_slot, w_slot)


@pytest.mark.gpu
def test_gpu_compaction_argument_errors():
    import torch

    L = _abi.lib()
    d = torch.zeros(16, dtype=torch.int64, device="cuda")
    p = C.c_void_p(d.data_ptr())
    assert L.kp_pool_keep(p, 10, 6, p, p, p, 1 << 20, None) == _abi.KP_ERR_ARG  # unit % 4
    assert L.kp_pool_keep(p, 10, 8, p, p, p, 4, None) == _abi.KP_ERR_ARG  # workspace
    assert L.kp_bin_by_key(p, 10, 0, p, None, None, p, 1 << 20, None) == _abi.KP_ERR_ARG
    assert "n_bins" in L.kp_last_error().decode()
