"""kp_bin_by_key_static (the binning entry point that allocates nothing)
against kp_bin_by_key, through the C ABI.

Its workspace is a fixed device array inside the library, shared by every
call on the GPU and left as the last call wrote it; two calls in a row, the
second smaller than the first, must each give kp_bin_by_key's result bit for
bit (offsets, order and slot are integer outputs)."""
import ctypes as C

import numpy as np
import pytest

from kratos_polrad_amd import _abi

pytestmark = pytest.mark.gpu

KP_BIN_STATIC_MAX = 1 << 21


def _keys(n, n_bins, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, n_bins, n).astype(np.int32)
    # records that stay (-1 and keys past the last bin) sort last in both calls
    stay = rng.random(n) < 0.1
    k[stay] = np.where(rng.random(int(stay.sum())) < 0.5, -1, n_bins + 3)
    return k


def _bin(keys, n_bins, static):
    import torch

    n = len(keys)
    d_keys = torch.from_numpy(keys).to("cuda")
    offs = torch.full((n_bins + 1,), -1, dtype=torch.int64, device="cuda")
    order = torch.full((max(n, 1),), -5, dtype=torch.int32, device="cuda")
    slot = torch.full((max(n, 1),), -5, dtype=torch.int32, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = _abi.lib()
    if static:
        st = L.kp_bin_by_key_static(C.c_void_p(d_keys.data_ptr()), n, n_bins,
                                    C.c_void_p(offs.data_ptr()), C.c_void_p(order.data_ptr()),
                                    C.c_void_p(slot.data_ptr()), s)
    else:
        wb = int(L.kp_compact_workspace_bytes(n, n_bins))
        ws = torch.empty(max(wb, 1), dtype=torch.uint8, device="cuda")
        st = L.kp_bin_by_key(C.c_void_p(d_keys.data_ptr()), n, n_bins,
                             C.c_void_p(offs.data_ptr()), C.c_void_p(order.data_ptr()),
                             C.c_void_p(slot.data_ptr()), C.c_void_p(ws.data_ptr()), wb, s)
    _abi.check(st)
    torch.cuda.synchronize()
    return offs.cpu().numpy(), order.cpu().numpy(), slot.cpu().numpy()


@pytest.mark.parametrize("sizes", [((KP_BIN_STATIC_MAX, 3000), (100_003, 37)),
                                   ((500_000, 70_000), (4097, 1))])
def test_static_binning_equals_workspace_binning_twice(sizes):
    """Two calls in a row on the shared array, the second smaller."""
    for k, (n, n_bins) in enumerate(sizes):
        keys = _keys(n, n_bins, seed=17 + k)
        got = _bin(keys, n_bins, static=True)
        want = _bin(keys, n_bins, static=False)
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)
        # and the result is the stable binning itself
        offs, order, slot = got
        valid = (keys >= 0) & (keys < n_bins)
        ref = np.nonzero(valid)[0][np.argsort(keys[valid], kind="stable")]
        np.testing.assert_array_equal(order[:offs[-1]], ref)
        np.testing.assert_array_equal(offs, np.searchsorted(np.sort(keys[valid]),
                                                            np.arange(n_bins + 1)))


def test_static_binning_rejects_sizes_beyond_its_array():
    L = _abi.lib()
    p = C.c_void_p(8)  # never dereferenced: the size check comes first
    st = L.kp_bin_by_key_static(p, KP_BIN_STATIC_MAX + 1, 4, p, None, None, None)
    assert st == _abi.KP_ERR_ARG
    assert b"KP_BIN_STATIC_MAX" in L.kp_last_error() or b"2097152" in L.kp_last_error()


def test_static_binning_of_nothing_clears_the_offsets():
    import torch

    offs = torch.full((6,), -1, dtype=torch.int64, device="cuda")
    _abi.check(_abi.lib().kp_bin_by_key_static(None, 0, 5, C.c_void_p(offs.data_ptr()), None,
                                                None, None))
    torch.cuda.synchronize()
    assert offs.cpu().numpy().tolist() == [0] * 6
