"""The benchmark workloads through the C ABI against the FP64 CPU oracle.

Every test runs a benchmark workload at its full mesh size with a reduced
packet count, on the GPU in FP64, and compares with the oracle on the same
packet ids and seeds:
  - config 2 (3-level AMR, 2,234 blocks) and config 5 (5-level AMR disk,
    28,435 blocks, three observers): per-packet end states bit-identical,
    counters identical, J and images equal to the order of FP64 sums;
  - config 3 (64 bands of uneven packet counts, one empty, captured into a
    hipGraph): end states of every band bit-identical, J per band and the
    64 images equal to the order of FP64 sums;
  - config 4 (2-level AMR adjoint mesh): the one-pass path-replay gradients
    for log rho, albedo and g against the CPU adjoint (rtol 1e-9) and the
    primal functional.
"""
import numpy as np
import pytest

from kratos_polrad_amd import _abi, configs
from kratos_polrad_amd.engine import BandSpec, Engine, RunSpec

pytestmark = pytest.mark.gpu

PERSISTENT = dict(kind=_abi.KP_SCHED_PERSISTENT, block_threads=256, blocks_per_cu=4, chunk=64)


def _cmp_end_states(got, ref):
    if np.array_equal(got, ref):
        return
    bad = np.nonzero(np.any(got != ref, axis=1))[0]
    k = bad[0]
    raise AssertionError(f"{len(bad)} of {len(ref)} packets differ; first slot {k}:\n"
                         f" gpu {got[k].tolist()}\n cpu {ref[k].tolist()}")


def _close(got, ref, rtol):
    np.testing.assert_allclose(got, ref, rtol=rtol, atol=rtol * max(np.abs(ref).max(), 1e-300))


def _gpu_forward(w, run, sched=PERSISTENT):
    import torch

    eng = Engine(w.mesh)
    eng.set_schedule(**sched)
    eng.set_media(w.kappa, w.albedo, w.g)
    out = eng.alloc_outputs(run, end_state=True)
    eng.forward(w.source, run, out)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["counters"] = res["counters"].astype(np.uint64)
    return res


@pytest.mark.parametrize("sched", [PERSISTENT, dict(kind=_abi.KP_SCHED_HISTORY)],
                         ids=["persistent", "history"])
def test_config2_full_mesh_fp64_equals_oracle(oracle_lib, sched):
    w = configs.config2(n_packets=40_000, precision=_abi.KP_F64)
    print(f"config2: {w.mesh.n_blocks} blocks, 40000 packets", flush=True)
    got = _gpu_forward(w, w.run, sched)
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _cmp_end_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"], ref["counters"])
    _close(got["J"], ref["J"], 1e-11)
    _close(got["image"], ref["image"], 1e-11)


def test_config5_full_mesh_fp64_equals_oracle(oracle_lib):
    w = configs.config5(n_packets=20_000, precision=_abi.KP_F64)
    print(f"config5: {w.mesh.n_blocks} blocks, 20000 packets", flush=True)
    got = _gpu_forward(w, w.run)
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _cmp_end_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"], ref["counters"])
    assert int(got["counters"][_abi.KP_C_LOST]) == 0
    _close(got["J"], ref["J"], 1e-11)
    _close(got["image"], ref["image"], 1e-11)


def test_config3_graph_fp64_equals_oracle(oracle_lib):
    """64 bands in one captured launch; band 5 is empty, the long-wavelength
    bands are optically thin."""
    import torch

    w = configs.config3(precision=_abi.KP_F64, scale_packets=4e-3)
    bands = [BandSpec(**{**b.__dict__}) for b in w.run.bands]
    bands[5].n_packets = 0
    run = RunSpec(**{**w.run.__dict__, "bands": bands})
    assert min(b.kappa_scale for b in bands) < 0.1 < max(b.kappa_scale for b in bands)
    print(f"config3: {len(bands)} bands, {run.total_packets} packets", flush=True)
    eng = Engine(w.mesh)
    eng.set_schedule(**PERSISTENT)
    eng.set_media(w.kappa, w.albedo, w.g)
    out = eng.alloc_outputs(run, end_state=True)
    gr = eng.graph(w.source, run, out)
    gr.launch()
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, run, end_state=True)
    _cmp_end_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"].astype(np.uint64), ref["counters"])
    nb = len(bands)
    J, Jr = got["J"].reshape(-1, nb), ref["J"].reshape(-1, nb)
    for b in range(nb):
        if bands[b].n_packets == 0:
            assert not J[:, b].any() and not got["image"][b].any()
            continue
        _close(J[:, b], Jr[:, b], 1e-11)
        _close(got["image"][b], ref["image"][b], 1e-11)
    # a second replay adds the same packets again
    gr.launch()
    torch.cuda.synchronize()
    _close(out["J"].cpu().numpy(), 2 * ref["J"], 1e-11)


def test_config4_adjoint_equals_cpu_adjoint(oracle_lib):
    import torch

    aw = configs.config4(n_packets=20_000)
    w = aw.base
    assert len(set(w.mesh.level.tolist())) == 2  # a real two-level mesh
    print(f"config4: {w.mesh.n_blocks} blocks, 20000 packets", flush=True)
    rng = np.random.default_rng(4)
    W = rng.normal(size=(1, 1, w.run.image_n, w.run.image_n, 3))
    W[..., 0] = 0.0  # the loss weights Q and U only
    eng = Engine(w.mesh)
    eng.set_media(w.kappa, w.albedo, w.g)
    gr = eng.adjoint(w.source, w.run, torch.tensor(W, dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    dl, da, dg, jl = oracle_lib.adjoint(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, W)
    assert abs(gr["jlin"].item() - jl) <= 1e-11 * abs(jl)
    for got, ref in ((gr["d_logrho"], dl), (gr["d_albedo"], da), (gr["d_g"], dg)):
        got = got.cpu().numpy()
        assert np.abs(ref).max() > 0
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())
