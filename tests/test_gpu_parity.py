"""GPU parity: the HIP kernels through the C ABI against the FP64 CPU oracle.

FP64 path: per-packet end states (position, direction, Stokes I/Q/U, event
and crossing counts, exit status, block, cell) must be bit-identical to the
oracle's on the same packet ids and seeds; tallies and images differ only by
the order of the FP64 atomic sums (relative 1e-12).
FP32 mixed path: tallies and images must agree with the FP64 path within the
combined Monte Carlo noise (3 sigma, batch means) -- its random numbers are
the same Philox words at 23-bit resolution, so histories track the FP64 ones.
"""
import numpy as np
import pytest

from kratos_polrad_amd import _abi, configs
from kratos_polrad_amd.engine import BandSpec, Engine, RunSpec
from kratos_polrad_amd.mesh import Mesh

pytestmark = pytest.mark.gpu


SCHEDULES = [dict(kind=_abi.KP_SCHED_PERSISTENT, block_threads=256, blocks_per_cu=4, chunk=64),
             dict(kind=_abi.KP_SCHED_HISTORY, block_threads=256, blocks_per_cu=1, chunk=1),
             # packets marched in emission-direction order: same histories
             dict(kind=_abi.KP_SCHED_PERSISTENT, block_threads=256, blocks_per_cu=4, chunk=64,
                  coherent=1),
             dict(kind=_abi.KP_SCHED_HISTORY, block_threads=256, blocks_per_cu=1, chunk=1,
                  coherent=1)]


def _run_gpu(w, end_state=True, precision=None, sched=None):
    import torch

    run = w.run
    if precision is not None:
        run = RunSpec(**{**run.__dict__, "precision": precision})
    eng = Engine(w.mesh)
    if sched is not None:
        eng.set_schedule(**sched)
    eng.set_media(w.kappa, w.albedo, w.g)
    out = eng.alloc_outputs(run, end_state=end_state)
    eng.forward(w.source, run, out)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["counters"] = res["counters"].astype(np.uint64)
    return res


def _compare_end_states(got, ref):
    if np.array_equal(got, ref):
        return
    bad = np.nonzero(np.any(got != ref, axis=1))[0]
    k = bad[0]
    raise AssertionError(
        f"{len(bad)} of {len(ref)} packets differ; first slot {k}:\n gpu {got[k].tolist()}\n"
        f" cpu {ref[k].tolist()}")


def _amr_workload(n_packets, precision=_abi.KP_F64, model=_abi.KP_HG_RAYLEIGH):
    """Static refinement around the centre: levels 0/1/2 with 2:1 balance."""
    mesh = Mesh(root_blocks=(4, 4, 4), n_cell=8)
    mesh.refine_region(1, (0.3, 0.3, 0.3), (0.7, 0.7, 0.7))
    mesh.refine_region(2, (0.42, 0.45, 0.4), (0.58, 0.55, 0.6))
    mesh.finalize()
    xc = mesh.cell_centers()
    r2 = np.sum((xc - 0.5) ** 2, axis=1)
    kappa = (0.6 / (r2 + 0.01)).astype(np.float32)
    kappa[r2 > 0.2] = 0.0
    rng = np.random.default_rng(3)
    albedo = rng.uniform(0.4, 0.9, mesh.n_cells).astype(np.float32)
    g = rng.uniform(0.2, 0.7, mesh.n_cells).astype(np.float32)
    run = RunSpec(bands=[BandSpec(n_packets=n_packets, cell_albedo=True, cell_g=True)],
                  seed=1234, scatter_model=model, p_max=0.5, precision=precision,
                  incl_deg=(90.0, 37.0), image_n=64, image_half=0.6, rr_weight=0.05,
                  rr_survive=0.25)
    return configs.Workload("amr_test", mesh, kappa, albedo, g, (0.51, 0.49, 0.5), run)


@pytest.mark.parametrize("sched", SCHEDULES, ids=["persistent", "history", "persistent_coherent",
                                                   "history_coherent"])
def test_f64_config1_bit_exact(oracle_lib, sched):
    w = configs.config1(n_packets=20000)
    got = _run_gpu(w, sched=sched)
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _compare_end_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"], ref["counters"])
    np.testing.assert_allclose(got["J"], ref["J"], rtol=1e-12, atol=1e-12 * ref["J"].max())
    np.testing.assert_allclose(got["image"], ref["image"], rtol=1e-11,
                               atol=1e-12 * np.abs(ref["image"]).max())


@pytest.mark.parametrize("sched", SCHEDULES, ids=["persistent", "history", "persistent_coherent",
                                                   "history_coherent"])
def test_f64_amr_bit_exact(oracle_lib, sched):
    w = _amr_workload(8000)
    got = _run_gpu(w, sched=sched)
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _compare_end_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"], ref["counters"])
    np.testing.assert_allclose(got["J"], ref["J"], rtol=1e-12, atol=1e-12 * ref["J"].max())
    np.testing.assert_allclose(got["image"], ref["image"], rtol=1e-11,
                               atol=1e-12 * np.abs(ref["image"]).max())


def _batch_stats(eng_run, n_batches):
    """mean and standard error over independent batches (different offsets)."""
    vals = np.stack(eng_run)
    return vals.mean(axis=0), vals.std(axis=0, ddof=1) / np.sqrt(n_batches)


def test_f32_matches_f64_within_noise():
    import torch

    w = configs.config1(n_packets=50000)
    eng = Engine(w.mesh)
    eng.set_media(w.kappa, w.albedo, w.g)
    res = {0: [], 1: []}
    nb = 8
    for prec in (_abi.KP_F64, _abi.KP_F32):
        for b in range(nb):
            run = RunSpec(**{**w.run.__dict__, "precision": prec, "packet_offset": b * 50000})
            out = eng.alloc_outputs(run)
            eng.forward(w.source, run, out)
            torch.cuda.synchronize()
            img = out["image"].cpu().numpy()[0, 0]
            # coarse 8x8 image of I, Q, U
            co = img.reshape(8, 16, 8, 16, 3).sum(axis=(1, 3))
            res[prec].append(co.ravel())
    m64, e64 = _batch_stats(res[0], nb)
    m32, e32 = _batch_stats(res[1], nb)
    d = np.abs(m32 - m64)
    # same random numbers: FP32 and FP64 differ far less than the MC noise
    assert np.all(d <= 3.0 * np.sqrt(e64 ** 2 + e32 ** 2) + 1e-9)


def test_f32_persistent_equals_history():
    """Both schedules run the same FP32 per-packet arithmetic: end states agree."""
    w = _amr_workload(4000, precision=_abi.KP_F32)
    a = _run_gpu(w, sched=SCHEDULES[0])
    b = _run_gpu(w, sched=SCHEDULES[1])
    np.testing.assert_array_equal(a["end_state"], b["end_state"])
    assert np.array_equal(a["counters"], b["counters"])
    np.testing.assert_allclose(a["J"], b["J"], rtol=2e-6, atol=1e-9 * np.abs(b["J"]).max())


@pytest.mark.parametrize("bt", [256, 128])
def test_f32_lockstep_coherent_equals_persistent(bt):
    """The mixed persistent schedule in emission-direction order (waves
    refilled together, J deposits of a run of lanes in one cell summed before
    one atomic, no per-lane accumulator) runs the same FP32 histories: end
    states and counters identical, image identical up to summation order, J up
    to the accumulator's FP32 rounding.  Two bands, so runs must not merge
    different bands' elements."""
    w = _amr_workload(6000, precision=_abi.KP_F32)
    run = RunSpec(**{**w.run.__dict__, "bands": [
        BandSpec(n_packets=6000, cell_albedo=True, cell_g=True),
        BandSpec(n_packets=4000, kappa_scale=0.2, cell_albedo=True, cell_g=True)]})
    w = w.__class__(**{**w.__dict__, "run": run})
    a = _run_gpu(w, sched=SCHEDULES[0])
    b = _run_gpu(w, sched=dict(kind=_abi.KP_SCHED_PERSISTENT, block_threads=bt, blocks_per_cu=4,
                               chunk=64, coherent=1))
    np.testing.assert_array_equal(a["end_state"], b["end_state"])
    assert np.array_equal(a["counters"], b["counters"])
    np.testing.assert_allclose(b["image"], a["image"], rtol=1e-11,
                               atol=1e-12 * np.abs(a["image"]).max())
    np.testing.assert_allclose(b["J"], a["J"], rtol=2e-6, atol=1e-9 * np.abs(a["J"]).max())


@pytest.mark.parametrize("chunk,bt,batch", [(1, 64, 1), (16, 256, 16), (1000, 128, 64)])
def test_persistent_refill_processes_each_packet_once(chunk, bt, batch):
    """Wave slot ranges never overlap and batched action phases keep every
    history intact: any chunk / batch size gives every packet exactly once,
    with the same end states as the one-thread-per-packet run."""
    w = _amr_workload(3000, precision=_abi.KP_F32)
    ref = _run_gpu(w, sched=SCHEDULES[1])
    got = _run_gpu(w, sched=dict(kind=_abi.KP_SCHED_PERSISTENT, block_threads=bt,
                                 blocks_per_cu=2, chunk=chunk, event_batch=batch))
    assert int(got["counters"][_abi.KP_C_PACKETS]) == 3000
    np.testing.assert_array_equal(got["end_state"], ref["end_state"])


def _periodic_workload(n_packets, precision=_abi.KP_F64):
    """Periodic in x and y, refined off-centre so packets cross level jumps
    and the periodic faces; J only (images need a bounded domain)."""
    mesh = Mesh(root_blocks=(3, 2, 2), n_cell=8, x_max=(1.5, 1.0, 1.0), periodic=(1, 1, 0))
    mesh.refine_region(1, (1.2, 0.0, 0.3), (1.5, 0.4, 0.7))
    mesh.finalize()
    xc = mesh.cell_centers()
    kappa = (1.5 + 0.6 * np.sin(7 * xc[:, 0]) + 0.6 * np.cos(5 * xc[:, 1])).astype(np.float32)
    albedo = np.full(mesh.n_cells, 0.9, dtype=np.float32)
    g = np.full(mesh.n_cells, 0.4, dtype=np.float32)
    run = RunSpec(bands=[BandSpec(n_packets=n_packets, albedo=0.9, g=0.4)], seed=77,
                  scatter_model=_abi.KP_HG_RAYLEIGH, p_max=0.5, precision=precision,
                  step_max=5000, incl_deg=())
    return configs.Workload("periodic_test", mesh, kappa, albedo, g, (1.37, 0.05, 0.5), run)


@pytest.mark.parametrize("sched", SCHEDULES, ids=["persistent", "history", "persistent_coherent",
                                                   "history_coherent"])
def test_f64_periodic_bit_exact(oracle_lib, sched):
    w = _periodic_workload(4000)
    got = _run_gpu(w, sched=sched)
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _compare_end_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"], ref["counters"])
    # packets wrap: some end far outside the refined box, none are lost
    assert int(got["counters"][_abi.KP_C_LOST]) == 0
    np.testing.assert_allclose(got["J"], ref["J"], rtol=1e-12, atol=1e-12 * ref["J"].max())


def test_f32_periodic_tracks_f64():
    w64 = _periodic_workload(20000)
    w32 = _periodic_workload(20000, precision=_abi.KP_F32)
    a = _run_gpu(w64, end_state=False)
    b = _run_gpu(w32, end_state=False)
    assert int(b["counters"][_abi.KP_C_PACKETS]) == 20000
    assert int(b["counters"][_abi.KP_C_LOST]) == 0
    # same random numbers: totals agree far inside the Monte Carlo noise
    assert abs(b["J"].sum() / a["J"].sum() - 1) < 2e-3
