"""The block-bucketed packet queue (KP_SCHED_BLOCKQ) through the C ABI.

Its passes sort the packets' records by (band, block), march every packet of
a block with the block's kappa_rho and a J tile in LDS until it leaves the
block, and compact the leavers into the next pass by wave ballot; the
packets still alive after the last pass finish one thread each.  It must run
the same histories as the other schedules:
  - FP64: end states bit-identical to the CPU oracle on meshes with level
    jumps, periodic faces, per-cell albedo and g, two observers and several
    bands (one empty, the first ones on the persistent kernel in the same
    launch), for few and many passes, several task sizes and packets sent
    back to the queue every few steps; J and images equal to the order of
    FP64 sums;
  - FP32: end states identical to the persistent mixed-precision kernel's,
    J to the FP32 rounding of the queue's per-task LDS tile;
  - a captured graph replays the same sequence, and a graph of several
    chunks equals the chunks launched one by one.
"""
import numpy as np
import pytest

from kratos_polrad_amd import _abi, configs
from kratos_polrad_amd.engine import BandSpec, Engine, RunSpec
from kratos_polrad_amd.mesh import Mesh

pytestmark = pytest.mark.gpu

PERSISTENT = dict(kind=_abi.KP_SCHED_PERSISTENT, block_threads=256, blocks_per_cu=4, chunk=64)


def _bq(passes=0, chunk=0, from_band=0, bpc=4, yld=0):
    return dict(kind=_abi.KP_SCHED_BLOCKQ, block_threads=256, blocks_per_cu=bpc, chunk=64,
                queue_from_band=from_band, queue_passes=passes, queue_chunk=chunk,
                queue_yield=yld)


def _amr_workload(n_packets, precision=_abi.KP_F64, bands=None):
    """Levels 0/1/2 with 2:1 balance, per-cell albedo and g, two observers."""
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
    if bands is None:
        bands = [BandSpec(n_packets=n_packets, cell_albedo=True, cell_g=True)]
    run = RunSpec(bands=bands, seed=1234, scatter_model=_abi.KP_HG_RAYLEIGH, p_max=0.5,
                  precision=precision, incl_deg=(90.0, 37.0), image_n=64, image_half=0.6,
                  rr_weight=0.05, rr_survive=0.25)
    return configs.Workload("amr_test", mesh, kappa, albedo, g, (0.51, 0.49, 0.5), run)


def _periodic_workload(n_packets, precision=_abi.KP_F64):
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


def _run(w, sched, run=None, end_state=True):
    import torch

    run = run or w.run
    eng = Engine(w.mesh)
    eng.set_schedule(**sched)
    eng.set_media(w.kappa, w.albedo, w.g)
    out = eng.alloc_outputs(run, end_state=end_state)
    eng.forward(w.source, run, out)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["counters"] = res["counters"].astype(np.uint64)
    return res


def _same_states(got, ref):
    if np.array_equal(got, ref):
        return
    bad = np.nonzero(np.any(got != ref, axis=1))[0]
    k = bad[0]
    raise AssertionError(f"{len(bad)} of {len(ref)} packets differ; first slot {k}:\n"
                         f" gpu {got[k].tolist()}\n ref {ref[k].tolist()}")


def _close(a, b, rtol):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("passes,chunk,yld", [(0, 0, 0), (2, 64, 0), (64, 256, 0),
                                              (1, 4096, 0), (200, 256, 3)],
                         ids=["default", "few_passes", "many_passes", "one_pass_big_tasks",
                              "yield_every_3_steps"])
def test_f64_amr_bit_exact(oracle_lib, passes, chunk, yld):
    w = _amr_workload(8000)
    print(f"blockq amr f64 passes={passes} chunk={chunk} yield={yld}", flush=True)
    got = _run(w, _bq(passes, chunk, yld=yld))
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _same_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"], ref["counters"])
    _close(got["J"], ref["J"], 1e-12)
    _close(got["image"], ref["image"], 1e-11)


def test_f64_bands_split_between_persistent_and_queue(oracle_lib):
    """Bands 0-1 on the persistent kernel, 2-4 in the queue (band 3 empty),
    uneven counts, different extinction scales: one launch."""
    bands = [BandSpec(n_packets=n, kappa_scale=k, cell_albedo=True, cell_g=True)
             for n, k in ((3000, 1.0), (500, 2.0), (2500, 0.3), (0, 1.0), (4000, 0.05))]
    w = _amr_workload(0, bands=bands)
    print("blockq bands split", flush=True)
    got = _run(w, _bq(passes=8, from_band=2))
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _same_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"], ref["counters"])
    _close(got["J"], ref["J"], 1e-12)
    _close(got["image"], ref["image"], 1e-11)


def test_f64_periodic_bit_exact(oracle_lib):
    w = _periodic_workload(4000)
    print("blockq periodic", flush=True)
    got = _run(w, _bq(passes=16, chunk=512))
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _same_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"], ref["counters"])
    assert int(got["counters"][_abi.KP_C_LOST]) == 0
    _close(got["J"], ref["J"], 1e-12)


def test_f64_config1_bit_exact(oracle_lib):
    w = configs.config1(n_packets=20000)
    print("blockq config1", flush=True)
    got = _run(w, _bq())
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _same_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"], ref["counters"])
    _close(got["J"], ref["J"], 1e-12)
    _close(got["image"], ref["image"], 1e-11)


@pytest.mark.parametrize("passes,yld,n_obs", [(3, 0, 2), (48, 0, 2), (100, 5, 2), (20, 0, 1)],
                         ids=["3_passes", "48_passes", "100_passes_yield_5", "one_observer"])
def test_f32_equals_persistent_mixed(passes, yld, n_obs):
    """Two observers: the peel-off kernel's waves claim rays from a counter;
    one observer: the rays in grid-stride order over a large grid."""
    w = _amr_workload(6000, precision=_abi.KP_F32)
    if n_obs == 1:
        w = configs.Workload(w.name, w.mesh, w.kappa, w.albedo, w.g, w.source,
                             RunSpec(**{**w.run.__dict__, "incl_deg": (90.0,)}))
    print(f"blockq f32 passes={passes} yield={yld} observers={n_obs}", flush=True)
    a = _run(w, PERSISTENT)
    b = _run(w, _bq(passes=passes, chunk=512, yld=yld))
    _same_states(b["end_state"], a["end_state"])
    assert np.array_equal(a["counters"], b["counters"])
    _close(b["image"], a["image"], 1e-9)
    # the queue sums each task's J in an FP32 LDS tile before the FP64 atomic
    _close(b["J"], a["J"], 3e-5)


def test_graph_replays_the_queue():
    import torch

    w = _amr_workload(5000, precision=_abi.KP_F32)
    print("blockq graph", flush=True)
    eng = Engine(w.mesh)
    eng.set_schedule(**_bq(passes=6, chunk=256))
    eng.set_media(w.kappa, w.albedo, w.g)
    eager = eng.alloc_outputs(w.run)
    eng.forward(w.source, w.run, eager)
    out = eng.alloc_outputs(w.run)
    g = eng.graph(w.source, w.run, out)
    for k in range(3):
        g.launch()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out["counters"].cpu().numpy(),
                                      (k + 1) * eager["counters"].cpu().numpy())
        # J: each task's FP32 LDS tile sums its packets in LDS-atomic order,
        # which differs between launches (FP32 rounding, ~1e-7)
        _close(out["J"].cpu().numpy(), (k + 1) * eager["J"].cpu().numpy(), 1e-5)
        _close(out["image"].cpu().numpy(), (k + 1) * eager["image"].cpu().numpy(), 1e-11)


@pytest.mark.parametrize("precision", [_abi.KP_F64, _abi.KP_F32], ids=["f64", "f32"])
def test_graph_of_chunks_equals_eager_chunks(precision):
    """A graph of three chunks (packet ids offset per chunk, as the large
    workload streams them) against the same chunks launched one by one."""
    import torch

    w = _amr_workload(3000, precision=precision)
    print(f"blockq graph of 3 chunks, precision {precision}", flush=True)
    eng = Engine(w.mesh)
    eng.set_schedule(**_bq(passes=6, chunk=512))
    eng.set_media(w.kappa, w.albedo, w.g)
    stride = w.run.total_packets
    ref = eng.alloc_outputs(w.run)
    for k in range(3):
        eng.forward(w.source, RunSpec(**{**w.run.__dict__, "packet_offset": k * stride}), ref)
    out = eng.alloc_outputs(w.run)
    g = eng.graph(w.source, w.run, out, n_chunks=3)
    g.launch()
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    exp = {k: v.cpu().numpy() for k, v in ref.items()}
    assert np.array_equal(got["counters"], exp["counters"])
    # J: FP64 sums in another order; FP32 also the tile's LDS-atomic order
    _close(got["J"], exp["J"], 1e-12 if precision == _abi.KP_F64 else 1e-5)
    _close(got["image"], exp["image"], 1e-11)


def test_config3_graph_queue_bands_fp64_equal_oracle(oracle_lib):
    """The 64-band workload with its optically thin half in the queue, in one
    captured launch."""
    import torch

    w = configs.config3(precision=_abi.KP_F64, scale_packets=2e-3)
    run = w.run
    first = next(b for b, s in enumerate(run.bands) if s.kappa_scale <= 0.5)
    print(f"blockq config3: queue from band {first}, {run.total_packets} packets", flush=True)
    eng = Engine(w.mesh)
    eng.set_schedule(**_bq(from_band=first))
    eng.set_media(w.kappa, w.albedo, w.g)
    out = eng.alloc_outputs(run, end_state=True)
    gr = eng.graph(w.source, run, out)
    gr.launch()
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, run, end_state=True)
    _same_states(got["end_state"], ref["end_state"])
    assert np.array_equal(got["counters"].astype(np.uint64), ref["counters"])
    _close(got["J"], ref["J"], 1e-11)
    _close(got["image"], ref["image"], 1e-11)


def test_schedule_rejections():
    from kratos_polrad_amd import _abi as A

    w = _amr_workload(10)
    eng = Engine(w.mesh)
    with pytest.raises(RuntimeError):
        eng.set_schedule(**{**_bq(), "queue_passes": -1})
    with pytest.raises(RuntimeError):
        eng.set_schedule(**{**_bq(), "queue_chunk": (1 << 20) + 1})
    with pytest.raises(RuntimeError):
        eng.set_schedule(**{**_bq(), "queue_yield": -2})
    big = Mesh(root_blocks=(1, 1, 1), n_cell=32).finalize()
    e2 = Engine(big)
    with pytest.raises(RuntimeError, match="4096"):
        e2.set_schedule(**_bq())
    assert A.KP_SCHED_BLOCKQ == 2
