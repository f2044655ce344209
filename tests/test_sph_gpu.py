"""GPU parity of the spherical streaming march (kp_sph.h, the reference's
photon_sph.h) against the CPU restatement (oracle/kp_sph_oracle.c).

Per-packet end states (position, direction, crossing count, exit status,
block, cell) must be bit-identical on the same packet ids and seeds; the J
tally differs only by the order of the FP64 atomic sums (relative 1e-12).
"""
import numpy as np
import pytest

import oracle
from kratos_polrad_amd import _abi, configs
from kratos_polrad_amd.engine import BandSpec, Engine, RunSpec

pytestmark = pytest.mark.gpu

SRC = (0.25, 0.15, 0.1)
SCHEDULES = [dict(kind=_abi.KP_SCHED_PERSISTENT, block_threads=256, blocks_per_cu=4, chunk=64),
             dict(kind=_abi.KP_SCHED_PERSISTENT, block_threads=128, blocks_per_cu=2, chunk=16,
                  event_batch=1),
             dict(kind=_abi.KP_SCHED_HISTORY, block_threads=256, blocks_per_cu=1, chunk=1),
             dict(kind=_abi.KP_SCHED_PERSISTENT, block_threads=256, blocks_per_cu=4, chunk=64,
                  coherent=1),
             dict(kind=_abi.KP_SCHED_HISTORY, block_threads=256, blocks_per_cu=1, chunk=1,
                  coherent=1)]


def _shell_mesh():
    return configs.sph_mesh(root_blocks=(4, 2, 4), n_cell=8, r_lim=(0.1, 1.0), r_ratio=-1.0,
                            refine=((1, (0.2, 1.0, 0.2), (0.4, 1.6, 0.9)),
                                    (2, (0.25, 1.2, 0.4), (0.3, 1.4, 0.6))))


def _wedge_mesh():
    # uniform r, a theta band and a phi wedge: packets also leave through
    # theta and phi faces
    return configs.sph_mesh(root_blocks=(2, 2, 2), n_cell=8, r_lim=(0.1, 1.0), r_ratio=1.0,
                            theta=(0.6, 2.2), phi=(-0.4, 1.4),
                            refine=((1, (0.2, 1.0, 0.2), (0.5, 1.6, 0.8)),))


def _gpu(mesh, run, sched=None, weight=1.0, src=SRC):
    import torch

    eng = Engine(mesh)
    if sched is not None:
        eng.set_schedule(**sched)
    out = eng.alloc_outputs(run, end_state=True)
    eng.forward(src, run, out, weight=weight)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["counters"] = res["counters"].astype(np.uint64)
    return res


def _check(got, ref):
    es, er = got["end_state"], ref["end_state"]
    if not np.array_equal(es, er):
        bad = np.nonzero(np.any(es != er, axis=1))[0]
        k = bad[0]
        raise AssertionError(f"{len(bad)} of {len(er)} packets differ; first slot {k}:\n"
                             f" gpu {es[k].tolist()}\n cpu {er[k].tolist()}")
    assert np.array_equal(got["counters"], ref["counters"])
    assert np.allclose(got["J"], ref["J"], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("sched", SCHEDULES, ids=["persist256", "persist128_b1", "history",
                                                   "persist_coherent", "history_coherent"])
def test_sph_bit_exact_shell(sched):
    m = _shell_mesh()
    run = RunSpec(bands=[BandSpec(n_packets=20000, kappa_scale=0.0),
                         BandSpec(n_packets=7000, kappa_scale=0.0)], seed=9, packet_offset=123,
                  step_max=100000)
    got = _gpu(m, run, sched, weight=1.5)
    ref = oracle.sph_forward(m, SRC, run, weight=1.5, end_state=True)
    _check(got, ref)
    c = got["counters"]
    assert c[_abi.KP_C_PACKETS] == 27000 and c[_abi.KP_C_ESCAPED] == 27000
    assert c[_abi.KP_C_LOST] == 0


def test_sph_bit_exact_wedge_and_step_budget():
    m = _wedge_mesh()
    src = (0.3, 0.25, 0.05)
    run = RunSpec(bands=[BandSpec(n_packets=15000, kappa_scale=0.0)], seed=3, step_max=12)
    got = _gpu(m, run, src=src)
    ref = oracle.sph_forward(m, src, run, end_state=True)
    _check(got, ref)
    st = got["end_state"][:, _abi.KP_ES_STATUS]
    assert np.any(st == _abi.KP_END_STEPS) and np.any(st == _abi.KP_END_ESCAPED)


@pytest.mark.parametrize("coherent", [0, 1])
def test_sph_graph_replay_matches_forwards(coherent):
    import torch

    m = _shell_mesh()
    run = RunSpec(bands=[BandSpec(n_packets=8000, kappa_scale=0.0),
                         BandSpec(n_packets=3000, kappa_scale=0.0)], seed=5, step_max=100000)
    eng = Engine(m)
    eng.set_schedule(coherent=coherent)
    out_g = eng.alloc_outputs(run)
    g = eng.graph(SRC, run, out_g, n_chunks=2)
    g.launch()
    out_e = eng.alloc_outputs(run)
    eng.forward(SRC, run, out_e)
    run2 = RunSpec(**{**run.__dict__, "packet_offset": 8000})  # chunk stride: max band count
    eng.forward(SRC, run2, out_e)
    torch.cuda.synchronize()
    assert torch.allclose(out_g["J"], out_e["J"], rtol=1e-12, atol=0)
    assert torch.equal(out_g["counters"], out_e["counters"])
    del g


def test_sph_rejections():
    import torch

    m = _shell_mesh()
    eng = Engine(m)
    ok = RunSpec(bands=[BandSpec(n_packets=10, kappa_scale=0.0)], step_max=100000)
    for bad in (RunSpec(bands=[BandSpec(n_packets=10, kappa_scale=1.0)]),
                RunSpec(bands=[BandSpec(n_packets=10, kappa_scale=0.0)], incl_deg=(90.0,)),
                RunSpec(bands=[BandSpec(n_packets=10, kappa_scale=0.0)],
                        precision=_abi.KP_F32)):
        with pytest.raises(_abi.KPError) as ei:
            eng.forward(SRC, bad, eng.alloc_outputs(bad))
        assert ei.value.status == _abi.KP_ERR_UNSUPPORTED
    with pytest.raises(_abi.KPError):
        eng.forward((0.01, 0.0, 0.0), ok, eng.alloc_outputs(ok))  # inside the inner hole
    with pytest.raises(_abi.KPError) as ei:
        eng.adjoint(SRC, ok, torch.zeros((1, 1, 128, 128, 3), dtype=torch.float64,
                                          device="cuda"))
    assert ei.value.status == _abi.KP_ERR_UNSUPPORTED
