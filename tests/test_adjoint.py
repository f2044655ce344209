"""Adjoint by path replay: the FP64 CPU adjoint against common-random-number
finite differences, and the GPU kernel against the CPU adjoint."""
import numpy as np
import pytest

from kratos_polrad_amd import _abi
from kratos_polrad_amd.engine import BandSpec, Engine, RunSpec
from kratos_polrad_amd.mesh import Mesh


def _setup(n_packets=40000, offset=0):
    m = Mesh(root_blocks=(2, 2, 2), n_cell=8)
    m.refine_region(1, (0.35, 0.35, 0.35), (0.65, 0.65, 0.65))
    m.finalize()
    xc = m.cell_centers()
    r2 = ((xc - 0.5) ** 2).sum(1)
    rng = np.random.default_rng(0)
    kappa = (3.0 * np.exp(-r2 / 0.05)).astype(np.float32)
    alb = rng.uniform(0.5, 0.9, m.n_cells).astype(np.float32)
    g = rng.uniform(0.2, 0.6, m.n_cells).astype(np.float32)
    run = RunSpec(bands=[BandSpec(n_packets=n_packets, cell_albedo=True, cell_g=True)], seed=5,
                  scatter_model=_abi.KP_HG_RAYLEIGH, p_max=0.5, incl_deg=(70.0, 20.0),
                  image_n=16, image_half=0.6, rr_weight=0.05, rr_survive=0.3,
                  packet_offset=offset)
    W = rng.normal(size=(1, 2, 16, 16, 3))
    return m, kappa, alb, g, run, W, r2 < 0.09


SRC = (0.47, 0.52, 0.5)


def test_replay_reproduces_forward_functional(oracle_lib):
    m, kappa, alb, g, run, W, sel = _setup(20000)
    dl, da, dg, jl = oracle_lib.adjoint(m, kappa, alb, g, SRC, run, W)
    fw = oracle_lib.forward(m, kappa, alb, g, SRC, run, J=False)
    assert abs(jl - (fw["image"] * W).sum()) <= 1e-9 * abs(jl)


@pytest.mark.parametrize("param", ["logrho", "albedo", "g"])
def test_adjoint_matches_crn_finite_differences(oracle_lib, param):
    """Directional derivative over the central cells, mean over 6 batches of
    independent packets: adjoint and CRN central differences agree within
    3 combined standard errors and 5 %."""
    ad, fd = [], []
    eps = 1e-2
    for b in range(6):
        m, kappa, alb, g, run, W, sel = _setup(30000, offset=b * 30000)
        dl, da, dg, jl = oracle_lib.adjoint(m, kappa, alb, g, SRC, run, W)

        def J(k, a, gg):
            return (oracle_lib.forward(m, k, a, gg, SRC, run, J=False)["image"] * W).sum()

        if param == "logrho":
            ad.append(dl[sel].sum())
            kp = (kappa * np.where(sel, np.exp(eps), 1.0)).astype(np.float32)
            km = (kappa * np.where(sel, np.exp(-eps), 1.0)).astype(np.float32)
            fd.append((J(kp, alb, g) - J(km, alb, g)) / (2 * eps))
        elif param == "albedo":
            ad.append(da[sel].sum())
            fd.append((J(kappa, np.where(sel, alb + eps, alb).astype(np.float32), g)
                       - J(kappa, np.where(sel, alb - eps, alb).astype(np.float32), g)) / (2 * eps))
        else:
            ad.append(dg[sel].sum())
            fd.append((J(kappa, alb, np.where(sel, g + eps, g).astype(np.float32))
                       - J(kappa, alb, np.where(sel, g - eps, g).astype(np.float32))) / (2 * eps))
    ad, fd = np.array(ad), np.array(fd)
    se = np.sqrt(ad.var(ddof=1) / len(ad) + fd.var(ddof=1) / len(fd))
    assert abs(ad.mean() - fd.mean()) < max(3 * se, 0.05 * abs(fd.mean()))


@pytest.mark.gpu
@pytest.mark.parametrize("kind,one_pass", [(_abi.KP_SCHED_PERSISTENT, 1),
                                           (_abi.KP_SCHED_PERSISTENT, 0),
                                           (_abi.KP_SCHED_HISTORY, 0)],
                         ids=["persistent_one_pass", "persistent_two_pass", "history"])
def test_gpu_adjoint_matches_cpu_adjoint(oracle_lib, kind, one_pass):
    import torch

    m, kappa, alb, g, run, W, sel = _setup(20000)
    eng = Engine(m)
    eng.set_schedule(kind=kind, peel_queue=one_pass)
    eng.set_media(kappa, alb, g)
    run64 = RunSpec(**{**run.__dict__, "precision": _abi.KP_F64})
    w = torch.tensor(W, dtype=torch.float64, device="cuda")
    gr = eng.adjoint(SRC, run64, w)
    torch.cuda.synchronize()
    dl, da, dg, jl = oracle_lib.adjoint(m, kappa, alb, g, SRC, run64, W)
    assert abs(gr["jlin"].item() - jl) <= 1e-11 * abs(jl)
    for got, ref in ((gr["d_logrho"], dl), (gr["d_albedo"], da), (gr["d_g"], dg)):
        got = got.cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())


@pytest.mark.gpu
def test_autograd_function_uses_the_adjoint():
    import torch

    from kratos_polrad_amd.autograd import polarized_image

    m, kappa, alb, g, run, W, sel = _setup(20000)
    eng = Engine(m)
    eng.set_media(kappa, alb, g)
    lr = torch.zeros(m.n_cells, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(alb, dtype=torch.float64, requires_grad=True)
    gg = torch.tensor(g, dtype=torch.float64, requires_grad=True)
    img = polarized_image(eng, SRC, run, kappa, lr, a, gg)
    obs = torch.zeros_like(img)
    loss = ((img[..., 1:] - obs[..., 1:]) ** 2).sum()
    loss.backward()
    # the same gradient straight from kp_adjoint with dL/dimage = 2 (S - obs) on Q, U
    w = torch.zeros_like(img)
    w[..., 1:] = 2 * (img.detach()[..., 1:] - obs[..., 1:])
    ref = eng.adjoint(SRC, RunSpec(**{**run.__dict__, "precision": _abi.KP_F64}), w.cuda())
    torch.testing.assert_close(lr.grad, ref["d_logrho"].cpu(), rtol=1e-9, atol=1e-12)
    torch.testing.assert_close(gg.grad, ref["d_g"].cpu(), rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
def test_autograd_two_images_each_use_their_own_medium():
    """Two images of different media in one loss: each backward replays its
    own medium (the engine holds the second one when backward starts)."""
    import torch

    from kratos_polrad_amd.autograd import polarized_image

    m, kappa, alb, g, run, W, sel = _setup(20000)
    eng = Engine(m)
    rng = np.random.default_rng(21)
    lr1 = torch.zeros(m.n_cells, dtype=torch.float64, requires_grad=True)
    lr2 = torch.tensor(0.3 * rng.normal(size=m.n_cells), dtype=torch.float64,
                       requires_grad=True)
    a = torch.tensor(alb, dtype=torch.float64, requires_grad=True)
    gg = torch.tensor(g, dtype=torch.float64)
    w1 = torch.tensor(W, dtype=torch.float64)
    w2 = torch.tensor(rng.normal(size=W.shape), dtype=torch.float64)
    img1 = polarized_image(eng, SRC, run, kappa, lr1, a, gg)
    img2 = polarized_image(eng, SRC, run, kappa, lr2, a, gg)
    ((img1 * w1).sum() + (img2 * w2).sum()).backward()
    run64 = RunSpec(**{**run.__dict__, "precision": _abi.KP_F64})
    want_a = torch.zeros(m.n_cells, dtype=torch.float64)
    for lr, w in ((lr1, w1), (lr2, w2)):
        k = (kappa.astype(np.float64) * np.exp(lr.detach().numpy())).astype(np.float32)
        eng.set_media(k, alb, g)
        ref = eng.adjoint(SRC, run64, w.cuda())
        torch.cuda.synchronize()
        torch.testing.assert_close(lr.grad, ref["d_logrho"].cpu(), rtol=1e-9, atol=1e-12)
        want_a += ref["d_albedo"].cpu()
    torch.testing.assert_close(a.grad, want_a, rtol=1e-9, atol=1e-12)


@pytest.mark.gpu
def test_gpu_f32_adjoint_tracks_f64_adjoint():
    """The FP32 replay draws the same Philox words: its region gradients agree
    with the FP64 replay far inside the Monte Carlo noise (0.5 %)."""
    import torch

    m, kappa, alb, g, run, W, sel = _setup(40000)
    eng = Engine(m)
    eng.set_media(kappa, alb, g)
    w = torch.tensor(W, dtype=torch.float64, device="cuda")
    out = {}
    for prec in (_abi.KP_F64, _abi.KP_F32):
        gr = eng.adjoint(SRC, RunSpec(**{**run.__dict__, "precision": prec}), w)
        torch.cuda.synchronize()
        out[prec] = {k: v.cpu().numpy() for k, v in gr.items()}
    for k in ("d_logrho", "d_albedo", "d_g"):
        a, b = out[_abi.KP_F32][k][sel].sum(), out[_abi.KP_F64][k][sel].sum()
        assert abs(a - b) <= 5e-3 * abs(b), (k, a, b)
    assert abs(out[_abi.KP_F32]["jlin"][0] / out[_abi.KP_F64]["jlin"][0] - 1) < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [_abi.KP_F32, _abi.KP_F64], ids=["f32", "f64"])
@pytest.mark.parametrize("n_incl", [1, 2])
@pytest.mark.parametrize("vertex_slots,wps", [(0, 0), (3, 0), (1, 1)],
                         ids=["default", "vq3_fallback", "vq1_fallback_wps1"])
def test_gpu_one_pass_replay_equals_two_pass(prec, n_incl, vertex_slots, wps):
    """The one-pass replay (deposits deferred: main-path vertices and
    peel-off rays queued, deposited by separate kernels once each packet's
    sum of peel-off contributions is known) deposits the same gradients as
    the two-pass replay: same histories, same arithmetic per deposit in
    either precision, B = Atot - C instead of a running difference,
    different summation order.  Short vertex queues (1 or 3 records per packet) send packets to
    the two-pass kernel, which must neither lose nor double-count them (their
    queued records are skipped)."""
    import torch

    m, kappa, alb, g, run, W, sel = _setup(40000)
    run = RunSpec(**{**run.__dict__, "precision": prec,
                     "incl_deg": (70.0, 20.0)[:n_incl]})
    eng = Engine(m)
    eng.set_media(kappa, alb, g)
    w = torch.tensor(W[:, :n_incl], dtype=torch.float64, device="cuda").contiguous()
    out = {}
    for q in (0, 1):
        eng.set_schedule(peel_queue=q, vertex_slots=vertex_slots, waves_per_simd=wps)
        gr = eng.adjoint(SRC, run, w)
        torch.cuda.synchronize()
        out[q] = {k: v.cpu().numpy() for k, v in gr.items()}
    for k in ("d_logrho", "d_albedo", "d_g"):
        a, b = out[1][k], out[0][k]
        assert np.any(b != 0.0), k
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-11 * np.abs(b).max(), err_msg=k)
    assert out[1]["jlin"][0] == pytest.approx(out[0]["jlin"][0], rel=1e-12)


@pytest.mark.gpu
def test_gpu_f32_one_pass_replay_many_records_with_fallback():
    """More queued peel-off rays than threads of the gradient kernels (each
    thread takes several), with many packets left to the two-pass kernel (2
    vertex records per packet): the records of those packets are interleaved
    with valid ones in image-plane order and must be skipped without ending a
    thread's share early."""
    import torch

    m, kappa, alb, g, run, W, sel = _setup(300000)
    run = RunSpec(**{**run.__dict__, "precision": _abi.KP_F32})
    eng = Engine(m)
    eng.set_media(kappa, alb, g)
    w = torch.tensor(W, dtype=torch.float64, device="cuda")
    out = {}
    for q in (0, 1):
        eng.set_schedule(peel_queue=q, vertex_slots=2)
        gr = eng.adjoint(SRC, run, w)
        torch.cuda.synchronize()
        out[q] = {k: v.cpu().numpy() for k, v in gr.items()}
    for k in ("d_logrho", "d_albedo", "d_g"):
        a, b = out[1][k], out[0][k]
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-11 * np.abs(b).max(), err_msg=k)
    assert out[1]["jlin"][0] == pytest.approx(out[0]["jlin"][0], rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [_abi.KP_F32, _abi.KP_F64], ids=["f32", "f64"])
@pytest.mark.parametrize("vertex_slots", [0, 2])
def test_gpu_one_pass_replay_in_chunks(prec, vertex_slots):
    """The one-pass replay in chunks of packets (the claim counter starting at
    each chunk's first index, the queues reused per chunk, left-over packets
    of every chunk replayed two-pass at the end) gives the same gradients as
    the two-pass replay."""
    import torch

    m, kappa, alb, g, run, W, sel = _setup(40000)
    run = RunSpec(**{**run.__dict__, "precision": prec})
    eng = Engine(m)
    eng.set_media(kappa, alb, g)
    w = torch.tensor(W, dtype=torch.float64, device="cuda")
    out = {}
    for q in (0, 1):
        eng.set_schedule(peel_queue=q, vertex_slots=vertex_slots, adjoint_chunk=11000)
        gr = eng.adjoint(SRC, run, w)
        torch.cuda.synchronize()
        out[q] = {k: v.cpu().numpy() for k, v in gr.items()}
    for k in ("d_logrho", "d_albedo", "d_g"):
        a, b = out[1][k], out[0][k]
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-11 * np.abs(b).max(), err_msg=k)
    assert out[1]["jlin"][0] == pytest.approx(out[0]["jlin"][0], rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("model", [_abi.KP_HG_RAYLEIGH, _abi.KP_THOMSON], ids=["hg", "thomson"])
def test_gpu_f32_one_pass_replay_several_bands(model):
    """Three bands with their own extinction scale, albedo and g (per-cell
    or per-band) and uneven packet counts, and an empty band between them:
    the deferred kernels find each record's band from its packet slot.  Same
    gradients as the two-pass replay, for HG-Rayleigh and for Thomson
    scattering (no g terms)."""
    import torch

    m, kappa, alb, g, run, W, sel = _setup(30000)
    bands = [BandSpec(n_packets=30000, cell_albedo=True, cell_g=True),
             BandSpec(n_packets=7000, kappa_scale=0.3, albedo=0.7, g=0.4),
             BandSpec(n_packets=0, kappa_scale=0.5),
             BandSpec(n_packets=15000, kappa_scale=2.0, cell_albedo=True, g=0.1)]
    run = RunSpec(**{**run.__dict__, "bands": bands, "precision": _abi.KP_F32,
                     "scatter_model": model})
    eng = Engine(m)
    eng.set_media(kappa, alb, g)
    rng = np.random.default_rng(7)
    w = torch.tensor(rng.normal(size=(4, 2, 16, 16, 3)), dtype=torch.float64, device="cuda")
    out = {}
    for q in (0, 1):
        eng.set_schedule(peel_queue=q)
        gr = eng.adjoint(SRC, run, w)
        torch.cuda.synchronize()
        out[q] = {k: v.cpu().numpy() for k, v in gr.items()}
    for k in ("d_logrho", "d_albedo", "d_g"):
        a, b = out[1][k], out[0][k]
        if model == _abi.KP_THOMSON and k == "d_g":
            assert not np.any(a) and not np.any(b)
            continue
        assert np.any(b != 0.0), k
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-11 * np.abs(b).max(), err_msg=k)
    assert out[1]["jlin"][0] == pytest.approx(out[0]["jlin"][0], rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [_abi.KP_F32, _abi.KP_F64], ids=["f32", "f64"])
def test_gpu_adjoint_without_packets_leaves_gradients(prec):
    """A run whose bands hold no packets returns at once: the gradients and
    J_lin handed in are left as they were."""
    import torch

    m, kappa, alb, g, run, W, sel = _setup(10)
    run = RunSpec(**{**run.__dict__, "precision": prec,
                     "bands": [BandSpec(n_packets=0, cell_albedo=True, cell_g=True)]})
    eng = Engine(m)
    eng.set_media(kappa, alb, g)
    w = torch.tensor(W, dtype=torch.float64, device="cuda")
    gr = {k: torch.full((m.n_cells,), 2.0, dtype=torch.float64, device="cuda")
          for k in ("d_logrho", "d_albedo", "d_g")}
    gr["jlin"] = torch.full((1,), 3.0, dtype=torch.float64, device="cuda")
    eng.adjoint(SRC, run, w, grads=gr)
    torch.cuda.synchronize()
    for k in ("d_logrho", "d_albedo", "d_g"):
        assert torch.all(gr[k] == 2.0), k
    assert gr["jlin"].item() == 3.0
