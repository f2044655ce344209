"""The photon module (include/kp_module.h): INI read, mesh init, step, save.

CPU: INI keys under the reference's section|key names, the refined mesh
against the same refinement through the mesh ABI, and the saved container
against an independent byte-level encoding of the documented format
(binary_io.h:6-32, serialize.h).  GPU: a module step equals kp_forward runs
of the same packets (FP64 tallies to 1e-12, counters exact), cycles use
consecutive packet ids, and the saved per-block fields equal the tallies.
"""
import struct

import numpy as np
import pytest

from kratos_polrad_amd import _abi
from kratos_polrad_amd.engine import BandSpec, Engine, RunSpec
from kratos_polrad_amd.mesh import Mesh
from kratos_polrad_amd.module import PhotonModule, read_container, read_string_map

INI = """
# photon module test input
[mesh]
x_min = 0 0 0
x_max = 1 1 1
n_cell_global = 32 32 32
n_cell_block = 8 8 8

[refine_region_centre]
level = 1
x_min = 0.3 0.3 0.3
x_max = 0.7 0.7 0.7

[particle]
n_hv = 2
n_par = 3000
n_cycle_stp = 2
step_max = 100000

[photon]
seed = 7
model = hg_rayleigh
p_max = 0.5
kappa = 1.5
albedo = 0.6 0.8
g = 0.5 0.3
kappa_scale = 1.0 0.4
source = 0.51 0.49 0.5
incl = 90 30
image_n = 16
image_half = 0.6
precision = f64
"""


def test_read_keys_and_defaults():
    m = PhotonModule(INI)
    i = m.info()
    assert (i["n_bands"], i["n_par"], i["n_cycle"], i["n_incl"], i["image_n"]) == (2, 3000, 2, 2, 16)
    assert i["initialised"] == 0


@pytest.mark.parametrize("bad,msg", [
    ("[mesh]\nx_min = 0 0 0\n", "x_max"),
    ("[mesh]\nx_min = 0 0 0\nx_max = 1 1 1\nn_cell_global = 30 32 32\nn_cell_block = 8 8 8\n",
     "indivisible"),
    (INI.replace("model = hg_rayleigh", "model = mie"), "model"),
    (INI.replace("albedo = 0.6 0.8", "albedo = 0.6 0.8 0.9"), "albedo"),
    (INI + "order = random\n", "order"),
])
def test_read_rejects_bad_input(bad, msg):
    with pytest.raises(_abi.KPError) as e:
        PhotonModule(bad)
    assert msg in str(e.value)


def test_init_builds_the_refined_mesh():
    m = PhotonModule(INI).init(device=-1)
    ref = Mesh(root_blocks=(4, 4, 4), n_cell=8)
    ref.refine_region(1, (0.3, 0.3, 0.3), (0.7, 0.7, 0.7))
    ref.finalize()
    i = m.info()
    assert i["n_blocks"] == ref.n_blocks and i["n_cells"] == ref.n_cells


def _expected_container(tags):
    """binary_io container bytes built from the format description alone."""
    items = sorted(tags.items(), key=lambda kv: kv[0].encode())
    hdr = struct.pack("<Q", len(items))
    off = 0
    for tag, (data, u) in items:
        hdr += struct.pack("<Q", len(tag)) + tag.encode() + struct.pack("<QbQ", len(data), u, off)
        off += len(data)
    head = bytes([1 | 8]) + struct.pack("<Q", 1 + 8 + len(hdr)) + hdr
    return head + b"".join(d for _, (d, _u) in items)


def test_save_container_bytes(tmp_path):
    ini = """[mesh]
x_min = 0 0 0
x_max = 2 1 1
n_cell_global = 8 4 4
n_cell_block = 4 4 4
[particle]
n_hv = 1
[photon]
kappa = 0.25
"""
    m = PhotonModule(ini).init(device=-1)
    p = tmp_path / "out.bin"
    m.save(str(p))
    n3 = 64
    args = {"mesh|x_min": "0 0 0", "mesh|x_max": "2 1 1", "mesh|n_cell_global": "8 4 4",
            "mesh|n_cell_block": "4 4 4", "particle|n_hv": "1", "photon|kappa": "0.25"}
    amap = struct.pack("<Q", len(args))
    for k in sorted(args, key=str.encode):
        amap += struct.pack("<Q", len(k)) + k.encode() + struct.pack("<Q", len(args[k])) + \
            args[k].encode()
    tags = {"input_args": (amap, 1)}
    i32 = lambda *v: struct.pack("<%di" % len(v), *v)
    for b, logic in enumerate([(0, 0, 0), (1, 0, 0)]):
        pre = f"block_{b}|"
        tags[pre + "i_logic"] = (i32(*logic), 4)
        tags[pre + "level"] = (i32(0), 4)
        for fld, nint, val in (("rad_flx_flux_", 1, 0.0), ("rad_flx_mfpi_", 1, 0.25)):
            tags[pre + fld + "n_int"] = (i32(nint), 4)
            tags[pre + fld + "n_step"] = (i32(1), 4)
            tags[pre + fld + "n_cell"] = (i32(4, 4, 4), 4)
            tags[pre + fld + "n_gh"] = (i32(0, 0, 0), 4)
            tags[pre + fld + "field"] = (np.full(n3 * nint, val).tobytes(), 8)
    tags["photon|counters"] = (np.zeros(8, np.uint64).tobytes(), 8)
    tags["photon|packets_per_band"] = (struct.pack("<q", 0), 8)
    assert p.read_bytes() == _expected_container(tags)
    c = read_container(str(p))
    assert read_string_map(c["input_args"][0]) == args


@pytest.mark.gpu
@pytest.mark.parametrize("order", ["", "order = direction\n"], ids=["default", "direction"])
def test_step_equals_forward_runs(tmp_path, order):
    """photon|order = direction marches each launch's packets in
    emission-direction order: same FP64 histories, same tallies."""
    import torch

    m = PhotonModule(INI + order).init(device=0)
    m.step()
    got = m.tallies()
    assert m.info()["packets_done"] == 6000
    # the same two cycles through the engine API
    mesh = Mesh(root_blocks=(4, 4, 4), n_cell=8)
    mesh.refine_region(1, (0.3, 0.3, 0.3), (0.7, 0.7, 0.7))
    mesh.finalize()
    n = mesh.n_cells
    eng = Engine(mesh)
    eng.set_media(np.full(n, 1.5, np.float32), np.full(n, 0.6, np.float32),
                  np.full(n, 0.5, np.float32))
    bands = [BandSpec(n_packets=3000, kappa_scale=1.0, albedo=0.6, g=0.5),
             BandSpec(n_packets=3000, kappa_scale=0.4, albedo=0.8, g=0.3)]
    out = None
    for c in range(2):
        run = RunSpec(bands=bands, seed=7, scatter_model=_abi.KP_HG_RAYLEIGH, p_max=0.5,
                      precision=_abi.KP_F64, packet_offset=3000 * c, step_max=100000,
                      incl_deg=(90.0, 30.0), image_n=16, image_half=0.6)
        out = out or eng.alloc_outputs(run)
        eng.forward((0.51, 0.49, 0.5), run, out)
    torch.cuda.synchronize()
    J = out["J"].cpu().numpy().reshape(-1, 2)
    np.testing.assert_allclose(got["J"], J, rtol=1e-12, atol=1e-12 * np.abs(J).max())
    np.testing.assert_allclose(got["image"], out["image"].cpu().numpy(), rtol=1e-11,
                               atol=1e-12 * np.abs(got["image"]).max())
    assert np.array_equal(got["counters"], out["counters"].cpu().numpy().astype(np.uint64))
    # saved fields are the tallies, block by block
    p = tmp_path / "step.bin"
    m.save(str(p))
    c = read_container(str(p))
    n3 = 512
    for b in (0, mesh.n_blocks - 1):
        f = np.frombuffer(c[f"block_{b}|rad_flx_flux_field"][0], dtype=np.float64)
        np.testing.assert_array_equal(f, got["J"][b * n3:(b + 1) * n3].ravel())
    img = np.frombuffer(c["photon|image"][0], dtype=np.float64)
    np.testing.assert_array_equal(img, got["image"].ravel())


SPH_INI = """
[mesh]
coord = spherical
r_ratio = -1
x_min = 0.1 0 0
x_max = 1.0 3.141592653589793 6.283185307179586
n_cell_global = 32 16 32
n_cell_block = 8 8 8

[boundary]
kinds = outflow outflow outflow outflow periodic periodic

[refine_region_src]
level = 1
x_min = 0.2 1.0 0.2
x_max = 0.4 1.6 0.9

[particle]
n_hv = 1
n_par = 5000
n_cycle_stp = 2
step_max = 100000

[photon]
seed = 3
source = 0.25 0.15 0.1
"""


def test_spherical_ini_builds_the_sph_mesh():
    m = PhotonModule(SPH_INI).init(device=-1)
    i = m.info()
    ref = Mesh(root_blocks=(4, 2, 4), n_cell=8, x_min=(0.1, 0.0, 0.0),
               x_max=(1.0, np.pi, 2 * np.pi), periodic=(0, 0, 1), coord="spherical",
               r_ratio=-1.0)
    ref.refine_region(1, (0.2, 1.0, 0.2), (0.4, 1.6, 0.9))
    ref.finalize()
    assert (i["n_blocks"], i["n_cells"]) == (ref.n_blocks, ref.n_cells)
    with pytest.raises(_abi.KPError):
        PhotonModule(SPH_INI.replace("seed = 3", "seed = 3\nincl = 90"))
    with pytest.raises(_abi.KPError):
        PhotonModule(SPH_INI.replace("coord = spherical", "coord = polar"))


@pytest.mark.gpu
def test_spherical_step_equals_forward_runs():
    import torch

    from kratos_polrad_amd import configs

    m = PhotonModule(SPH_INI).init(device=0)
    m.step()
    got = m.tallies()
    mesh = configs.sph_mesh(root_blocks=(4, 2, 4), n_cell=8, r_lim=(0.1, 1.0), r_ratio=-1.0,
                            refine=((1, (0.2, 1.0, 0.2), (0.4, 1.6, 0.9)),))
    eng = Engine(mesh)
    out = None
    for c in range(2):
        run = RunSpec(bands=[BandSpec(n_packets=5000, kappa_scale=0.0)], seed=3,
                      packet_offset=5000 * c, step_max=100000)
        out = out or eng.alloc_outputs(run)
        eng.forward((0.25, 0.15, 0.1), run, out)
    torch.cuda.synchronize()
    J = out["J"].cpu().numpy().reshape(-1, 1)
    np.testing.assert_allclose(got["J"], J, rtol=1e-12, atol=1e-12 * np.abs(J).max())
    assert np.array_equal(got["counters"], out["counters"].cpu().numpy().astype(np.uint64))
