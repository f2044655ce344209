"""Synthetic workloads (meshes + media) for the photon transport benchmarks.

Config 1 (the reference's CPU-runnable case): uniform 64^3 unit cube as
4x4x4 blocks of 16^3, a homogeneous Thomson sphere of radius 0.4 with radial
optical depth 1 (kappa_rho = 2.5 inside, cell-centre test), albedo 1, an
isotropic unpolarized point source at the centre, seed 42, a 128x128 I/Q/U
peel-off image at inclination 90 degrees.

Config 2: 3-level block AMR, rho ~ r^-2 times a lognormal field (sigma 0.5,
seeded Gaussian random field with a k^-11/3 power spectrum), refinement of
every block whose max |grad rho| dx / rho exceeds 0.5, HG g=0.6, albedo 0.5,
Rayleigh-form linear polarization with p_max 0.5.

Config 3 (batched): the config-2 mesh and medium seen at 64 wavelengths
log-spaced over 0.3-3 um: extinction scaled by (lambda/0.55 um)^-1.7, HG g
falling linearly from 0.7 to 0.3 across the bands, albedo 0.5, uneven packet
counts drawn log-uniform in [1e5, 1e7] from the seed; all bands in one launch.

Config 5 (large): a flared disk rho ~ (r/r0)^-1 exp(-z^2 / 2h^2), h = 0.1 r^1.25
(cylindrical r from the domain axis, unit-cube units, r0 = 0.1), an inner
hole r < 0.05, times a seeded lognormal field (sigma 0.3); 5-level block AMR
on an 8^3 root grid of 16^3 blocks (128^3 root cells), refined where
max |grad rho| dx / (rho + 2e-3) > 0.2 (~28,000 blocks, 1.2e8 cells); HG
g = 0.5, albedo 0.6, p_max 0.5; the central source; 256^2 I/Q/U images at
three inclinations; kappa_rho normalised to a midplane optical depth of 20
from the inner rim to the edge.

Spherical (the reference's spherical photon, photon_sph.h, on meshgen::sph_t):
a shell r in [0.1, 1] with logarithmic r spacing, the full theta range and a
periodic phi axis, 8x8x16 root blocks of 16^3 cells, two levels of static
refinement around an off-centre point source at (0.25, 0.15, 0.1); pure
streaming with the path-length tally.

Config 4 (adjoint): the config-2 field on a 2-level mesh of 3^3 root blocks
(96^3 effective at level 1), where the 8 root blocks whose max
|grad rho| dx / rho exceeds 1.4 are refined; the "observed" Q/U image is
rendered from the same field times an extra seeded lognormal factor
(sigma 0.2); loss = sum of squared Q and U residuals.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _abi
from .engine import BandSpec, RunSpec
from .mesh import Mesh


@dataclass
class Workload:
    name: str
    mesh: Mesh
    kappa: np.ndarray
    albedo: np.ndarray
    g: np.ndarray
    source: tuple
    run: RunSpec


def config1(n_packets: int = 1_000_000, precision: int = _abi.KP_F64, seed: int = 42,
            n_cell: int = 16, root: int = 4, image_n: int = 128) -> Workload:
    mesh = Mesh(root_blocks=(root, root, root), n_cell=n_cell).finalize()
    xc = mesh.cell_centers()
    r = np.linalg.norm(xc - 0.5, axis=1)
    R, tau = 0.4, 1.0
    kappa = np.where(r < R, tau / R, 0.0).astype(np.float32)
    albedo = np.ones(mesh.n_cells, dtype=np.float32)
    g = np.zeros(mesh.n_cells, dtype=np.float32)
    run = RunSpec(bands=[BandSpec(n_packets=n_packets, albedo=1.0)], seed=seed,
                  scatter_model=_abi.KP_THOMSON, p_max=1.0, precision=precision,
                  incl_deg=(90.0,), image_n=image_n, image_half=0.5)
    return Workload("config1_uniform64_thomson_sphere", mesh, kappa, albedo, g,
                    (0.5, 0.5, 0.5), run)


def gaussian_random_field(n: int, seed: int, slope: float = -11.0 / 3.0) -> np.ndarray:
    """Unit-variance periodic Gaussian random field on n^3 with P(k) ~ k^slope."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n) * n
    kz = np.fft.rfftfreq(n) * n
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = np.zeros_like(kk)
    nz = kk > 0
    amp[nz] = kk[nz] ** (slope / 2.0)
    noise = rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))
    f = np.fft.irfftn(noise * amp, s=(n, n, n), axes=(0, 1, 2))
    f -= f.mean()
    f /= f.std()
    return f


def _sample_field(field: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Periodic trilinear interpolation of a unit-cube field at points x [m,3]
    (field values sit at cell centres (i+0.5)/n)."""
    n = field.shape[0]
    u = x * n - 0.5
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    out = np.zeros(len(x))
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1 - f[:, 0]
        ix = (i0[:, 0] + dx) % n
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1 - f[:, 1]
            iy = (i0[:, 1] + dy) % n
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1 - f[:, 2]
                iz = (i0[:, 2] + dz) % n
                out += wx * wy * wz * field[ix, iy, iz]
    return out


def amr_density(root: int = 4, n_cell: int = 16, levels: int = 3, sigma: float = 0.5,
                seed: int = 7, threshold: float = 0.5, field_n: int = 112,
                r_soft: float = 0.02, extra_sigma: float = 0.0, extra_seed: int = 11):
    """Criterion-refined block mesh and the density rho ~ r^-2 * lognormal on it.

    A block at level l is refined when max over its cells of
    |grad rho| * dx / rho exceeds ``threshold`` (central differences of the
    density sampled at that block's cell centres), 2:1 balance kept by the
    tree.  The random field lives on a periodic field_n^3 grid (its k cutoff)
    and is interpolated trilinearly.  Returns (mesh, rho at cell centres,
    rho as a function of position)."""
    grf = gaussian_random_field(field_n, seed)
    grf2 = gaussian_random_field(field_n, extra_seed) if extra_sigma > 0 else None

    def rho_at(x):
        r2 = np.sum((x - 0.5) ** 2, axis=1) + r_soft ** 2
        ln = sigma * _sample_field(grf, x) - 0.5 * sigma ** 2
        if grf2 is not None:
            ln = ln + extra_sigma * _sample_field(grf2, x) - 0.5 * extra_sigma ** 2
        return np.exp(ln) / r2

    refine = {}
    # decide refinement level by level on a provisional uniform set of blocks
    cur = [(0, i, j, k) for i in range(root) for j in range(root) for k in range(root)]
    for lvl in range(levels - 1):
        nxt = []
        for (l, i, j, k) in cur:
            h = 1.0 / (root << l) / n_cell
            o = np.array([i, j, k], dtype=np.float64) * n_cell * h
            c = (np.arange(n_cell) + 0.5) * h
            gx, gy, gz = np.meshgrid(c, c, c, indexing="ij")
            pts = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1) + o
            rho = rho_at(pts).reshape(n_cell, n_cell, n_cell)
            gr = np.sqrt(sum(np.gradient(rho, axis=a) ** 2 for a in range(3)))
            if np.max(gr / rho) > threshold:
                refine.setdefault(l + 1, []).append((i, j, k))
                for di in range(2):
                    for dj in range(2):
                        for dk in range(2):
                            nxt.append((l + 1, 2 * i + di, 2 * j + dj, 2 * k + dk))
        cur = nxt
    mesh = Mesh(root_blocks=(root, root, root), n_cell=n_cell)
    for lvl in sorted(refine):
        # refining a block means inserting its children at lvl
        kids = []
        for (i, j, k) in refine[lvl]:
            for di in range(2):
                for dj in range(2):
                    for dk in range(2):
                        kids.append((2 * i + di, 2 * j + dj, 2 * k + dk))
        mesh.refine(lvl, np.array(kids, dtype=np.int32))
    mesh.finalize()
    rho = rho_at(mesh.cell_centers())
    return mesh, rho, rho_at


def centre_tau_norm(rho_at, n: int = 4096) -> float:
    """Mean over the six axis directions of the column of rho from the domain
    centre to the edge (midpoint rule on n points)."""
    t = (np.arange(n) + 0.5) / n * 0.5
    cols = []
    for a in range(3):
        for s in (1.0, -1.0):
            x = np.full((n, 3), 0.5)
            x[:, a] += s * t
            cols.append(np.sum(rho_at(x)) * (0.5 / n))
    return float(np.mean(cols))


def config2(n_packets: int = 100_000_000, precision: int = _abi.KP_F32, seed: int = 42,
            tau_centre: float = 5.0, **kw) -> Workload:
    """kappa_rho = kappa0 * rho with kappa0 set so that the mean centre-to-edge
    optical depth along the six axes is ``tau_centre``."""
    mesh, rho, rho_at = amr_density(**kw)
    kappa = (rho * (tau_centre / centre_tau_norm(rho_at))).astype(np.float32)
    albedo = np.full(mesh.n_cells, 0.5, dtype=np.float32)
    g = np.full(mesh.n_cells, 0.6, dtype=np.float32)
    run = RunSpec(bands=[BandSpec(n_packets=n_packets, albedo=0.5, g=0.6)], seed=seed,
                  scatter_model=_abi.KP_HG_RAYLEIGH, p_max=0.5, precision=precision,
                  incl_deg=(90.0,), image_n=128, image_half=0.5)
    return Workload("config2_amr3_r2_lognormal_hg", mesh, kappa, albedo, g, (0.5, 0.5, 0.5), run)


def config3_bands(n_bands: int = 64, seed: int = 42, n_min: float = 1e5, n_max: float = 1e7,
                  lam_min: float = 0.3, lam_max: float = 3.0, lam_ref: float = 0.55,
                  slope: float = -1.7, g_first: float = 0.7, g_last: float = 0.3,
                  albedo: float = 0.5, scale_packets: float = 1.0):
    """Per-wavelength bands: (wavelengths um, [BandSpec])."""
    lam = lam_min * (lam_max / lam_min) ** (np.arange(n_bands) / max(n_bands - 1, 1))
    g = g_first + (g_last - g_first) * np.arange(n_bands) / max(n_bands - 1, 1)
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_bands)) * scale_packets
    bands = [BandSpec(n_packets=max(1, int(round(n[k]))), kappa_scale=float((lam[k] / lam_ref) ** slope),
                      albedo=albedo, g=float(g[k])) for k in range(n_bands)]
    return lam, bands


def config3(precision: int = _abi.KP_F32, seed: int = 42, tau_centre: float = 5.0,
            scale_packets: float = 1.0, n_bands: int = 64, **kw) -> Workload:
    """64-band batched forward on the config-2 mesh (kappa_rho of config 2 is
    the 0.55 um extinction)."""
    mesh, rho, rho_at = amr_density(**kw)
    kappa = (rho * (tau_centre / centre_tau_norm(rho_at))).astype(np.float32)
    albedo = np.full(mesh.n_cells, 0.5, dtype=np.float32)
    g = np.full(mesh.n_cells, 0.6, dtype=np.float32)
    _, bands = config3_bands(n_bands=n_bands, seed=seed, scale_packets=scale_packets)
    run = RunSpec(bands=bands, seed=seed, scatter_model=_abi.KP_HG_RAYLEIGH, p_max=0.5,
                  precision=precision, incl_deg=(90.0,), image_n=128, image_half=0.5)
    return Workload("config3_amr3_64band_batched", mesh, kappa, albedo, g, (0.5, 0.5, 0.5), run)


def lognormal_factor(x: np.ndarray, sigma: float, seed: int, field_n: int = 112) -> np.ndarray:
    """exp(sigma G - sigma^2/2) of a unit-variance k^-11/3 random field at x."""
    grf = gaussian_random_field(field_n, seed)
    return np.exp(sigma * _sample_field(grf, x) - 0.5 * sigma ** 2)


# Refinement threshold of the adjoint mesh.  At the 48^3 root resolution the
# lognormal field's cell-to-cell gradients alone put max |grad rho| dx / rho
# above 1.0 in every root block (config 2's 0.5 would refine all 27 and leave
# a uniform 96^3 mesh); 1.4 refines the 8 blocks with the steepest gradients,
# the source block among them, so packets cross level jumps from the start.
CONFIG4_THRESHOLD = 1.4


@dataclass
class AdjointWorkload:
    base: Workload
    kappa_obs: np.ndarray  # extinction of the perturbed ("observed") medium


def config4(n_packets: int = 10_000_000, seed: int = 42, tau_centre: float = 5.0,
            obs_sigma: float = 0.2, obs_seed: int = 11,
            threshold: float = CONFIG4_THRESHOLD) -> AdjointWorkload:
    mesh, rho, rho_at = amr_density(root=3, levels=2, threshold=threshold)
    k0 = tau_centre / centre_tau_norm(rho_at)
    kappa = (rho * k0).astype(np.float32)
    kappa_obs = (rho * k0 * lognormal_factor(mesh.cell_centers(), obs_sigma, obs_seed)
                 ).astype(np.float32)
    albedo = np.full(mesh.n_cells, 0.5, dtype=np.float32)
    g = np.full(mesh.n_cells, 0.6, dtype=np.float32)
    run = RunSpec(bands=[BandSpec(n_packets=n_packets, cell_albedo=True, cell_g=True)],
                  seed=seed, scatter_model=_abi.KP_HG_RAYLEIGH, p_max=0.5,
                  precision=_abi.KP_F64, incl_deg=(90.0,), image_n=128, image_half=0.5)
    base = Workload("config4_amr2_96_adjoint", mesh, kappa, albedo, g, (0.5, 0.5, 0.5), run)
    return AdjointWorkload(base, kappa_obs)


def disk_rho(x: np.ndarray, grf: np.ndarray, sigma: float = 0.3, r_in: float = 0.05,
             r0: float = 0.1, h0: float = 0.1, flare: float = 1.25) -> np.ndarray:
    """Flared-disk density at points x [m,3] of the unit cube (centre 0.5)."""
    c = x - 0.5
    r = np.sqrt(c[:, 0] ** 2 + c[:, 1] ** 2)
    rr = np.maximum(r, 1e-6)
    h = h0 * rr ** flare
    ln = sigma * _sample_field(grf, x) - 0.5 * sigma ** 2
    rho = (rr / r0) ** -1 * np.exp(-c[:, 2] ** 2 / (2 * h * h) + ln)
    rho[r < r_in] = 0.0
    return rho


def disk_mesh(grf: np.ndarray, root: int = 8, n_cell: int = 16, levels: int = 5,
              threshold: float = 0.2, floor: float = 2e-3, samples: int = 8) -> Mesh:
    """Criterion-refined block mesh of the disk: a block at level l is refined
    when max |grad rho| (cell width) / (rho + floor) over samples^3 points of
    the block exceeds threshold (2:1 balance kept by the tree)."""
    refine = {}
    cur = [(i, j, k) for i in range(root) for j in range(root) for k in range(root)]
    for lvl in range(levels - 1):
        bs = 1.0 / (root << lvl)
        c = (np.arange(samples) + 0.5) / samples * bs
        gx, gy, gz = np.meshgrid(c, c, c, indexing="ij")
        off = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
        blk = np.array(cur, dtype=np.float64) * bs
        nxt = []
        for s0 in range(0, len(blk), 2000):
            B = blk[s0:s0 + 2000]
            pts = (B[:, None, :] + off[None]).reshape(-1, 3)
            rho = disk_rho(pts, grf).reshape(len(B), samples, samples, samples)
            gr = np.sqrt(sum(np.gradient(rho, axis=a) ** 2 for a in (1, 2, 3)))
            crit = (gr * (samples / n_cell) / (rho + floor)).reshape(len(B), -1).max(axis=1)
            for q in np.nonzero(crit > threshold)[0]:
                i, j, k = cur[s0 + q]
                refine.setdefault(lvl + 1, []).append((i, j, k))
                nxt += [(2 * i + a, 2 * j + b, 2 * k + d)
                        for a in range(2) for b in range(2) for d in range(2)]
        cur = nxt
    mesh = Mesh(root_blocks=(root, root, root), n_cell=n_cell)
    for lvl in sorted(refine):
        kids = [(2 * i + a, 2 * j + b, 2 * k + d) for (i, j, k) in refine[lvl]
                for a in range(2) for b in range(2) for d in range(2)]
        mesh.refine(lvl, np.array(kids, dtype=np.int32))
    return mesh.finalize()


def cell_field(mesh: Mesh, f, chunk_blocks: int = 2048) -> np.ndarray:
    """f evaluated at every cell centre, block chunk by block chunk (FP32)."""
    n = mesh.n_cell
    c = np.arange(n) + 0.5
    gi, gj, gk = np.meshgrid(c, c, c, indexing="ij")
    # cell order i0 + n*(i1 + n*i2): i0 fastest
    off = np.stack([gi.transpose(2, 1, 0).ravel(), gj.transpose(2, 1, 0).ravel(),
                    gk.transpose(2, 1, 0).ravel()], 1)
    out = np.empty(mesh.n_cells, dtype=np.float32)
    n3 = n ** 3
    for b0 in range(0, mesh.n_blocks, chunk_blocks):
        b1 = min(b0 + chunk_blocks, mesh.n_blocks)
        pts = (mesh.xf0[b0:b1, None, :] + off[None] * mesh.dx0[b0:b1, None, :]).reshape(-1, 3)
        out[b0 * n3:b1 * n3] = f(pts)
    return out


def _disk_rho_torch(x, grf_t, sigma=0.3, r_in=0.05, r0=0.1, h0=0.1, flare=1.25):
    """disk_rho on a torch device (FP64): periodic trilinear field lookup."""
    import torch

    n = grf_t.shape[0]
    u = x * n - 0.5
    i0 = torch.floor(u)
    f = u - i0
    i0 = i0.long()
    acc = torch.zeros(x.shape[0], dtype=torch.float64, device=x.device)
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1 - f[:, 0]
        ix = (i0[:, 0] + dx) % n
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1 - f[:, 1]
            iy = (i0[:, 1] + dy) % n
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1 - f[:, 2]
                iz = (i0[:, 2] + dz) % n
                acc += wx * wy * wz * grf_t[ix, iy, iz]
    c = x - 0.5
    r = torch.sqrt(c[:, 0] ** 2 + c[:, 1] ** 2)
    rr = torch.clamp(r, min=1e-6)
    h = h0 * rr ** flare
    rho = (rr / r0) ** -1 * torch.exp(-c[:, 2] ** 2 / (2 * h * h) + sigma * acc - 0.5 * sigma ** 2)
    return torch.where(r < r_in, torch.zeros_like(rho), rho)


def cell_field_disk(mesh: Mesh, grf: np.ndarray, scale: float) -> np.ndarray:
    """scale * disk_rho at every cell centre; on the GPU when torch sees one."""
    try:
        import torch

        use_gpu = torch.cuda.is_available()
    except ImportError:
        use_gpu = False
    if not use_gpu:
        return cell_field(mesh, lambda x: disk_rho(x, grf) * scale)
    dev = torch.device("cuda")
    grf_t = torch.from_numpy(grf).to(dev)
    n = mesh.n_cell
    c = torch.arange(n, dtype=torch.float64, device=dev) + 0.5
    g2, g1, g0 = torch.meshgrid(c, c, c, indexing="ij")  # i0 fastest in the flattening
    off = torch.stack([g0.reshape(-1), g1.reshape(-1), g2.reshape(-1)], 1)
    xf0 = torch.from_numpy(np.ascontiguousarray(mesh.xf0)).to(dev)
    dx0 = torch.from_numpy(np.ascontiguousarray(mesh.dx0)).to(dev)
    out = np.empty(mesh.n_cells, dtype=np.float32)
    n3 = n ** 3
    for b0 in range(0, mesh.n_blocks, 1024):
        b1 = min(b0 + 1024, mesh.n_blocks)
        pts = (xf0[b0:b1, None, :] + off[None] * dx0[b0:b1, None, :]).reshape(-1, 3)
        out[b0 * n3:b1 * n3] = (_disk_rho_torch(pts, grf_t) * scale).float().cpu().numpy()
    return out


def config5(n_packets: int = 1_000_000_000, precision: int = _abi.KP_F32, seed: int = 42,
            tau_mid: float = 20.0, field_seed: int = 21) -> Workload:
    grf = gaussian_random_field(112, field_seed)
    mesh = disk_mesh(grf)
    # midplane column from the inner rim to the edge along +x
    t = 0.05 + (np.arange(20000) + 0.5) / 20000 * 0.45
    line = np.stack([0.5 + t, np.full_like(t, 0.5), np.full_like(t, 0.5)], 1)
    col = float(np.sum(disk_rho(line, grf)) * 0.45 / 20000)
    k0 = tau_mid / col
    kappa = cell_field_disk(mesh, grf, k0)
    albedo = np.full(mesh.n_cells, 0.6, dtype=np.float32)
    g = np.full(mesh.n_cells, 0.5, dtype=np.float32)
    run = RunSpec(bands=[BandSpec(n_packets=n_packets, albedo=0.6, g=0.5)], seed=seed,
                  scatter_model=_abi.KP_HG_RAYLEIGH, p_max=0.5, precision=precision,
                  incl_deg=(30.0, 60.0, 85.0), image_n=256, image_half=0.5)
    return Workload("config5_amr5_disk_1e9", mesh, kappa, albedo, g, (0.5, 0.5, 0.5), run)


def sph_mesh(root_blocks=(4, 4, 8), n_cell: int = 16, r_lim=(0.1, 1.0), r_ratio: float = -1.0,
             theta=(0.0, np.pi), phi=(0.0, 2.0 * np.pi), refine=()) -> Mesh:
    """Spherical block mesh (axes r, theta, phi); ``refine`` is a sequence of
    (level, lo, hi) boxes in (r, theta, phi).  phi is periodic when it spans
    2 pi."""
    per = (0, 0, 1 if abs(phi[1] - phi[0] - 2.0 * np.pi) < 1e-12 else 0)
    mesh = Mesh(root_blocks=root_blocks, n_cell=n_cell, x_min=(r_lim[0], theta[0], phi[0]),
                x_max=(r_lim[1], theta[1], phi[1]), periodic=per, coord="spherical",
                r_ratio=r_ratio)
    for lvl, lo, hi in refine:
        mesh.refine_region(lvl, lo, hi)
    return mesh.finalize()


SPH_SOURCE = (0.25, 0.15, 0.1)


def config_sph(n_packets: int = 100_000_000, seed: int = 42, n_bands: int = 1,
               root_blocks=(8, 8, 16)) -> Workload:
    """Spherical streaming workload (see the module docstring)."""
    mesh = sph_mesh(root_blocks=root_blocks,
                    refine=((1, (0.2, 0.9, 0.2), (0.45, 1.6, 0.9)),
                            (2, (0.26, 1.1, 0.4), (0.36, 1.4, 0.7))))
    per_band = max(1, int(n_packets) // n_bands)
    run = RunSpec(bands=[BandSpec(n_packets=per_band, kappa_scale=0.0) for _ in range(n_bands)],
                  seed=seed, precision=_abi.KP_F64, step_max=100000)
    z = np.zeros(0, dtype=np.float32)
    return Workload("sph_shell_amr2_stream", mesh, z, z, z, SPH_SOURCE, run)
