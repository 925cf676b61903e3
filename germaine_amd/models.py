"""Seeded synthetic models and acquisition geometries of the benchmark workloads.

Grids include the PML: the par-file interior is (nx - 2 NPML) x (ny - 2 NPML)
with FREE_SURF = 0 (GERMAINE adds NPML on every side, src/AC/forward_AC.c:
44-45), and the model is extended into the PML by copying the edge values
like readmod (src/readmod.c:62-93).  All generators use numpy's PCG64 with a
fixed seed, so every machine produces the same arrays.
"""
from dataclasses import dataclass, field

import numpy as np

EPS0 = 8.854187817e-12


@dataclass
class Workload:
    name: str
    physics: int
    nx: int
    ny: int
    npml: int
    dh: float
    freqs: list
    m1: np.ndarray = field(repr=False)          # vp | eps (absolute)
    m2: np.ndarray = field(repr=False)          # rho | sigma
    src: list = field(repr=False)               # grid rows k of the point sources
    rec: list = field(repr=False)               # grid rows k of the receivers
    a0_pml: float = 90.0
    s: float = 0.0


def extend_pml(interior, npml):
    """Edge-copy extension of an interior model into the PML (src/readmod.c:62-93)."""
    return np.pad(interior, npml, mode="edge").astype(np.float32)


def acquisition_row(nx, npml, dh, row, count, margin=2):
    """Grid rows of `count` points evenly spaced along interior row `row`, using
    the reference's point-source index rule k = nxsrc + (nysrc-1)*NX with
    nxsrc = iround(x/DH) + NPML, nysrc = iround(y/DH) + FSSHIFT (FSSHIFT = NPML)
    (src/AC/RHS_source_AC.c:23-32, src/AC/RHS_source_AC_adj.c:28)."""
    nxi = nx - 2 * npml
    xs = np.linspace(margin, nxi - 1 - margin, count)
    ks = []
    for x in xs:
        xm = np.float32(np.round(x) * dh)
        ix = int(np.floor(np.float32(xm / np.float32(dh)) + 0.5)) + npml
        iy = row + npml
        ks.append(ix + (iy - 1) * nx)
    return ks


def inclusion_interior(nxi, nyi, v0=2000.0, v1=2500.0, radius=20, rho=1000.0):
    j, i = np.mgrid[0:nyi, 0:nxi]
    ci, cj = (nxi - 1) / 2.0, (nyi - 1) / 2.0
    vp = np.where((i - ci) ** 2 + (j - cj) ** 2 <= radius ** 2, v1, v0)
    return vp.astype(np.float32), np.full((nyi, nxi), rho, np.float32)


def layered_fault_interior(nxi, nyi, seed, n_layers=12, vmin=2000.0, vmax=6000.0, n_faults=3, min_thick=3):
    """Layered model with dipping fault offsets; rho from Gardner (310 vp^0.25)."""
    rng = np.random.default_rng(seed)
    while True:
        depths = np.sort(rng.integers(min_thick, nyi - min_thick, n_layers - 1))
        if np.all(np.diff(np.concatenate([[0], depths, [nyi]])) >= min_thick):
            break
    vel = np.sort(rng.uniform(vmin, vmax, n_layers))
    j, i = np.mgrid[0:nyi, 0:nxi].astype(np.float64)
    depth = j.copy()                       # depth coordinate used to look up the layer
    for _ in range(n_faults):
        x0 = rng.uniform(0.15 * nxi, 0.85 * nxi)
        dip = np.deg2rad(rng.uniform(30.0, 70.0))
        throw = int(rng.integers(3, 11))
        side = (i - x0) * np.sin(dip) - j * np.cos(dip) > 0     # hanging wall
        depth = np.where(side, depth - throw, depth)
    layer = np.searchsorted(depths, depth, side="right")
    vp = vel[np.clip(layer, 0, n_layers - 1)].astype(np.float32)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    return vp, rho


def gaussian_smooth(field2d, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(field2d.astype(np.float64), sigma, mode="nearest").astype(np.float32)


def te_interior(nxi, nyi, seed, n_ellipses=20):
    """Layered relative permittivity with random ellipses; returns absolute eps, sigma."""
    rng = np.random.default_rng(seed)
    j, i = np.mgrid[0:nyi, 0:nxi].astype(np.float64)
    nl = 6
    depths = np.sort(rng.integers(5, nyi - 5, nl - 1))
    er_layers = np.sort(rng.uniform(1.0, 12.0, nl))
    sg_layers = rng.uniform(0.0, 0.02, nl)
    layer = np.searchsorted(depths, j, side="right")
    er = er_layers[layer]
    sg = sg_layers[layer]
    for _ in range(n_ellipses):
        cx, cy = rng.uniform(0, nxi), rng.uniform(0, nyi)
        a, b = rng.uniform(5, 60), rng.uniform(5, 60)
        th = rng.uniform(0, np.pi)
        xr = (i - cx) * np.cos(th) + (j - cy) * np.sin(th)
        yr = -(i - cx) * np.sin(th) + (j - cy) * np.cos(th)
        inside = (xr / a) ** 2 + (yr / b) ** 2 <= 1.0
        er = np.where(inside, rng.uniform(1.0, 25.0), er)
        sg = np.where(inside, rng.uniform(0.0, 0.05), sg)
    return (er * EPS0).astype(np.float32), sg.astype(np.float32)


def config(name):
    """The five benchmark workloads (grid sizes include the PML)."""
    if name == "ac_201x101_inclusion":      # 201x101, h=10, PML 20, 5 Hz, 10 sources, 100 receivers on row 5
        nx, ny, npml, dh = 201, 101, 20, 10.0
        vp, rho = inclusion_interior(nx - 2 * npml, ny - 2 * npml)
        return Workload(name, 1, nx, ny, npml, dh, [5.0], extend_pml(vp, npml), extend_pml(rho, npml),
                        acquisition_row(nx, npml, dh, 5, 10), acquisition_row(nx, npml, dh, 5, 100))
    if name == "ac_801x187_4f":      # 801x187, h=25, PML 20, 3/5/7/10 Hz, 100 shots, 400 receivers
        nx, ny, npml, dh = 801, 187, 20, 25.0
        vp, rho = layered_fault_interior(nx - 2 * npml, ny - 2 * npml, seed=2)
        return Workload(name, 1, nx, ny, npml, dh, [3.0, 5.0, 7.0, 10.0], extend_pml(vp, npml),
                        extend_pml(rho, npml), acquisition_row(nx, npml, dh, 1, 100),
                        acquisition_row(nx, npml, dh, 1, 400))
    if name == "ac_1601x401_8f":      # 1601x401, h=12.5, PML 30, 2..16 Hz, 200 shots
        nx, ny, npml, dh = 1601, 401, 30, 12.5
        vp, rho = layered_fault_interior(nx - 2 * npml, ny - 2 * npml, seed=3)
        return Workload(name, 1, nx, ny, npml, dh, [2.0 * k for k in range(1, 9)], extend_pml(vp, npml),
                        extend_pml(rho, npml), acquisition_row(nx, npml, dh, 1, 200), [])
    if name == "ac_801x201_fwi":      # 801x201, h=20, PML 20, 3..9 Hz, 80 shots, 320 receivers; smoothed start model
        nx, ny, npml, dh = 801, 201, 20, 20.0
        vp, rho = layered_fault_interior(nx - 2 * npml, ny - 2 * npml, seed=4)
        return Workload(name, 1, nx, ny, npml, dh, [3.0, 5.0, 7.0, 9.0], extend_pml(vp, npml),
                        extend_pml(rho, npml), acquisition_row(nx, npml, dh, 1, 80),
                        acquisition_row(nx, npml, dh, 1, 320))
    if name == "te_2401x801_6f":      # TE 2401x801, h=0.04, PML 40, 100..600 MHz, 150 shots, 600 receivers
        nx, ny, npml, dh = 2401, 801, 40, 0.04
        eps, sig = te_interior(nx - 2 * npml, ny - 2 * npml, seed=5)
        return Workload(name, 3, nx, ny, npml, dh, [1e8 * k for k in range(1, 7)], extend_pml(eps, npml),
                        extend_pml(sig, npml), acquisition_row(nx, npml, dh, 1, 150),
                        acquisition_row(nx, npml, dh, 1, 600))
    raise ValueError(name)
