"""Small seeded models shared by the tests (grid includes the PML)."""
import numpy as np


def smooth_model(nx, ny, seed=0, v0=2000.0, dv=600.0, rho0=1000.0, drho=300.0):
    rng = np.random.default_rng(seed)
    j, i = np.mgrid[0:ny, 0:nx].astype(np.float64)
    vp = v0 + dv * (0.5 + 0.5 * np.sin(i / (3 + rng.random() * 4) + 0.3) * np.cos(j / (2 + rng.random() * 3)))
    vp += 50.0 * rng.random((ny, nx))
    rho = rho0 + drho * (0.5 + 0.5 * np.cos(i / 5.0 - j / 7.0)) + 20.0 * rng.random((ny, nx))
    return vp.astype(np.float32), rho.astype(np.float32)


def te_model(nx, ny, seed=0):
    rng = np.random.default_rng(seed)
    eps0 = 8.854187817e-12
    epsr = 1.0 + 24.0 * rng.random((ny, nx))
    sigma = 0.05 * rng.random((ny, nx))
    return (epsr * eps0).astype(np.float32), sigma.astype(np.float32)
