"""The oracle's restatements of two optional pieces of the shot path, pinned
by what the mathematics fixes, and the product's host code against them.

* READ_WAVELET = 1 (src/AC/RHS_source_AC.c:41-43, src/read_stf_dft.c:34-70):
  the source amplitude is sum_j data[j] (cos, sin)(2 pi f (j-1) dt) dt / DH^2
  over j = 1..ns.  Pinned by a single sample (it enters at one sample time
  earlier, sample 0 never counts), the Dirichlet closed form of a constant
  trace, a float64 DFT of a random trace, and the TE time unit; the library's
  gfdfd_wavelet_amp equals the oracle bit for bit; the SU reader finds ns and
  dt at the header's bytes 114 and 116.
* SWS_TAPER_CIRCULAR_PER_SHOT (src/AC/ass_grad_AC.c:37-43):
  taper_grad_shot's profile is f0 = 1e-3 at the source, 0.1 at half the
  radius (the exponent's defining property), 1 beyond the radius, centred at
  (ix DH, iy DH) = srcpos; taper_grad_shot1 writes 1e-5 on the union of the
  (2 FILTSIZE + 1)^2 boxes, clipped to the grid.
"""
import numpy as np
import pytest

import germaine_amd as G
from germaine_amd import fdfd, formats
from oracle import oracle as O


def amp_closed(data, k, f, dt, dh):
    th = 2.0 * np.pi * f * (k - 1) * dt
    return data[k] * complex(np.cos(th), np.sin(th)) * dt / dh ** 2


def test_wavelet_single_sample_and_skipped_first():
    f, dh, ndt = 7.5, 10.0, 2000
    dt = ndt / 1e6
    for k in (1, 5, 37):
        d = np.zeros(64, np.float32)
        d[k] = 3.0
        a = O.wavelet_amp(O.AC, f, dh, d, ndt)
        assert abs(a - amp_closed(d, k, f, np.float32(dt), dh)) <= 1e-6 * abs(a)
    d = np.zeros(64, np.float32)
    d[0] = 5.0                        # data[0] is read but never summed
    assert O.wavelet_amp(O.AC, f, dh, d, ndt) == 0


def test_wavelet_constant_trace_dirichlet():
    f, dh, ndt, ns = 11.0, 12.5, 1000, 301
    dt = np.float32(ndt / 1e6)
    th = 2 * np.pi * f * float(dt)
    m = ns - 1                        # data[1..ns-1] at t = 0..(ns-2) dt; the unread data[ns] adds nothing
    re = np.sin(m * th / 2) * np.cos((m - 1) * th / 2) / np.sin(th / 2)
    im = np.sin(m * th / 2) * np.sin((m - 1) * th / 2) / np.sin(th / 2)
    a = O.wavelet_amp(O.AC, f, dh, np.ones(ns, np.float32), ndt)
    want = complex(re, im) * float(dt) / dh ** 2
    assert abs(a - want) <= 1e-5 * abs(want)


def test_wavelet_random_trace_and_te_unit():
    rng = np.random.default_rng(4)
    d = rng.standard_normal(500).astype(np.float32)
    for phys, f, ndt, unit, dh in ((O.AC, 4.0, 1000, 1e-6, 20.0), (O.TE, 2.5e8, 50, 1e-12, 0.04)):
        dt = float(np.float32(ndt * unit))
        t = np.arange(len(d) - 1) * dt
        want = np.sum(d[1:].astype(np.float64) * np.exp(2j * np.pi * f * t)) * dt / dh ** 2
        a = O.wavelet_amp(phys, f, dh, d, ndt)
        assert abs(a - want) <= 1e-5 * abs(want), (phys, a, want)


def test_library_wavelet_amp_equals_oracle():
    rng = np.random.default_rng(8)
    for phys in (G.AC, G.TE):
        for _ in range(5):
            d = rng.standard_normal(int(rng.integers(2, 700))).astype(np.float32)
            f = float(rng.uniform(1, 20)) * (1e7 if phys == G.TE else 1.0)
            ndt = int(rng.integers(100, 4000))
            assert fdfd.source_amp(phys, f, 0.0, 7.5, (d, ndt)) == O.wavelet_amp(phys, f, 7.5, d, ndt)


def test_read_su_wavelet(tmp_path):
    d = np.linspace(-1, 1, 123).astype(np.float32)
    head = np.zeros(240, np.uint8)
    head[114:116] = np.frombuffer(np.uint16(123).tobytes(), np.uint8)
    head[116:118] = np.frombuffer(np.uint16(4000).tobytes(), np.uint8)
    p = tmp_path / "wavelet.su"
    with open(p, "wb") as f:
        f.write(head.tobytes())
        f.write(d.tobytes())
    got, ndt = formats.read_su_wavelet(str(p))
    assert ndt == 4000 and np.array_equal(got, d)
    with open(p, "wb") as f:
        f.write(head.tobytes())
        f.write(d[:50].tobytes())
    with pytest.raises(ValueError):
        formats.read_su_wavelet(str(p))


def test_taper_grad_shot_profile():
    nx, ny, dh, R = 41, 31, 10.0, 40.0
    g = np.ones(nx * ny, np.float32)
    O.taper_grad_shot(g, nx, ny, dh, R, 20 * dh, 15 * dh, [], [])
    G2 = g.reshape(ny, nx)
    # centre at (ix, iy) = (20, 15), 1-based: the array's [14, 19]
    assert G2[14, 19] == np.float32(1e-3)
    assert abs(G2[14, 21] - 0.1) <= 1e-6          # r = R / 2
    assert abs(G2[16, 19] - 0.1) <= 1e-6
    j, i = np.mgrid[1:ny + 1, 1:nx + 1]
    r = np.hypot(i * dh - 20 * dh, j * dh - 15 * dh)
    assert np.all(G2[r >= R] == 1.0) and np.all(G2[r < R] < 1.0)     # at r = R the profile reaches 1
    # monotone in r inside the radius
    inside = np.argsort(r[r <= R].ravel())
    assert np.all(np.diff(G2[r <= R].ravel()[inside]) >= -1e-7)
    # receivers multiply on top, each at recpos * DH
    g2 = np.ones(nx * ny, np.float32)
    O.taper_grad_shot(g2, nx, ny, dh, R, 20 * dh, 15 * dh, [5, 35], [3, 3])
    G3 = g2.reshape(ny, nx)
    assert G3[2, 4] == np.float32(1e-3) and G3[2, 34] == np.float32(1e-3)
    assert G3[14, 19] == np.float32(1e-3)


def test_taper_grad_shot1_boxes():
    nx, ny, dh, F = 30, 20, 5.0, 2
    g = np.full(nx * ny, 7.0, np.float32)
    # a source near the left edge (box clipped) and one receiver
    O.taper_grad_shot1(g, nx, ny, dh, F, [1 * dh, 15 * dh], [10 * dh, 10 * dh], [25], [4])
    G2 = g.reshape(ny, nx)
    want = np.zeros((ny, nx), bool)
    for (i1, j1) in ((1, 10), (15, 10), (25, 4)):
        for i in range(i1 - F, i1 + F + 1):
            for j in range(j1 - F, j1 + F + 1):
                if 1 <= i <= nx and 1 <= j <= ny:
                    want[j - 1, i - 1] = True
    assert np.array_equal(G2 == np.float32(1e-5), want)
    assert np.all(G2[~want] == 7.0)
