"""On-disk formats (src/AC/write_seis_AC.c:50-63, src/AC/read_seis_AC.c:25-42,
src/readmod.c:29-91, src/writemod_true.c:27-33).  Files are built byte by
byte with struct in the nesting order of the reference's loops, and readmod's
PML extension is replayed with its four copy loops in the reference's order
(left, right, top, bottom: the later loops overwrite the corners), so a
transposed axis, a swapped (re, im) or a wrong corner fails."""
import struct

import numpy as np
import pytest

from germaine_amd import formats as FM


def test_seismogram_file_layout(tmp_path):
    nf, ns, ntr = 3, 4, 5
    val = lambda k, j, i: complex(100 * k + 10 * j + i, -(k + 0.5 * j + 0.25 * i))
    p = tmp_path / FM.seis_filename("picks/obs", 2).replace("picks/", "")
    assert p.name == "obs_p_stage_2.bin"
    with open(p, "wb") as f:                      # write_seis_AC.c:53-61 loop order
        for k in range(nf):
            for j in range(ns):
                for i in range(ntr):
                    v = val(k, j, i)
                    f.write(struct.pack("<f", v.real) + struct.pack("<f", v.imag))
    a = FM.read_seis(p, nf, ns, ntr)
    assert a.dtype == np.complex64 and a.shape == (nf, ns, ntr)
    assert all(a[k, j, i] == np.complex64(val(k, j, i)) for k in range(nf) for j in range(ns) for i in range(ntr))
    q = tmp_path / "copy.bin"
    FM.write_seis(q, [a[k] for k in range(nf)])
    assert q.read_bytes() == p.read_bytes()
    with pytest.raises(ValueError):
        FM.read_seis(p, nf, ns, ntr + 1)


def _readmod_loops(m0, npml, free_surf):
    """readmod's placement and its four extension loops, 1-based as written."""
    ny0, nx0 = m0.shape
    yshift = 0 if free_surf else npml
    NX, NY = nx0 + 2 * npml, ny0 + npml + yshift
    v = np.full((NY + 1, NX + 1), np.nan, np.float32)
    for i in range(1, nx0 + 1):
        for j in range(1, ny0 + 1):
            v[j + yshift, i + npml] = m0[j - 1, i - 1]
    for i in range(1, npml + 1):
        for j in range(1, NY + 1):
            v[j, i] = v[j, npml + 1]
    for i in range(npml + nx0 + 1, NX + 1):
        for j in range(1, NY + 1):
            v[j, i] = v[j, npml + nx0]
    if not free_surf:
        for i in range(1, NX + 1):
            for j in range(1, npml + 1):
                v[j, i] = v[npml + 1, i]
    for i in range(1, NX + 1):
        for j in range(NY - npml + 1, NY + 1):
            v[j, i] = v[NY - npml, i]
    return v[1:, 1:]


@pytest.mark.parametrize("free_surf", [0, 1])
def test_model_files(tmp_path, free_surf):
    nx0, ny0, npml = 7, 5, 3
    rng = np.random.default_rng(5)
    vp = (1500 + 1000 * rng.random((ny0, nx0))).astype(np.float32)
    rho = (1000 + 500 * rng.random((ny0, nx0))).astype(np.float32)
    for name, m in (("vp", vp), ("rho", rho)):
        with open(tmp_path / ("model." + name), "wb") as f:      # readmod.c:47-54: x outer, y inner
            for i in range(nx0):
                for j in range(ny0):
                    f.write(struct.pack("<f", m[j, i]))
    gvp, grho = FM.read_model(str(tmp_path / "model"), "ac", nx0, ny0, npml, free_surf)
    assert gvp.shape[::-1] == FM.grid_shape(nx0, ny0, npml, free_surf)
    ref = _readmod_loops(vp, npml, free_surf)
    assert not np.isnan(ref).any()
    assert np.array_equal(gvp, ref) and np.array_equal(grho, _readmod_loops(rho, npml, free_surf))
    # writemod_true writes the model without PML back, byte for byte
    FM.write_model(tmp_path / "out.vp", gvp, npml, free_surf, interior=True)
    assert (tmp_path / "out.vp").read_bytes() == (tmp_path / "model.vp").read_bytes()
    # writemod (format 3) writes the whole grid, x outer
    FM.write_model(tmp_path / "grid.vp", gvp)
    g = np.fromfile(tmp_path / "grid.vp", "<f4")
    assert g[1] == gvp[1, 0] and g[gvp.shape[0]] == gvp[0, 1]
