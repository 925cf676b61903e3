"""The reference's on-disk formats, so a run of this library reads and writes
the same files as GERMAINE's own harness.  Host-side numpy; nothing here
touches the device.

* Seismograms ``<PICKS_FILE>_p_stage_<n>.bin`` (src/AC/write_seis_AC.c:50-63,
  src/AC/read_seis_AC.c:25-42): float32 (re, im) pairs, little-endian, in the
  order frequency -> shot -> receiver, with no header.  A complex64 array of
  shape [nfreq, nshots, ntr] has exactly this byte layout.
* Model files ``<MFILE>.vp`` / ``.rho`` (AC, src/readmod.c:29-91) and ``.sig``
  / ``.eps`` (TE, src/TE/readmod_TE.c:30-92): float32, x outer and y inner
  (each grid column contiguous), over the NX0 x NY0 model without PML.  The
  reader extends the model into the PML by copying the nearest edge value.
* Model output (src/writemod.c:28-33 with format 3 for the whole grid,
  src/writemod_true.c:27-33 for the model without PML): the same ordering.

Arrays in this package are [ny, nx] row-major (y rows, x columns), so a file
holds the transpose of one.
"""
import numpy as np

MODEL_SUFFIXES = {"ac": ("vp", "rho"), "te": ("sig", "eps")}


def seis_filename(prefix, stage):
    """``<prefix>_p_stage_<stage>.bin`` (src/AC/write_seis_AC.c:50)."""
    return "%s_p_stage_%d.bin" % (prefix, stage)


def write_seis(path, data):
    """data: complex [nfreq, nshots, ntr] (or a list of [nshots, ntr] per frequency)."""
    a = np.stack([np.asarray(d) for d in data]) if isinstance(data, (list, tuple)) else np.asarray(data)
    if a.ndim != 3:
        raise ValueError("write_seis: data must be [nfreq, nshots, ntr]")
    np.ascontiguousarray(a, "<c8").tofile(path)


def read_seis(path, nfreq, nshots, ntr):
    """The [nfreq, nshots, ntr] complex64 samples of a seismogram file; raises
    when the file holds fewer or more than nfreq * nshots * ntr pairs."""
    a = np.fromfile(path, "<c8")
    want = nfreq * nshots * ntr
    if a.size != want:
        raise ValueError("read_seis: %s holds %d samples, expected %d" % (path, a.size, want))
    return a.reshape(nfreq, nshots, ntr).astype(np.complex64)


def grid_shape(nx0, ny0, npml, free_surf):
    """(NX, NY) of the grid with PML (src/TE/forward_TE.c:39-44): PML on the left,
    right and bottom, and on top unless the top is a free surface."""
    fsshift = 0 if free_surf else npml
    return nx0 + 2 * npml, ny0 + npml + fsshift


def extend_model(m0, npml, free_surf):
    """[ny0, nx0] model -> [NY, NX] grid with the PML filled by the nearest
    edge value, corners by the corner value (src/readmod.c:56-91)."""
    fsshift = 0 if free_surf else npml
    return np.pad(np.asarray(m0, np.float32), ((fsshift, npml), (npml, npml)), mode="edge")


def read_model_file(path, nx0, ny0):
    """One float32 model file of the NX0 x NY0 model, as [ny0, nx0]."""
    a = np.fromfile(path, "<f4")
    if a.size != nx0 * ny0:
        raise ValueError("read_model_file: %s holds %d values, expected %d" % (path, a.size, nx0 * ny0))
    return np.ascontiguousarray(a.reshape(nx0, ny0).T)


def read_model(prefix, physics, nx0, ny0, npml, free_surf):
    """readmod / readmod_TE: the two material files of ``prefix`` ("ac": .vp
    and .rho; "te": .sig and .eps), each extended into the PML.  Returns two
    float32 [NY, NX] arrays in the order of the suffixes."""
    sfx = MODEL_SUFFIXES[physics]
    return tuple(extend_model(read_model_file("%s.%s" % (prefix, s), nx0, ny0), npml, free_surf) for s in sfx)


def write_model(path, m, npml=0, free_surf=1, interior=False):
    """writemod (format 3: the whole [NY, NX] grid) or, with interior=True,
    writemod_true (the model without PML) -- x outer, y inner, float32."""
    m = np.asarray(m, np.float32)
    if interior:
        fsshift = 0 if free_surf else npml
        ny, nx = m.shape
        m = m[fsshift:ny - npml, npml:nx - npml]
    np.ascontiguousarray(m.T, "<f4").tofile(path)


def read_su_wavelet(path):
    """The time-domain source wavelet of READ_WAVELET = 1 (src/read_stf_dft.c:25-58):
    the first trace of an SU file -- a 240-byte trace header whose bytes 114-115
    hold ns and 116-117 the sample interval (native-endian unsigned 16-bit
    fields of the SU header), then ns float32 samples.  Returns (samples, ndt);
    fdfd.source_amp(..., wavelet=(samples, ndt)) turns it into the per-frequency
    source amplitude."""
    with open(path, "rb") as f:
        head = f.read(240)
        if len(head) < 240:
            raise ValueError(f"{path}: shorter than one SU trace header")
        ns = int(np.frombuffer(head, np.uint16, 1, 114)[0])
        ndt = int(np.frombuffer(head, np.uint16, 1, 116)[0])
        data = np.fromfile(f, np.float32, ns)
    if len(data) < ns:
        raise ValueError(f"{path}: the header announces {ns} samples, the file holds {len(data)}")
    return data, ndt
