"""gfdfd_solve_noalloc, the solve that allocates no device memory, against
gfdfd_solve through the C-ABI.

The no-allocation entry point runs the single-vector sweep kernels with its
work buffers in a file-scope device array; gfdfd_solve runs the same sweeps
but, on the levels of few large fronts, multiplies by inverse pivot triangles
that it allocates once per factorisation.  Both solve exactly, so their
results agree to rounding: 1e-12 relative with FP64 factors, 1e-5 with FP32.
Two calls in a row -- a 61 x 41 grid with three right-hand sides, then a
smaller 45 x 33 grid with one (the array still holds the first call's
contents) -- for FP64 and for FP32 factors; neither call may change the
library's or torch's device allocations.  A fourth right-hand side more than
the entry point takes, and a grid whose buffers exceed the array, are
rejected with GFDFD_ERR_ARG.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import germaine_amd as G
from germaine_amd import fdfd
from tests._grids import smooth_model

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ptr(t):
    return C.c_void_p(t.data_ptr())


def rel(x, ref):
    return float((x - ref).abs().max() / ref.abs().max())


def library_bytes():
    cur = C.c_longlong(0)
    G.check(G.lib().gfdfd_device_memory(C.byref(cur), None, 0), "gfdfd_device_memory")
    return cur.value


def system(nx, ny, precision, freq=7.0, seed=3):
    m1, m2 = smooth_model(nx, ny, seed=seed)
    pl = fdfd.Plan(nx, ny, small_max=16)
    vals = pl.assemble(G.AC, 10.0, [freq], torch.as_tensor(m1, device="cuda"), torch.as_tensor(m2, device="cuda"),
                       npml=5)
    return pl, pl.factor(vals, precision)


@pytest.mark.parametrize("precision", [G.FP64, G.FP32])
def test_noalloc_solve_equals_solve(precision):
    L = G.lib()
    st = torch.cuda.current_stream()
    cases = [(61, 41, 3, 0), (45, 33, 1, 2)]      # (nx, ny, right-hand sides, sys); the second is smaller
    systems, rhs, refs = [], [], []
    rng = np.random.default_rng(11)
    for nx, ny, nrhs, sysno in cases:
        pl, F = system(nx, ny, precision)
        B = torch.as_tensor(rng.standard_normal((nrhs, pl.n)) + 1j * rng.standard_normal((nrhs, pl.n)),
                            device="cuda")
        systems.append((pl, F))
        rhs.append(B)
        refs.append(F.solve(0, B, sys=sysno))      # gfdfd_solve
    outs = [torch.empty_like(B) for B in rhs]
    torch.cuda.synchronize()
    lib0, torch0 = library_bytes(), torch.cuda.memory_allocated()
    for (nx, ny, nrhs, sysno), (pl, F), B, X in zip(cases, systems, rhs, outs):
        say("noalloc solve", precision, nx, ny, nrhs, sysno)
        G.check(L.gfdfd_solve_noalloc(F.h, 0, sysno, nrhs, ptr(B), ptr(X), C.c_void_p(st.cuda_stream)),
                "gfdfd_solve_noalloc")
    torch.cuda.synchronize()
    assert library_bytes() == lib0 and torch.cuda.memory_allocated() == torch0
    bar = 1e-12 if precision == G.FP64 else 1e-5
    for X, R in zip(outs, refs):
        assert rel(X, R) <= bar
    # in place (d_x == d_b), as gfdfd_solve allows
    (pl, F), B = systems[1], rhs[1].clone()
    G.check(L.gfdfd_solve_noalloc(F.h, 0, cases[1][3], cases[1][2], ptr(B), ptr(B), C.c_void_p(st.cuda_stream)),
            "gfdfd_solve_noalloc")
    torch.cuda.synchronize()
    assert rel(B, refs[1]) <= bar


def test_noalloc_solve_rejects_what_does_not_fit():
    L = G.lib()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    pl, F = system(45, 33, G.FP64)
    B = torch.zeros((5, pl.n), dtype=torch.complex128, device="cuda")
    assert L.gfdfd_solve_noalloc(F.h, 0, 0, 5, ptr(B), ptr(B), st) == -1         # GFDFD_ERR_ARG
    assert L.gfdfd_solve_noalloc(F.h, 0, 0, 0, ptr(B), ptr(B), st) == -1
    assert L.gfdfd_solve_noalloc(F.h, 1, 0, 1, ptr(B), ptr(B), st) == -1         # one frequency only
    # 1801 x 601 in FP64: n + work rows = 4.69M entries, 71.6 MiB of buffers against 64 MiB
    big, Fb = system(1801, 601, G.FP64, freq=2.0)
    need = (big.n * 16 + 255) // 256 * 256 + int(big.info.work_rows) * 16
    assert need > L.gfdfd_solve_noalloc_limit() == 64 << 20
    Bb = torch.zeros((1, big.n), dtype=torch.complex128, device="cuda")
    lib0 = library_bytes()
    assert L.gfdfd_solve_noalloc(Fb.h, 0, 0, 1, ptr(Bb), ptr(Bb), st) == -1
    assert library_bytes() == lib0
    say("rejected:", L.gfdfd_last_error().decode())
