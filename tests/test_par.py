"""The input contract (germaine_amd/par.py) on the CPU.

* known answers from GERMAINE's shipped acoustic two-inclusion run
  (tests/golden/par/, copied by tests/golden/make_par_fixtures.py): every
  par field, the eight workflow stages, the 264 sources and receivers, and
  their solution rows;
* the par scanner against glibc's own fscanf: a small C program with the
  reference's counting loop (a line counts when the character after a
  newline is not '#') and the same conversions scans synthetic par files
  with blank lines, octal and hex integers, a missing '=' and short lines,
  and both must agree field by field.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from germaine_amd import par as P
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "par")


def test_shipped_par_file_known_answers():
    p = P.read_par(os.path.join(GOLD, "GERMAINE_incl2.inp"))
    want = dict(INVMAT=0, PHYSICS=1, NPROCFREQ=1, NX=250, NY=125, DH=40.0, SOURCE_FILE="./source/source_inc.dat",
                READ_WAVELET=0, WAVELET_NAME="start/wavelet.su", READMOD=1, MFILE="start/inclusions_true",
                FREE_SURF=0, NPML=10, A0_PML=90.0, SNAP=0, IDX=1, IDY=1, SNAP_FORMAT=3, SNAP_FILE="./snap/incl2",
                SEISMO=1, REC_FILE="./receiver/receiver_inc", READ_REC=0, REFREC1=0.0, REFREC2=0.0, N_STREAMER=0,
                REC_INCR_X=80.0, REC_INCR_Y=0.0, PICKS_FILE="./seis/incl2", LOG_FILE="log/GERMAINE.log", LOG=1,
                ITERMAX=10, STF_INV=0, JACOBIAN="gradient/gradient", DATA_DIR="seis/incl2_true/incl2", TAPER=0,
                TAPERLENGTH=4, HESSIAN=2, EPS_HESS=float(np.float32(0.01)), GRADT1=22, GRADT2=25, GRADT3=490,
                GRADT4=500, SWS_TAPER_GRAD_VERT=0, SWS_TAPER_GRAD_HOR=0, EXP_TAPER_GRAD_HOR=0.0,
                SWS_TAPER_GRAD_SOURCES=0, SWS_TAPER_CIRCULAR_PER_SHOT=0, SRTSHAPE=1, SRTRADIUS=80.0, FILTSIZE=1,
                SWS_TAPER_FILE=0, INV_MODELFILE="model/modelTest", nfstart=2000, nf=2000, nfstart_jac=2000,
                nf_jac=1, MAT1_LOW=300.0, MAT1_UP=6000.0, MAT2_LOW=1.25, MAT2_UP=2000.0, MAT1_NORM=1.0,
                MAT2_NORM=1.0, GRAD_METHOD=2, PCG_BETA=2, NLBFGS=10, MODEL_FILTER=0, FILT_SIZE=5, LINESEARCH=2,
                EPS_SCALE=float(np.float32(0.01)), C1=float(np.float32(1e-4)), C2=float(np.float32(0.9)), STEPMAX=6,
                SCALEFAC=2.0, MISFIT_LOG_FILE="LOG_TEST.dat", MIN_ITER=0, VP0_1=1500.0, VP0_2=2501.0, DVP0=1.0,
                GRAD0_1=0.0, GRAD0_2=float(np.float32(1.6)), DGRAD0=float(np.float32(0.1)),
                GRIDSEARCH_FILE="./gridsearch/1D_gridsearch.dat", FSSHIFT=10)
    assert set(want) <= set(p)
    for k, v in want.items():
        assert p[k] == v, (k, p[k], v)


def test_shipped_workflow_known_answers():
    st = P.read_workflow(os.path.join(GOLD, "GERMAINE_workflow_incl2.inp"))
    assert len(st) == 8
    lows = [4.0, 6.3, 8.6, 10.9, 13.1, 15.4, 17.7, 20.0]
    for s, lo in zip(st, lows):
        assert s["FC_low"] == float(np.float32(lo)) and s["FC_high"] == 20.0 and s["NF"] == 1
        assert s["MISFIT"] == 1 and s["SPATFILTER"] == 1 and s["VREF"] == 4000.0 and s["S"] == 0.0
        assert s["FILT_SIZE_GRAD"] == 1.0 and s["LAMBDA_1"] == 0.0 and s["BETA_MAT1"] == 0.0
        assert P.stage_frequencies(s) == [float(np.float32(lo))]
    # NF > 1: FC_low, then (FC_high - FC_low) / (NF - 1) steps in float (src/AC/forward_AC.c:120-131)
    fr = P.stage_frequencies(dict(NF=4, FC_low=2.0, FC_high=8.0))
    assert fr == [2.0, 4.0, 6.0, 8.0]


def test_shipped_sources_and_receivers_known_answers():
    src = P.read_sources(os.path.join(GOLD, "source_inc.dat"))
    assert src.shape == (264, 8)
    assert np.array_equal(src[0], np.float32([500.0, 400.0, 0.0, 0.0, 5.0, 1.0, 0.0, 1.0]))
    assert np.array_equal(src[2, :2], np.float32([400.0, 500.0]))
    rec = P.read_receivers(os.path.join(GOLD, "receiver_inc.dat"), 40.0)
    assert rec.shape == (264, 2)
    # iround(12.5) = 13, iround(237.5) = 238 (floor(x + 0.5))
    assert tuple(rec[0]) == (13, 10) and tuple(rec[1]) == (238, 10) and tuple(rec[2]) == (10, 13)
    # solution rows use the point-source rule of src/AC/RHS_source_AC.c:23-32
    k = P.grid_rows(rec[:3, 0], rec[:3, 1], 10, 10, 270)
    assert k == [O.source_index(x, y, 40.0, 10, 10, 270) for x, y in ((500.0, 400.0), (9500.0, 400.0),
                                                                          (400.0, 500.0))]


def test_setup_from_shipped_files(tmp_path):
    # the model files the par file names (start/inclusions_true.vp/.rho), written here
    for f in os.listdir(GOLD):
        shutil.copy(os.path.join(GOLD, f), tmp_path)
    p = P.read_par(os.path.join(GOLD, "GERMAINE_incl2.inp"))
    os.makedirs(tmp_path / "source")
    os.makedirs(tmp_path / "receiver")
    os.makedirs(tmp_path / "start")
    shutil.copy(os.path.join(GOLD, "source_inc.dat"), tmp_path / "source")
    shutil.copy(os.path.join(GOLD, "receiver_inc.dat"), tmp_path / "receiver")
    vp = np.full((p["NY"], p["NX"]), 2000.0, np.float32)
    vp[60:70, 100:120] = 2500.0
    rho = np.full_like(vp, 1000.0)
    for sfx, m in (("vp", vp), ("rho", rho)):
        np.ascontiguousarray(m.T, "<f4").tofile(tmp_path / "start" / f"inclusions_true.{sfx}")
    run = P.setup(str(tmp_path / "GERMAINE_incl2.inp"), str(tmp_path / "GERMAINE_workflow_incl2.inp"))
    assert (run["nx"], run["ny"], run["npml"], run["fsshift"]) == (270, 145, 10, 10)
    assert len(run["src_rows"]) == 264 and len(run["rec_rows"]) == 264
    assert run["src_rows"][0] == 23 + 19 * 270
    assert len(run["stages"]) == 8 and run["stages"][0]["freqs"] == [4.0]
    assert run["m1"].shape == (145, 270)
    assert np.array_equal(run["m1"][10:135, 10:260], vp) and run["m1"][0, 0] == vp[0, 0]


C_SCAN = r"""
#include <stdio.h>
/* counting loop and conversions of a GERMAINE par file, printed one field per line */
int main(int argc, char **argv) {
    FILE *fp = fopen(argv[1], "r");
    char s[256], str[256];
    int c, lineno = 0, i1, i2, i3, i4, n;
    float f1, f2, f3;
    while ((c = getc(fp)) != EOF) {
        if ((c == '\n') && (getc(fp) != '#')) {
            lineno++;
            i1 = i2 = i3 = i4 = -777; f1 = f2 = f3 = -777.0f; str[0] = 0;
            switch (lineno) {
            case 1: n = fscanf(fp, "%s =%i", s, &i1); printf("%d i %d\n", lineno, i1); break;
            case 2: n = fscanf(fp, "%s =%f", s, &f1); printf("%d f %.9g\n", lineno, f1); break;
            case 3: n = fscanf(fp, "%s =%s", s, str); printf("%d s %s\n", lineno, str); break;
            case 4: n = fscanf(fp, "%s =%f ,%f", s, &f1, &f2); printf("%d ff %.9g %.9g\n", lineno, f1, f2); break;
            case 5: n = fscanf(fp, "%s =%i, %i, %i, %i", s, &i1, &i2, &i3, &i4);
                    printf("%d iiii %d %d %d %d\n", lineno, i1, i2, i3, i4); break;
            case 6: n = fscanf(fp, "%s =%f, %f, %f", s, &f1, &f2, &f3);
                    printf("%d fff %.9g %.9g %.9g\n", lineno, f1, f2, f3); break;
            case 7: n = fscanf(fp, "%s =%i", s, &i1); printf("%d i %d\n", lineno, i1); break;
            case 8: n = fscanf(fp, "%s =%i", s, &i1); printf("%d i %d\n", lineno, i1); break;
            default: break;
            }
        }
    }
    return 0;
}
"""
# the same eight conversions through germaine_amd.par's scanner
SCAN_KINDS = [("i",), ("f",), ("s",), ("f ,f",), ("i, i, i, i",), ("f, f, f",), ("i",), ("i",)]

SYNTHETIC = [
    "#\nFIELD_a = 010\nFIELD_b = 2.5e-3\nFIELD_c = ./a/b.dat\nREF = 1.5 , -2\n"
    "T = 22, 25, 490, 500\nR = 1500.0, 2501.0, 1.0\nH = 0x1F\nX = -7\n",
    "# comment\n#\n  A = 42\n\nB=  7.25\nC =\n text\nD = 3,4\nE = 1 , 2 , 3 , 4\nF = 1.0\nG 5\nH = +9;\n",
    "#x\nA = 1\n#skip\nB = 1e5\n#\nC = x\nD = .5,.25\nE = 1,2,3\nF = nan, inf, -inf\nG = 0\nH = 08\n",
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc for the glibc fscanf harness")
@pytest.mark.parametrize("case", range(len(SYNTHETIC)))
def test_par_scanner_matches_glibc_fscanf(tmp_path, case):
    src = tmp_path / "scan.c"
    src.write_text(C_SCAN)
    exe = tmp_path / "scan"
    subprocess.check_call(["gcc", "-O0", "-o", str(exe), str(src)])
    parf = tmp_path / "p.inp"
    parf.write_text(SYNTHETIC[case])
    lines = subprocess.check_output([str(exe), str(parf)]).decode().split("\n")
    got_c = {}
    for ln in lines:
        if ln.strip():
            parts = ln.split()
            got_c[int(parts[0])] = parts[2:]
    # python: the same counting loop with the scanner's conversions
    sc = P._Scan(SYNTHETIC[case])
    got_py = {}
    lineno = 0
    while True:
        c = sc.getc()
        if c is None:
            break
        if c == "\n" and sc.getc() != "#":
            lineno += 1
            if lineno <= len(SCAN_KINDS):
                conv = SCAN_KINDS[lineno - 1][0]
                vals = P._scan_field(sc, conv)
                kinds = [k.strip() for k in conv.split(",")]
                out = []
                for i, k in enumerate(kinds):
                    if i < len(vals):
                        out.append(vals[i])
                    else:
                        out.append("" if k == "s" else -777)
                got_py[lineno] = out
    assert set(got_py) == set(got_c)
    for ln, cv in got_c.items():
        pv = got_py[ln]
        for a, b in zip(cv, pv):
            if isinstance(b, str):
                assert a == b or (a == "" and b == ""), (ln, cv, pv)
            elif isinstance(b, int):
                assert int(a) == b, (ln, cv, pv)
            else:
                fa = float(a)
                assert (np.isnan(fa) and np.isnan(b)) or np.float32(fa) == np.float32(b), (ln, cv, pv)
