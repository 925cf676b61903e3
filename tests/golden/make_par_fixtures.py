"""Copies GERMAINE's shipped input files (data, not code) into tests/golden/par/
as fixtures for the input-contract tests: the acoustic two-inclusion par file
and its workflow table, and its source and receiver geometry files.

Run from the repository root in a container that has the reference:
    python tests/golden/make_par_fixtures.py /root/reference
"""
import os
import shutil
import sys

FILES = ["par/GERMAINE_incl2.inp", "par/GERMAINE_workflow_incl2.inp", "par/source/source_inc.dat",
         "par/receiver/receiver_inc.dat"]


def main(ref):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "par")
    os.makedirs(out, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(ref, f), os.path.join(out, os.path.basename(f)))
        print("copied", f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
