"""The reference's own regression (test.sh:6-23) through BTEcmd on the GPU: the five command lines,
with --stop 1e-4 (the shipped logs stop at the first margin below 1e-4), their stdout compared
line by line with tests/*/*.sol -- the banner exactly, the margins and heat fluxes numerically to
the golden tolerances (6 printed digits).  Lines the logs' build did not print (the separator
before each margin, StaticBTESolver.cpp:789, and the timing lines) are dropped first.  Also the
--temperature file in the Tempcell.dat layout (StaticBTESolver.cpp:818-831): x/Lx y/Ly [z/Lz] T per
cell, checked against the oracle's cell centres and temperatures.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from staticbtesolver_amd import capi
from tests.cases import G, check_against_golden, golden_sol

pytestmark = pytest.mark.gpu

# test.sh:6-23 (paths relative to the golden directory)
COMMANDS = {
    "3DM3DG": "-x 1e-8 -y 1e-8 -z 1e-8 -t 4 -p 4 -d 3 -I 1000 -g {d}/mesh.mphtxt -b {d}/inputbc.dat -m {d}/band.dat",
    "3DM2DG": "-x 1e-8 -y 1e-8 -z 0 -t 4 -p 4 -d 3 -I 1000 -g {d}/mesh.mphtxt -b {d}/inputbc.dat -m {d}/band.dat",
    "3DM1DG": "-x 1e-8 -t 8 -p 1 -d 3 -I 1000 -g 100 -b {d}/inputbc.dat -m {d}/band.dat",
    "2DM2DG": "-x 1e-8 -y 1e-8 -t 8 -p 8 -d 2 -I 10000 -g {d}/mesh.mphtxt -b {d}/inputbc.dat -m {d}/band.dat",
    "2DM1DG": "-x 1e-8 -t 8 -p 1 -d 2 -I 10000 -g 100 -b {d}/inputbc.dat -m {d}/band.dat",
}

NUM = r"([-+0-9.eE]+|-?nan|-?inf)"


def say(*a):
    print(*a, flush=True)


def normalise(text):
    keep = []
    for line in text.splitlines():
        if line.startswith("-----") or line.startswith("Time taken"):
            continue
        keep.append(line)
    while keep and not keep[-1].strip():
        keep.pop()
    return keep


def shape_of(lines):
    """Each line with its numbers replaced by #, and the numbers of the margin / flux lines."""
    shape, margins, fluxes = [], [], []
    for line in lines:
        m = re.fullmatch(r"Iteration #(\d+)\t Margin per band per cell: " + NUM, line)
        q = re.fullmatch(r"heat flux of Boundary Condition #(\d+): " + NUM, line)
        if m:
            shape.append(f"Iteration #{m.group(1)}\t Margin per band per cell: #")
            margins.append(float(m.group(2)))
        elif q:
            shape.append(f"heat flux of Boundary Condition #{q.group(1)}: #")
            fluxes.append(float(q.group(2)))
        else:
            shape.append(line)
    return shape, np.array(margins), np.array(fluxes)


@pytest.mark.parametrize("name", sorted(COMMANDS))
def test_btecmd_stdout_matches_golden_log(name, tmp_path):
    say("BTEcmd", name)
    d = os.path.join(G, name)
    args = COMMANDS[name].format(d=d).split() + ["--stop", "1e-4", "--device", "0"]
    r = subprocess.run([capi.BTECMD] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = normalise(r.stdout)
    want = normalise(open(os.path.join(d, name + ".sol")).read())
    gs, gm, gq = shape_of(got)
    ws, wm, wq = shape_of(want)
    assert gs == ws
    gm_ref, gq_ref = golden_sol(name)
    np.testing.assert_array_equal(wm, gm_ref)
    # printed at 6 digits: the golden tolerances of the library-level check apply unchanged
    check_against_golden(gm, gq, wm, wq)


@pytest.mark.parametrize("name", ["2DM2DG", "3DM3DG"])
def test_temperature_file_layout(name, tmp_path):
    say("Tempcell", name)
    d = os.path.join(G, name)
    out = tmp_path / "Tempcell.dat"
    args = COMMANDS[name].format(d=d).split() + ["--stop", "1e-4", "--temperature", str(out)]
    r = subprocess.run([capi.BTECMD] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [line.split() for line in out.read_text().splitlines()]
    kw = dict(geometry=os.path.join(d, "mesh.mphtxt"), band_file=os.path.join(d, "band.dat"),
              bc_file=os.path.join(d, "inputbc.dat"), DM=3 if name.startswith("3D") else 2,
              ntheta=4 if name.startswith("3D") else 8, nphi=4 if name.startswith("3D") else 8, T_ref=300.0,
              Lx=1e-8, Ly=1e-8, Lz=1e-8 if name == "3DM3DG" else 0.0)
    o = O.OracleCase(**kw)
    o.set_inner(O.INNER_EXACT)
    o.set_stop(1e-4)
    ref = o.solve(10000)
    dim = o.dim
    assert len(rows) == o.N_cell and all(len(r_) == dim + 1 for r_ in rows)
    vals = np.array([[float(x) for x in r_] for r_ in rows])
    cen = o.geometry()["centre"][:, :dim] / 1e-8
    # ostream's default 6 significant digits
    np.testing.assert_allclose(vals[:, :dim], cen, rtol=6e-6, atol=1e-6)
    np.testing.assert_allclose(vals[:, dim], ref["T"], rtol=2e-6)
