"""The reference's golden cases (test.sh command lines) and a parser for its stdout logs."""
import os
import re

import numpy as np

from tests.conftest import REF_GOLDEN

G = REF_GOLDEN

# test.sh:6-23 command lines of the reference's golden regression
GOLDEN_CASES = {
    "2DM1DG": dict(geometry="100", Lx=1e-8, DM=2, ntheta=8, nphi=1),
    "3DM1DG": dict(geometry="100", Lx=1e-8, DM=3, ntheta=8, nphi=1),
    "2DM2DG": dict(geometry=os.path.join(G, "2DM2DG", "mesh.mphtxt"), Lx=1e-8, Ly=1e-8, DM=2, ntheta=8, nphi=8),
    "3DM2DG": dict(geometry=os.path.join(G, "3DM2DG", "mesh.mphtxt"), Lx=1e-8, Ly=1e-8, DM=3, ntheta=4, nphi=4),
    "3DM3DG": dict(geometry=os.path.join(G, "3DM3DG", "mesh.mphtxt"), Lx=1e-8, Ly=1e-8, Lz=1e-8, DM=3,
                   ntheta=4, nphi=4),
}


def case_files(name):
    return dict(band_file=os.path.join(G, name, "band.dat"), bc_file=os.path.join(G, name, "inputbc.dat"))


def case_kwargs(name):
    kw = dict(GOLDEN_CASES[name])
    kw.update(case_files(name))
    return kw


def parse_sol(path):
    """Margins per outer iteration and heat flux per BC from a reference stdout log."""
    text = open(path).read()
    margins = [float(x) for x in re.findall(r"Margin per band per cell: (\S+)", text)]
    flux = [float(x) for x in re.findall(r"heat flux of Boundary Condition #\d+: (\S+)", text)]
    return np.array(margins), np.array(flux)


def golden_sol(name):
    return parse_sol(os.path.join(G, name, name + ".sol"))


def check_against_golden(margins, flux, gm, gf, margin_rel=3e-5, tail_rel=2e-3, flux_rel=1e-5,
                         small_flux_floor=2e-6):
    """Golden logs print 6 significant digits and stop at margin < 1e-4.  All margins but the last
    two must agree to margin_rel (the printing rounds to +-5e-6); the last two sit at the inner
    solver's noise floor and get tail_rel.  Heat fluxes: the main (isothermal) fluxes to flux_rel,
    the adiabatic ones (~1e-6 of the main ones, pure solver noise) to small_flux_floor * max|q|."""
    assert len(margins) == len(gm), (len(margins), len(gm))
    rel = np.abs(margins - gm) / np.abs(gm)
    assert rel[:-2].max() <= margin_rel, rel
    assert rel[-2:].max() <= tail_rel, rel[-2:]
    qmax = np.abs(gf).max()
    main = np.abs(gf) > 1e-3 * qmax
    assert (np.abs(flux - gf)[main] / np.abs(gf)[main]).max() <= flux_rel
    assert (np.abs(flux - gf)[~main]).max(initial=0) <= small_flux_floor * qmax
