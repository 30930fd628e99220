"""Slab observables of sre/analysis.py and the thermo line of sre/apps.py (CPU)."""
import re

import numpy as np

from sre import analysis, apps


def test_block_average_known_answer():
    # 35 samples, 4 blocks of 8: the last 3 are dropped; block means
    # 3.5, 11.5, 19.5, 27.5 around 15.5 -> sqrt((144+16+16+144) / 12)
    b = analysis.BlockAverage(np.arange(35.0), 4)
    assert b.n_used == 32
    assert b.avg == 15.5
    assert abs(b.error - np.sqrt(320.0 / 12.0)) < 1e-14
    # independent noise: the block error estimates sigma / sqrt(n)
    x = np.random.default_rng(0).normal(0.0, 2.0, 64000)
    e = analysis.BlockAverage(x, 16).error
    assert 0.6 * 2.0 / np.sqrt(64000) < e < 1.5 * 2.0 / np.sqrt(64000)


def test_slab_densities_regions_and_wrap():
    box = (150.0, 21.0, 21.0)
    x = np.array([
        10.0, 20.0, 130.0, 149.9, -5.0,     # gas: <= 20 or >= 130; -5 wraps to 145
        70.0, 75.0, 80.0, 225.0,            # liquid: 70 <= x <= 80; 225 wraps to 75
        50.0, 69.9, 80.1, 100.0, 129.9,     # neither
    ])
    g, l = analysis.slab_densities(x, box, 70.0, 80.0, 20.0, 130.0)
    assert g == 5 / (21.0 * 21.0 * 40.0)
    assert l == 4 / (21.0 * 21.0 * 10.0)


def test_thermo_line_layout_and_tension():
    box = np.array([150.0, 21.0, 21.0])
    vol = box.prod()
    # isotropic kinetic tensor and virial: no tension
    text, row = apps.thermo_line(200, 0.005, 16000, box, [-5.0e4, -3.0e3, -3.0e3, -3.0e3],
                                 [6800.0] * 3, 12.5, 8.1e6)
    assert abs(row[10]) < 1e-12
    assert re.fullmatch(r"000000200 1\.000000e\+00( -?\d\.\d{7}e[+-]\d\d){9} 8\.10e\+06\n",
                        text), text
    # temperature = 2 K / 3 / (N - 3)
    assert abs(row[4] - 2 * 20400.0 / 3 / 15997.0) < 1e-12
    # a normal-pressure excess of dP over the lateral ones: tension dP * Lx / 2
    dP = 0.01
    w = np.array([-3.0e3 - dP * vol, -3.0e3, -3.0e3])
    _, row = apps.thermo_line(200, 0.005, 16000, box, [-5.0e4, *w], [6800.0] * 3, 0.0, 0)
    assert abs(row[7] - row[8] - dP) < 1e-12
    assert abs(row[10] - dP * 75.0) < 1e-9


def test_error_profile_bins_and_rms():
    ep = analysis.ErrorProfile((150.0, 21.0, 21.0), 2.0)
    assert ep.nx == 75 and ep.hx == 2.0
    x = np.array([0.5, 1.5, 151.0, -1.0, 3.0])       # 151 wraps to 1, -1 to 149
    dF = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0],
                   [0.0, 2.0, 0.0], [0.0, 0.0, 0.0]])
    ep.deposit(x, dF)
    xs, rms = ep.x_avg()
    assert xs[0] == 1.0 and xs[-1] == 149.0
    assert abs(rms[0] - np.sqrt((25.0 + 1.0 + 1.0) / 3.0)) < 1e-15
    assert rms[1] == 0.0 and rms[74] == 2.0
    assert ep.count.sum() == 5
