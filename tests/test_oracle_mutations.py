"""The oracle's pins catch plausible restatement mistakes.

Each case plants one mistake into a copy of oracle/smaug_oracle.c (CPU only:
the copy is built with gcc into a temporary directory and loaded in place of
the oracle), then reruns the closed-form checks of test_oracle.py against
it.  A case passes when at least one check fails on the planted mistake, and
every check still passes on the unchanged source.
"""
import os
import shutil
import subprocess

import pytest

import test_oracle as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "smaug_oracle.c")

CHECKS = T.SOURCE_CHECKS

# (name, exact text in smaug_oracle.c, replacement)
MUTANTS = [
    ("rho_source_sign",
     "out[c] += (nuR[c] * tA[c] - nuL[c] * tB[c]) / g->dx[dim][coord(dim, i_, j_, k_)];",
     "out[c] -= (nuR[c] * tA[c] - nuL[c] * tB[c]) / g->dx[dim][coord(dim, i_, j_, k_)];"),
    ("energy_source_sign",
     "re[c] += (nuR[c] * tA[c] - nuL[c] * tB[c]) / g->dx[dim][coord(dim, i_, j_, k_)];",
     "re[c] -= (nuR[c] * tA[c] - nuL[c] * tB[c]) / g->dx[dim][coord(dim, i_, j_, k_)];"),
    ("momentum_diagonal_sign",
     "om[c] += (rr * nuR[c] * tA[c] - rl * nuL[c] * tB[c]) / dxl / 2;",
     "om[c] -= (rr * nuR[c] * tA[c] - rl * nuL[c] * tB[c]) / dxl / 2;"),
    ("momentum_diagonal_heating_dropped",
     "re[c] += (mr * nuR[c] * tA[c] - ml * nuL[c] * tB[c]) / dxl / 2;",
     "(void)mr; (void)ml;"),
    ("momentum_offdiagonal_factor",
     "double g6 = tA[c] * (nuR[c] + nuL[c]) / 4.0;",
     "double g6 = tA[c] * (nuR[c] + nuL[c]) / 2.0;"),
    ("momentum_offdiagonal_heating_dropped",
     "re[c] += tA[c];",
     "re[c] += 0.0 * tA[c];"),
    ("momentum_offdiagonal_transposed",
     "g13(&G, g, tB, ii, tA);",
     "g13(&G, g, tB, dim, tA);"),
    ("magnetic_diagonal_sign",
     "ob[c] += 1.0 * (tA[c] * nuR[c] - tB[c] * nuL[c]) / dxl;",
     "ob[c] += -1.0 * (tA[c] * nuR[c] - tB[c] * nuL[c]) / dxl;"),
    ("magnetic_diagonal_heating_dropped",
     "re[c] += 1.0 * (br * tA[c] * nuR[c] - bl * tB[c] * nuL[c]) / dxl;",
     "(void)br; (void)bl;"),
    ("magnetic_offdiagonal_sign",
     "ob[c] += -1.0 * tA[c];",
     "ob[c] += 1.0 * tA[c];"),
    ("magnetic_offdiagonal_nu_mean",
     "double t3 = tA[c] * (nuL[c] + nuR[c]) / 2;",
     "double t3 = tA[c] * (nuL[c] + nuR[c]) / 4;"),
    ("nu_left_window_centre",
     "int a = (side == 0) ? x : x - 1;",
     "int a = (side == 0) ? x : x;"),
    ("nu_uses_wrong_cmax",
     "nu_from_q(&G, g, q, fld, dim, cmax[dim], nuR, nuL);\n                for (int ii1",
     "nu_from_q(&G, g, q, fld, dim, cmax[0], nuR, nuL);\n                for (int ii1"),
    ("numax_right_side_only",
     "if (nuL[c] > m) m = nuL[c];",
     "(void)0;"),
    ("numax_left_side_only",
     "if (nuR[c] > m) m = nuR[c];",
     "(void)0;"),
    ("numax_visits_b_along_own_dim",
     "if (fld >= F_B(D, 0) && fld - F_B(D, 0) == dim) continue;",
     "(void)0;"),
    ("numax_one_cmax",
     "nu_from_q(&G, g, q, fld, dim, cmax[dim], nuR, nuL);\n            for (long c",
     "nu_from_q(&G, g, q, fld, dim, cmax[0], nuR, nuL);\n            for (long c"),
    ("dtvisc_dx_not_squared",
     "double t = 0.25 * g->dx[d][x] * g->dx[d][x] / numax;",
     "double t = 0.25 * g->dx[d][x] / numax;"),
    ("dtvisc_first_cell_only",
     "for (int x = 0; x < G.n[d]; x++) {\n            double t = 0.25",
     "for (int x = 0; x < 1; x++) {\n            double t = 0.25"),
    ("cont_copies_second_cell",
     "if (x == 0 || x == 1) src = 2;",
     "if (x == 0 || x == 1) src = 3;"),
    ("fixed_faces_not_restored",
     "if (ghost && qf) q[c] = qf[c];",
     "(void)qf;"),
    ("maxcourant_edge_dx",
     "cfast_at(&G, g, w, bg, c, d) / g->dx[d][coord(d, i_, j_, k_)];",
     "cfast_at(&G, g, w, bg, c, d) / g->dx[d][0];"),
]


def _build(tmp_path, name, text):
    d = tmp_path / name
    d.mkdir()
    shutil.copy(os.path.join(ROOT, "oracle", "smaug_oracle.h"), d / "smaug_oracle.h")
    (d / "smaug_oracle.c").write_text(text)
    so = d / "liboracle_mut.so"
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11",
                           "-shared", "-o", str(so), str(d / "smaug_oracle.c"), "-lm"])
    return str(so)


def _failing(lib_path):
    bad = []
    for chk in CHECKS:
        try:
            chk(lib_path)
        except AssertionError:
            bad.append(chk.__name__)
    return bad


def test_unchanged_source_passes_every_check(tmp_path):
    so = _build(tmp_path, "plain", open(SRC).read())
    assert _failing(so) == []


@pytest.mark.parametrize("name,orig,mut", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(tmp_path, name, orig, mut):
    text = open(SRC).read()
    assert text.count(orig) == 1, f"{name}: the planted site is not unique in the oracle"
    so = _build(tmp_path, name, text.replace(orig, mut))
    assert _failing(so), f"{name}: no check fails on the planted mistake"
