"""Spec tables: the oracle's and the HIP library's copies against the values
parsed from the reference sources (tests/golden/spec_tables.json, written by
tests/golden/make_table_fixture.py), plus closed-form properties."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "spec_tables.json")))
CHECKED = ["GAUSSIAN_SEQUENCE", "UPSCALE_FILTER", "SGR_PARAMS", "CDEF_DIRECTIONS", "CDEF_UV_DIR",
           "NUM_4X4_BLOCKS_WIDE", "NUM_4X4_BLOCKS_HIGH", "TX_WIDTH", "TX_HEIGHT",
           "SUBSAMPLED_SIZE", "MAX_TX_SIZE_RECT"]


def _ck(vals):
    return sum((i + 1) * v for i, v in enumerate(vals))


@pytest.mark.parametrize("name", CHECKED + ["DIV_TABLE"])
def test_oracle_tables_match_reference(name, oracle_mod):
    assert oracle_mod.load().oo_table_checksum(name.encode()) == _ck(FIX[name])


@pytest.mark.parametrize("name", CHECKED)
def test_library_tables_match_reference(name):
    from oxav1 import abi
    assert abi.load().oxav1_pf_table_checksum(name.encode()) == _ck(FIX[name])


def test_fixture_matches_reference_sources_when_present():
    """Regenerate the fixture from the reference (when mounted) and compare."""
    ref = "/root/reference/src/film_grain_tables.rs"
    if not os.path.exists(ref):
        pytest.skip("reference sources not mounted here")
    import subprocess
    import sys
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        script = os.path.join(ROOT, "tests", "golden", "make_table_fixture.py")
        code = open(script).read().replace("HERE = os.path.dirname(os.path.abspath(__file__))",
                                           f"HERE = {td!r}")
        p = os.path.join(td, "gen.py")
        open(p, "w").write(code)
        subprocess.run([sys.executable, p, "/root/reference"], check=True, capture_output=True)
        fresh = json.load(open(os.path.join(td, "spec_tables.json")))
    assert fresh == FIX


def test_upscale_filter_closed_forms():
    f = np.array(FIX["UPSCALE_FILTER"]).reshape(64, 8)
    assert (f.sum(axis=1) == 128).all()  # superres.rs:533-541
    assert list(f[0]) == [0, 0, 0, 128, 0, 0, 0, 0]
    # the half-sample phase is symmetric about the centre taps
    assert f[32][3] == f[32][4] == 79 and f[32][2] == f[32][5] == -20
    # phase k/64 moves the centre of mass by k/64 sample (x128), up to rounding
    com = (f * np.arange(8)).sum(axis=1)
    assert com[0] == 3 * 128 and (np.abs(com - (384 + 2 * np.arange(64))) <= 8).all()


def test_div_table_and_cdef_closed_forms():
    assert FIX["DIV_TABLE"] == [0] + [840 // i for i in range(1, 9)]
    uv = np.array(FIX["CDEF_UV_DIR"]).reshape(2, 2, 8)
    assert list(uv[0][0]) == list(range(8)) and list(uv[1][1]) == list(range(8))
    d = np.array(FIX["CDEF_DIRECTIONS"]).reshape(8, 2, 2)
    # the radius-2 tap continues the radius-1 step along the direction
    for k in range(8):
        assert abs(d[k][1][0]) <= 2 and abs(d[k][1][1]) <= 2
    assert FIX["CDEF_PRI_TAPS"] == [4, 2, 3, 3] and FIX["CDEF_SEC_TAPS"] == [2, 1, 2, 1]


def test_gaussian_sequence_shape():
    g = np.array(FIX["GAUSSIAN_SEQUENCE"])
    assert len(g) == 2048 and (g % 4 == 0).all() and g.min() >= -2048 and g.max() <= 2047
    assert abs(g.mean()) < 40 and 400 < g.std() < 600


def test_sgr_params_closed_forms():
    s = np.array(FIX["SGR_PARAMS"]).reshape(16, 4)
    assert set(s[:, 0]) <= {0, 2} and set(s[:, 2]) <= {0, 1}
    assert list(s[10]) == [0, 0, 1, 5] and list(s[14]) == [2, 30, 0, 0]
    assert FIX["WIENER_TAPS_MIN"] == [-5, -23, -17] and FIX["WIENER_TAPS_MAX"] == [10, 8, 46]
    assert FIX["SGRPROJ_XQD_MIN"] == [-96, -32] and FIX["SGRPROJ_XQD_MAX"] == [31, 95]


def test_block_tables_closed_forms():
    w = np.array(FIX["NUM_4X4_BLOCKS_WIDE"])
    h = np.array(FIX["NUM_4X4_BLOCKS_HIGH"])
    assert ((w & (w - 1)) == 0).all() and ((h & (h - 1)) == 0).all()
    ss = np.array(FIX["SUBSAMPLED_SIZE"]).reshape(22, 2, 2)
    assert (ss[:, 0, 0] == np.arange(22)).all()  # luma keeps its size
    tw, th = np.array(FIX["TX_WIDTH"]), np.array(FIX["TX_HEIGHT"])
    for b, t in enumerate(FIX["MAX_TX_SIZE_RECT"]):
        assert tw[t] == min(64, 4 * w[b]) and th[t] == min(64, 4 * h[b])
