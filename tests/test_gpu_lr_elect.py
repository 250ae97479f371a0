"""The encoder's loop-restoration election on the GPU (encoder/lr_elect.rs):
oxav1_pf_lr_elect against the oracle's elect_lr field for field — the plan
(elected flag, frame restoration types, every unit's filter, d, d_pre) and
every unit's candidate record (the SSD of the unfiltered unit, of the fitted
Wiener filter and of all 16 self-guided sets, the fitted taps and weights) —
across bit depths, subsamplings, monochrome, odd sizes, detail and noise
content and lambdas from 0 to one that elects nothing; and
oxav1_pf_lr_apply_plan against the oracle's apply_lr_plan (restored planes
and SSD)."""
import concurrent.futures as cf

import numpy as np
import pytest

from oxav1 import abi
from oxav1.elect import apply_lr_plan, elect_lr, synthetic_lr_input

pytestmark = pytest.mark.gpu

CASES = [  # seed, width, height, bit depth, ssx, ssy, mono, lambda, detail
    (1, 100, 76, 8, 1, 1, False, 40, True),
    (2, 130, 98, 10, 1, 0, False, 300, True),
    (3, 96, 100, 12, 0, 0, True, 800, True),
    (4, 200, 130, 10, 0, 0, False, 0, False),
    (5, 257, 161, 8, 1, 1, False, 150, True),
    (6, 192, 128, 8, 1, 1, False, 613, True),
    (7, 150, 70, 12, 1, 1, False, 1 << 40, False),
    (8, 64, 64, 10, 1, 1, False, 60, False),
]


def _id(c):
    return f"{c[1]}x{c[2]}-{c[3]}bit-{c[4]}{c[5]}{'-mono' if c[6] else ''}-l{c[7]}{'' if c[8] else '-noise'}"


def _compare(gpu, ref, label):
    print(f"compare {label}", flush=True)
    assert gpu.elected == ref.elected, label
    assert gpu.frame_restoration_type == ref.frame_restoration_type, label
    assert (gpu.uses_lr, gpu.uses_chroma_lr) == (ref.uses_lr, ref.uses_chroma_lr), label
    assert gpu.loop_restoration_size == ref.loop_restoration_size, label
    assert (gpu.unit_rows, gpu.unit_cols) == (ref.unit_rows, ref.unit_cols), label
    for p in range(len(ref.cands)):
        for k, (cg, cr) in enumerate(zip(gpu.cands[p], ref.cands[p])):
            assert cg == cr, (label, p, k)
        assert gpu.units[p] == ref.units[p], (label, p)
    assert (gpu.d, gpu.d_pre) == (ref.d, ref.d_pre), label


@pytest.mark.parametrize("c", CASES, ids=_id)
def test_lr_elect_matches_oracle(c, oracle_mod):
    seed, w, h, bd, sx, sy, mono, lam, detail = c
    inp = synthetic_lr_input(seed, w, h, bd, sx, sy, mono, lambda_=lam, detail=detail)
    print(f"lr elect {_id(c)}", flush=True)
    gpu = elect_lr(inp)
    ref = elect_lr(inp, oracle_mod.load(), "oo_lr_elect")
    _compare(gpu, ref, _id(c))
    rg, sg = apply_lr_plan(gpu, inp)
    rr, sr = apply_lr_plan(ref, inp, oracle_mod.load(), "oo_lr_apply_plan")
    assert sg == sr == gpu.d
    for p in range(inp.num_planes):
        assert (rg[p] == rr[p]).all(), p


def test_lr_elect_640x360(oracle_mod):
    """A 640x360 10-bit 4:2:0 frame (61 units a plane set): the oracle runs
    in a worker thread while the GPU elects, then field for field."""
    inp = synthetic_lr_input(21, 640, 360, 10, 1, 1, lambda_=400)
    print("lr elect 640x360 10-bit 4:2:0", flush=True)
    with cf.ThreadPoolExecutor(1) as ex:
        fut = ex.submit(elect_lr, inp, oracle_mod.load(), "oo_lr_elect")
        gpu = elect_lr(inp)
        ref = fut.result()
    _compare(gpu, ref, "640x360")
    assert gpu.elected


def test_lr_elect_repeatable_and_rejects_bad_input():
    inp = synthetic_lr_input(9, 100, 76, 10, 1, 1, lambda_=100)
    a, b = elect_lr(inp), elect_lr(inp)
    assert (a.units, a.cands, a.d) == (b.units, b.cands, b.d)
    inp.restoration_type_cdf = (0, 22581, 32768, 0)
    with pytest.raises(abi.Oxav1Error) as e:
        elect_lr(inp)
    assert e.value.status == abi.ERR_INVALID_ARG
    inp.restoration_type_cdf = (9413, 22581, 32768, 0)
    inp.cdef[1] = inp.cdef[1][:, :-1]
    inp.curr[1] = inp.curr[1][:, :-1]
    inp.input[1] = inp.input[1][:, :-1]
    with pytest.raises(abi.Oxav1Error) as e:
        elect_lr(inp)
    assert e.value.status == abi.ERR_UNSUPPORTED
