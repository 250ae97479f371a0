"""Output-exact pins of the oracle's deblocking edge decisions
(loop_filter_edge, loop_filter.rs:302-423), and mutants of the oracle that
each pin catches.

The cases (tests/lf_edge_cases.py) state by hand which 4-sample edge
segments a geometry filters, with which filter size and strength; the
expected plane is the input with exactly those samples filtered by the
sample-level primitives.  The mutation test rebuilds the oracle with one
plausible mistake in the edge decisions at a time and requires at least one
case to fail on it, so the pins are not vacuous.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import lf_edge_cases as lc

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.fixture(scope="module")
def oo(oracle_mod):
    return oracle_mod.load()


def _mismatch(case, got, want):
    for p, (g, w) in enumerate(zip(got, want)):
        if not np.array_equal(g, w):
            y, x = np.argwhere(g != w)[0]
            return f"{case.name}: plane {p} first differs at ({y},{x}): {g[y, x]} vs {w[y, x]}"
    return None


def check_cases(walk, prim, cases=None):
    """Names of the cases whose walk output differs from the expected plane."""
    bad = []
    for case in cases or lc.all_cases():
        _, got = case.run(walk)
        msg = _mismatch(case, got, case.expected(prim))
        if msg:
            bad.append(msg)
    return bad


@pytest.mark.parametrize("case", lc.all_cases(), ids=lambda c: c.name)
def test_edge_case_expected_output(case, oo):
    want = case.expected(oo)
    changed = sum(int((w != p).sum()) for w, p in zip(want, case.planes))
    assert changed > 0, "the case's edges filter nothing: it would pin nothing"
    for fn in (oo.oo_loop_filter_frame, oo.oo_loop_filter_frame_parallel,
               oo.oo_loop_filter_frame_reversed):
        _, got = case.run(fn)
        msg = _mismatch(case, got, want)
        assert msg is None, msg


def test_edge_cases_cover_every_decision():
    """The cases together exercise chroma planes, all three filter sizes, the
    skip/intra combinations and a frame smaller than its padded planes."""
    cases = lc.all_cases()
    sizes = {e[4] for c in cases for e in c.edges}
    assert sizes == {4, 8, 16}
    assert any(e[0] > 0 for c in cases for e in c.edges)
    assert any(c.ctx.frame_width < c.ctx.mi_cols * 4 for c in cases)


# One plausible mistake each in oo_loop_filter_edge / oo_lf_filter_size
# (oracle/oxav1_oracle.c): (what, original text, mutated text).
MUTANTS = [
    ("chroma reads the even mi cell (no row |= subY, col |= subX)",
     "row |= sy;\n    col |= sx;", ""),
    ("onScreen compares x > FrameWidth",
     "if (x >= ctx->frame_width) return;", "if (x > ctx->frame_width) return;"),
    ("onScreen uses the padded width",
     "if (x >= ctx->frame_width) return;", "if (x >= ctx->mi_cols * MI_SIZE) return;"),
    ("onScreen ignores FrameHeight",
     "if (y >= ctx->frame_height) return;", ""),
    ("chroma planeSize = MiSize (before the r387 fix)",
     "int32_t plane_size = ps >= 0 ? ps : (int32_t)mi_size;",
     "int32_t plane_size = (int32_t)mi_size; (void)ps;"),
    ("applyFilter ignores skip",
     "int apply = is_tx_edge ? (is_block_edge || !skip || is_intra) : 0;",
     "int apply = is_tx_edge;"),
    ("applyFilter ignores is_intra",
     "int apply = is_tx_edge ? (is_block_edge || !skip || is_intra) : 0;",
     "int apply = is_tx_edge ? (is_block_edge || !skip) : 0;"),
    ("isTxEdge from the block width",
     "uint32_t tw = oo_tx_width[tx_sz];", "uint32_t tw = lf_block_width(plane_size);"),
    ("filter size from the current side only",
     "uint32_t fs = oo_lf_filter_size(tx_sz, prev_tx_sz, pass, plane);",
     "uint32_t fs = oo_lf_filter_size(tx_sz, tx_sz, pass, plane);"),
    ("no prev-side level fallback",
     "if (s.lvl == 0)\n        s = oo_lf_adaptive_filter_strength",
     "if (0)\n        s = oo_lf_adaptive_filter_strength"),
    ("strength from the prev side always",
     "oo_lf_strength s = oo_lf_adaptive_filter_strength(ctx, g, row, col, plane, pass);",
     "oo_lf_strength s = oo_lf_adaptive_filter_strength(ctx, g, (uint32_t)prev_row, "
     "(uint32_t)prev_col, plane, pass);"),
]


@pytest.fixture(scope="module")
def mutant_builder(tmp_path_factory):
    gcc = shutil.which("gcc") or shutil.which("cc")
    if gcc is None:
        pytest.skip("no C compiler")
    src = open(os.path.join(ORACLE_DIR, "oxav1_oracle.c")).read()
    root = tmp_path_factory.mktemp("lf_mutants")

    def build(i, old, new):
        assert src.count(old) == 1, f"mutation site not unique: {old!r}"
        d = root / f"m{i}"
        d.mkdir()
        c = d / "oxav1_oracle.c"
        c.write_text(src.replace(old, new))
        so = d / "libmut.so"
        subprocess.run([gcc, "-O1", "-std=c11", "-fPIC", "-shared", "-w", f"-I{ORACLE_DIR}",
                        "-o", str(so), str(c), "-lm"], check=True)
        return str(so)
    return build


@pytest.mark.parametrize("i", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_pins_catch_mutant(i, mutant_builder, oo, oracle_mod):
    what, old, new = MUTANTS[i]
    mut = oracle_mod.load_path(mutant_builder(i, old, new))
    bad = check_cases(mut.oo_loop_filter_frame, oo)
    assert bad, f"no edge pin fails on the mutant '{what}'"
