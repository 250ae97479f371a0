"""Independent checks of the oracle that the reference's own unit tests leave
open (CPU only).

Deblocking: the reference walks edges sequentially and in place
(loop_filter.rs:271-294).  The HIP kernels filter every edge of a pass at
once, which is the same computation only when no edge reads samples another
edge of the same pass writes (true for legal, size-aligned block and
transform geometry).  The oracle's walk is run forward, with every edge of a
pass reading the pass-start samples, and in reverse order on generated
frames, and all three must agree; on a transform grid that is not size-aligned the two must
differ, which shows the comparison can fail.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import gen_config

FORMATS = [  # (bit_depth, sx, sy, mono)
    (8, 1, 1, 0), (10, 1, 1, 0), (12, 0, 0, 0), (8, 1, 0, 0), (10, 1, 1, 1),
]
SIZES = [(96, 80), (100, 76), (200, 130), (256, 128)]


def _frame(seed, **kw):
    from oxav1 import Frame
    return Frame.generate(gen_config(**kw), seed)


def _lf_walks(oracle_mod, f):
    """(forward, all-edges-from-snapshot, reversed) deblocked planes."""
    from oxav1.frames import planes_struct
    lib = oracle_mod.load()
    ctx, g, keep = f.lf_args()
    out = []
    for fn in (lib.oo_loop_filter_frame, lib.oo_loop_filter_frame_parallel,
               lib.oo_loop_filter_frame_reversed):
        pl = [p.copy() for p in f.planes]
        fn(C.byref(ctx), C.byref(g), planes_struct(pl), len(pl))
        out.append(pl)
    return out


def _lf_both(oracle_mod, f):
    seq, par, _ = _lf_walks(oracle_mod, f)
    return seq, par


@pytest.mark.parametrize("geo", [0, 1], ids=["quadtree", "decoder_geometry"])
@pytest.mark.parametrize("fi", range(len(FORMATS)))
@pytest.mark.parametrize("si", range(len(SIZES)))
def test_deblock_sequential_equals_parallel(fi, si, geo, oracle_mod):
    """decoder_geometry adds 4:1 and A/B partitions, 64-sample transforms and
    var-tx trees (the adjacent-footprint argument holds for any legal
    transform tiling: the next edge along a row is a whole transform width
    away, and that width bounds both edges' filter lengths)."""
    bd, sx, sy, mono = FORMATS[fi]
    f = _frame(0x5E0 + 31 * fi + si + 7777 * geo, upscaled_width=SIZES[si][0],
               frame_height=SIZES[si][1], bit_depth=bd, subsampling_x=sx, subsampling_y=sy,
               mono=mono, noise_amp=8 << (bd - 8), exercise_segments=si % 2,
               exercise_delta_lf=fi % 2, sb128=si % 2, decoder_geometry=geo)
    seq, par, rev = _lf_walks(oracle_mod, f)
    assert any(not np.array_equal(a, b) for a, b in zip(seq, f.planes)), "nothing was filtered"
    for p, (a, b, c) in enumerate(zip(seq, par, rev)):
        assert np.array_equal(a, b), f"plane {p}: parallel walk, {(a != b).sum()} samples differ"
        assert np.array_equal(a, c), f"plane {p}: reversed walk, {(a != c).sum()} samples differ"


def test_deblock_parallel_differs_on_overlapping_edges(oracle_mod):
    """A transform grid that is not size-aligned (16x16 and 4x4 cells mixed at
    random on 4x4 blocks) puts a 16-wide transform edge next to a 4-wide one,
    so one edge's taps read samples its neighbour wrote: the in-place walk and
    the parallel walk must then diverge."""
    rng = np.random.default_rng(5)
    f = _frame(0x5E7, upscaled_width=64, frame_height=64, bit_depth=8, noise_amp=2)
    f.mi_size[:] = 0          # BLOCK_4X4: a block edge every 4 samples
    f.skip[:] = 0
    f.lf_tx_size[0][:] = rng.choice([0, 2], size=f.lf_tx_size[0].shape)  # TX_4X4 / TX_16X16
    f.planes[0][:] = 128 + rng.integers(0, 3, size=f.planes[0].shape)    # flat enough to filter
    seq, par = _lf_both(oracle_mod, f)
    assert not np.array_equal(seq[0], f.planes[0])
    assert not np.array_equal(seq[0], par[0])


def test_frame_driver_mirror_equals_oracle_chain(oracle_mod):
    """The Python frame-driver mirror over the oracle's five entry points
    equals the oracle's own chain, so the GPU test of the mirror over the
    drop-in calls compares like with like."""
    from conftest import gen_config
    from oxav1 import Frame, dropin
    calls = oracle_mod.calls()
    for kw, seed in ((dict(), 1),
                     (dict(bit_depth=10, noise_amp=32, superres_denom=12, upscaled_width=160,
                           frame_height=72, enable_fg=1, fg_num_y_points=8, fg_ar_lag=3,
                           fg_chroma_points=4, fg_overlap=1), 2),
                     (dict(mono=1, bit_depth=12, noise_amp=128), 3),
                     (dict(enable_lr=0, superres_denom=9, upscaled_width=120, frame_height=64), 4)):
        f = Frame.generate(gen_config(**kw), seed)
        ref_o, out_o, _ = oracle_mod.chain(f)
        ref_d, out_d = dropin.frame_driver_chain(calls, f)
        for a, b in zip(ref_o + out_o, ref_d + out_d):
            assert np.array_equal(a, b), kw


CHAIN_MATRIX = [  # (bit_depth, sx, sy, mono, superres_denom, fg)
    (8, 1, 1, 0, 8, 1), (10, 1, 1, 0, 12, 1), (12, 1, 1, 0, 16, 1), (8, 1, 0, 0, 9, 0),
    (10, 1, 0, 0, 8, 1), (12, 0, 0, 0, 13, 1), (8, 0, 0, 0, 8, 1), (10, 1, 1, 1, 11, 1),
    (12, 1, 1, 1, 8, 0),
]


@pytest.mark.parametrize("k", range(len(CHAIN_MATRIX)))
def test_chain_outputs_in_range(k, oracle_mod):
    """Every stage writes Clip1-ed values (the reason u8/u16 device storage
    is lossless): the chain's pre-grain and grain outputs stay within
    [0, 2^bd - 1] over depths, subsamplings, superres and film grain, with
    the decoder's geometry and the strongest noise.  Under the UBSan build
    of the oracle (tests/test_oracle_ubsan.py) this also runs every stage's
    arithmetic through the overflow and shift checks."""
    bd, sx, sy, mono, den, fg = CHAIN_MATRIX[k]
    f = _frame(0xC4A1 + k, upscaled_width=[136, 200, 96][k % 3], frame_height=[72, 104, 64][k % 3],
               bit_depth=bd, subsampling_x=sx, subsampling_y=sy, mono=mono, superres_denom=den,
               noise_amp=min(255, 64 << (bd - 8)), decoder_geometry=k % 2, sb128=k % 3 == 2,
               exercise_segments=1, exercise_delta_lf=1, cdef_skip_units=1, enable_fg=fg,
               fg_num_y_points=14, fg_ar_lag=3, fg_chroma_points=10 * (1 - mono),
               fg_overlap=k % 2, lr_unit_shift=k % 3, lr_uv_shift=k % 2)
    ref, out, _ = oracle_mod.chain(f)
    hi = (1 << bd) - 1
    for p in range(f.num_planes):
        assert ref[p].shape == out[p].shape == f.out_shape(p)
        for a in (ref[p], out[p]):
            assert a.min() >= 0 and a.max() <= hi, (p, a.min(), a.max())
