"""GPU parity of the five drop-in entry points against the CPU oracle.

Each test generates a seeded frame, runs one stage through the HIP library
(int32 PlaneBuffers staged to HBM) and through the oracle on the same
inputs, and requires byte-identical planes.  The matrix covers bit depths
8/10/12, subsampling 4:2:0/4:2:2/4:4:4/mono, odd frame sizes, superres
denominators, restoration unit sizes and the film-grain options.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import gen_config

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 0], ids=["narrowed", "int32"], autouse=True)
def staging(request):
    """Every test runs twice: with the calls' narrowed staging (u8/u16
    planes and kernels, the default) and with the int32 path forced."""
    from oxav1 import abi
    lib = abi.load()
    lib.oxav1_pf_dropin_narrowing(request.param)
    yield request.param
    lib.oxav1_pf_dropin_narrowing(1)

FORMATS = [  # (bit_depth, sx, sy, mono)
    (8, 1, 1, 0), (10, 1, 1, 0), (12, 1, 1, 0), (8, 1, 0, 0), (10, 0, 0, 0), (12, 1, 0, 0),
    (8, 1, 1, 1), (10, 0, 0, 0),
]
SIZES = [(96, 80), (100, 76), (64, 64), (200, 130)]


def _frame(seed, **kw):
    from oxav1 import Frame
    return Frame.generate(gen_config(**kw), seed)


def _eq(a_list, b_list, what):
    assert len(a_list) == len(b_list)
    for p, (a, b) in enumerate(zip(a_list, b_list)):
        assert a.shape == b.shape, (what, p, a.shape, b.shape)
        if not np.array_equal(a, b):
            diff = np.argwhere(a != b)
            y, x = diff[0]
            raise AssertionError(f"{what}: plane {p}: {len(diff)} samples differ, first at "
                                 f"({y},{x}): gpu {a[y, x]} oracle {b[y, x]}")


def _fmt_kw(fmt, size):
    bd, sx, sy, mono = fmt
    return dict(upscaled_width=size[0], frame_height=size[1], bit_depth=bd, subsampling_x=sx,
                subsampling_y=sy, mono=mono, noise_amp=8 << (bd - 8))


@pytest.mark.parametrize("geo", [0, 1], ids=["quadtree", "decoder_geometry"])
@pytest.mark.parametrize("fi", range(len(FORMATS)))
@pytest.mark.parametrize("si", range(len(SIZES)))
def test_loop_filter_parity(fi, si, geo, gpu_calls, oracle_mod):
    """decoder_geometry: 4:1 and A/B partitions, 64-sample transforms, var-tx."""
    seed = 0x1F00 + 17 * fi + si + 0x3000 * geo
    f = _frame(seed, exercise_segments=si % 2, exercise_delta_lf=(si + fi) % 2,
               decoder_geometry=geo, sb128=geo * (si % 2), **_fmt_kw(FORMATS[fi], SIZES[si]))
    got = gpu_calls.loop_filter_frame(f, f.planes)
    want = oracle_mod.calls().loop_filter_frame(f, f.planes)
    assert any(not np.array_equal(w, p) for w, p in zip(want, f.planes)), "oracle filtered nothing"
    _eq(got, want, "loop_filter_frame")


@pytest.mark.parametrize("fi", range(len(FORMATS)))
@pytest.mark.parametrize("si", range(len(SIZES)))
def test_cdef_parity(fi, si, gpu_calls, oracle_mod):
    seed = 0x2F00 + 17 * fi + si
    f = _frame(seed, cdef_skip_units=1, **_fmt_kw(FORMATS[fi], SIZES[si]))
    src = oracle_mod.calls().loop_filter_frame(f, f.planes)
    got = gpu_calls.cdef_frame(f, src)
    want = oracle_mod.calls().cdef_frame(f, src)
    _eq(got, want, "cdef_frame")


@pytest.mark.parametrize("seed", [0x5C01, 0x5C02])
def test_cdef_extremes_12bit(seed, gpu_calls, oracle_mod):
    """12-bit CDEF over flat, 0/4095 checkerboard and full-range noise bands
    (the largest tap differences, |d| = 4095, and saturated constrain), every
    unit filtered: the packed 16-bit filter holds differences, clamped taps
    and their weighted sums in 16-bit halves."""
    f = _frame(seed, bit_depth=12, cdef_skip_units=0, upscaled_width=200, frame_height=136)
    rng = np.random.default_rng(seed)
    src = [_extreme_plane(p.shape, rng, 4095) for p in f.planes]
    _eq(gpu_calls.cdef_frame(f, src), oracle_mod.calls().cdef_frame(f, src), "cdef_frame")


def _extreme_plane(shape, rng, maxv):
    """Flat maxv regions (the largest self-guided B: a2 = 1 over flat
    windows of the largest samples), 0/maxv checkerboards (the largest
    variance, A = 256) and full-range noise, in bands of 24 rows."""
    y = np.arange(shape[0])[:, None] // 24 % 3
    x = np.arange(shape[1])[None, :]
    checker = np.where((x + np.arange(shape[0])[:, None]) % 2 == 0, 0, maxv)
    noise = rng.integers(0, maxv + 1, size=shape)
    return np.where(y == 0, maxv, np.where(y == 1, checker, noise)).astype(np.int32)


@pytest.mark.parametrize("seed", [0x5A01, 0x5A02])
def test_loop_restoration_self_guided_extremes_12bit(seed, gpu_calls, oracle_mod):
    """Self-guided units only, 12-bit, on CDEF and deblocked planes built
    from the extremes above: the packed A/B words of neighbouring positions
    are summed before their fields are split, which is exact only while B
    stays below 2^20 (flat windows of 4095) and A at most 256."""
    f = _frame(seed, lr_types=4, bit_depth=12, upscaled_width=260, frame_height=140)
    deb, cdf = _lr_inputs(f, oracle_mod)
    rng = np.random.default_rng(seed)
    deb = [_extreme_plane(d.shape, rng, 4095) for d in deb]
    cdf = [_extreme_plane(c.shape, rng, 4095) for c in cdf]
    _eq(gpu_calls.loop_restoration_frame(f, deb, cdf),
        oracle_mod.calls().loop_restoration_frame(f, deb, cdf), "loop_restoration_frame")


@pytest.mark.parametrize("denom", [9, 12, 16])
@pytest.mark.parametrize("fi", [0, 1, 2, 3, 4, 6])
def test_superres_parity(denom, fi, gpu_calls, oracle_mod):
    seed = 0x3F00 + denom * 7 + fi
    f = _frame(seed, superres_denom=denom, **_fmt_kw(FORMATS[fi], (150, 70)))
    ctx = f.sr_ctx()
    shapes = [f.out_shape(p) for p in range(f.num_planes)]
    st_g, got = gpu_calls.upscale_frame(ctx, f.planes, shapes)
    st_o, want = oracle_mod.calls().upscale_frame(ctx, f.planes, shapes)
    assert st_g == st_o == 0
    _eq(got, want, "upscale_frame")


def test_superres_errors(gpu_calls, oracle_mod):
    from oxav1 import abi
    f = _frame(5, superres_denom=12, upscaled_width=96, frame_height=32)
    ctx = f.sr_ctx()
    shapes = [f.out_shape(p) for p in range(f.num_planes)]
    # UpscaledNotLarger
    bad = abi.SrCtx.from_buffer_copy(ctx)
    bad.upscaled_width = bad.frame_width - 2
    st_g, _ = gpu_calls.upscale_frame(bad, f.planes, shapes)
    st_o, _ = oracle_mod.calls().upscale_frame(bad, f.planes, shapes)
    assert st_g == st_o == abi.ERR_UPSCALED_NOT_LARGER
    # shape mismatch on plane 1: plane 0 is still written (superres.rs:309-316)
    shapes_bad = list(shapes)
    shapes_bad[1] = (shapes[1][0], shapes[1][1] + 1)
    st_g, got = gpu_calls.upscale_frame(ctx, f.planes, shapes_bad)
    st_o, want = oracle_mod.calls().upscale_frame(ctx, f.planes, shapes_bad)
    assert st_g == st_o == abi.ERR_PLANE_SHAPE_MISMATCH
    assert np.array_equal(got[0], want[0]) and want[0].any()
    # plane count
    st_g, _ = gpu_calls.upscale_frame(ctx, f.planes[:1], shapes[:1])
    st_o, _ = oracle_mod.calls().upscale_frame(ctx, f.planes[:1], shapes[:1])
    assert st_g == st_o == abi.ERR_PLANE_COUNT_MISMATCH


def _lr_inputs(f, oracle_mod):
    o = oracle_mod.calls()
    deb = o.loop_filter_frame(f, f.planes)
    cdf = o.cdef_frame(f, deb)
    return deb, cdf


@pytest.mark.parametrize("fi", range(len(FORMATS)))
@pytest.mark.parametrize("unit", [(0, 0), (1, 1), (2, 0)])
def test_loop_restoration_parity(fi, unit, gpu_calls, oracle_mod):
    seed = 0x4F00 + 31 * fi + unit[0] * 3 + unit[1]
    f = _frame(seed, lr_unit_shift=unit[0], lr_uv_shift=unit[1],
               **_fmt_kw(FORMATS[fi], (200, 150)))
    deb, cdf = _lr_inputs(f, oracle_mod)
    got = gpu_calls.loop_restoration_frame(f, deb, cdf)
    want = oracle_mod.calls().loop_restoration_frame(f, deb, cdf)
    _eq(got, want, "loop_restoration_frame")


@pytest.mark.parametrize("types", [2, 4, 6])
def test_loop_restoration_types(types, gpu_calls, oracle_mod):
    f = _frame(0x4A00 + types, lr_types=types, bit_depth=10, noise_amp=64,
               upscaled_width=260, frame_height=140)
    deb, cdf = _lr_inputs(f, oracle_mod)
    _eq(gpu_calls.loop_restoration_frame(f, deb, cdf),
        oracle_mod.calls().loop_restoration_frame(f, deb, cdf), "loop_restoration_frame")


@pytest.mark.parametrize("fi", [0, 1, 2, 3, 4, 6])
@pytest.mark.parametrize("opts", [(0, 0, 8), (1, 3, 6), (1, 2, 0), (0, 1, 10)])
def test_film_grain_parity(fi, opts, gpu_calls, oracle_mod):
    overlap, lag, cpts = opts
    seed = 0x5F00 + 13 * fi + lag
    f = _frame(seed, enable_fg=1, fg_overlap=overlap, fg_ar_lag=lag, fg_num_y_points=10,
               fg_chroma_points=cpts, **_fmt_kw(FORMATS[fi], (150, 98)))
    d = f.desc
    planes = [f.planes[p][: f.out_shape(p)[0], : f.out_shape(p)[1]].copy()
              for p in range(f.num_planes)]
    args = (d.fg, d.bit_depth, d.num_planes, d.subsampling_x, d.subsampling_y, d.matrix_coefficients)
    got = gpu_calls.film_grain_synthesis(*args, planes)
    want = oracle_mod.calls().film_grain_synthesis(*args, planes)
    assert any(not np.array_equal(w, p) for w, p in zip(want, planes))
    _eq(got, want, "film_grain_synthesis")


@pytest.mark.parametrize("clip", [0, 1], ids=["full_range", "clip_restricted"])
def test_film_grain_extremes_12bit(clip, gpu_calls, oracle_mod):
    """12-bit grain at full scaling over flat 0 / 4095, checkerboard and
    full-range bands: every blend saturates against one end of the range
    (full, or the restricted luma/chroma limits with clip_to_restricted_range)."""
    f = _frame(0x5E01 + clip, enable_fg=1, fg_overlap=1, fg_ar_lag=3, fg_num_y_points=6,
               fg_chroma_points=4, bit_depth=12, upscaled_width=200, frame_height=100)
    d = f.desc
    fg = d.fg
    fg.clip_to_restricted_range = clip
    for i in range(fg.num_y_points):
        fg.point_y_scaling[i] = 255
    for i in range(fg.num_cb_points):
        fg.point_cb_scaling[i] = 255
    for i in range(fg.num_cr_points):
        fg.point_cr_scaling[i] = 255
    rng = np.random.default_rng(0x5E01 + clip)
    planes = [_extreme_plane(f.out_shape(p), rng, 4095) for p in range(f.num_planes)]
    planes[0][:8] = 0  # a flat band at the low end as well
    args = (fg, d.bit_depth, d.num_planes, d.subsampling_x, d.subsampling_y, d.matrix_coefficients)
    want = oracle_mod.calls().film_grain_synthesis(*args, planes)
    assert any(not np.array_equal(w, p) for w, p in zip(want, planes))
    _eq(gpu_calls.film_grain_synthesis(*args, planes), want, "film_grain_synthesis")


def test_film_grain_chroma_from_luma(gpu_calls, oracle_mod):
    f = _frame(0x5A11, enable_fg=1, fg_overlap=1, fg_ar_lag=3, fg_num_y_points=6,
               upscaled_width=130, frame_height=70)
    fg = f.desc.fg
    fg.chroma_scaling_from_luma = 1
    fg.num_cb_points = fg.num_cr_points = 0
    planes = [f.planes[p][: f.out_shape(p)[0], : f.out_shape(p)[1]].copy() for p in range(3)]
    args = (fg, 8, 3, 1, 1, 0)
    _eq(gpu_calls.film_grain_synthesis(*args, planes),
        oracle_mod.calls().film_grain_synthesis(*args, planes), "fg chroma_scaling_from_luma")


def test_film_grain_scaling_points_out_of_order(gpu_calls, oracle_mod):
    """Scaling points that are not increasing (a stream the conformance rules
    forbid) build the tables the reference builds (film_grain.rs:480-526):
    ramps in segment order, skipped segments, later ramps winning."""
    f = _frame(0x5A22, enable_fg=1, fg_overlap=0, fg_ar_lag=1, fg_num_y_points=6, fg_chroma_points=4,
               upscaled_width=120, frame_height=66)
    fg = f.desc.fg
    xs, ys = [20, 90, 60, 60, 200, 150], [40, 200, 10, 90, 255, 30]
    for i, (x, y) in enumerate(zip(xs, ys)):
        fg.point_y_value[i], fg.point_y_scaling[i] = x, y
    fg.point_cb_value[0], fg.point_cb_value[1], fg.point_cb_value[2] = 100, 50, 120
    planes = [f.planes[p][: f.out_shape(p)[0], : f.out_shape(p)[1]].copy() for p in range(3)]
    args = (fg, 8, 3, 1, 1, 1)
    _eq(gpu_calls.film_grain_synthesis(*args, planes),
        oracle_mod.calls().film_grain_synthesis(*args, planes), "fg scaling points out of order")


@pytest.mark.parametrize("case", [
    dict(cfg=0),
    dict(kw=dict(bit_depth=10, noise_amp=32, superres_denom=12, upscaled_width=200, frame_height=96,
                 enable_fg=1, fg_num_y_points=8, fg_ar_lag=3, fg_chroma_points=4, fg_overlap=1)),
    dict(kw=dict(bit_depth=12, noise_amp=128, subsampling_y=0, superres_denom=10, upscaled_width=160,
                 frame_height=70, lr_unit_shift=1)),
    dict(kw=dict(mono=1, bit_depth=8, enable_fg=1, fg_num_y_points=5, fg_ar_lag=1)),
], ids=["cfg0", "sr_fg_10bit", "sr_12bit_422", "mono_fg"])
def test_frame_driver_through_dropin_calls(case, gpu_calls, oracle_mod):
    """The frame driver's post-filter sequence (frame_driver.rs:790-1116)
    driven through the five drop-in calls on the GPU equals the oracle's
    chain: the boundary composes the way the decoder would use it."""
    from oxav1 import Frame, dropin
    frame = Frame.generate(case["cfg"], 0x0A1F0000) if "cfg" in case else _frame(21, **case["kw"])
    ref_g, out_g = dropin.frame_driver_chain(gpu_calls, frame)
    ref_o, out_o, _ = oracle_mod.chain(frame)
    _eq(ref_g, ref_o, "pre-grain reference")
    _eq(out_g, out_o, "film-grain output")
    # the driver's planes reused over two frames (their previous contents
    # must not leak into the result)
    bufs = dropin.driver_buffers(frame)
    for rep in range(2):
        ref_b, out_b = dropin.frame_driver_chain(gpu_calls, frame, bufs=bufs)
        _eq(ref_b, ref_o, f"pre-grain reference, reused planes {rep}")
        _eq(out_b, out_o, f"film-grain output, reused planes {rep}")


@pytest.mark.parametrize("case", __import__("lf_edge_cases").all_cases(), ids=lambda c: c.name)
def test_loop_filter_edge_cases(case, gpu_calls, oracle_mod):
    """The hand-derived edge-decision cases (onScreen gates, chroma
    planeSize, applyFilter, filter size from both sides, prev-side level
    fallback) through the drop-in call: the GPU output is the expected
    plane exactly."""
    from oxav1 import abi
    want = case.expected(oracle_mod.load())
    st, got = case.run(abi.load().oxav1_pf_loop_filter_frame)
    abi.check(st, "loop_filter_frame")
    _eq(got, want, case.name)


@pytest.mark.parametrize("stage", ["lf", "cdef", "sr", "lr", "fg"])
def test_out_of_range_samples_take_the_int32_path(stage, gpu_calls, oracle_mod):
    """A sample outside [0, 2^bd) (here 2^bd + 4 and -3) makes a call stage
    int32 planes and run the int32 kernels: the output still equals the
    oracle's on the same int32 input."""
    f = _frame(0x0F0 + len(stage), upscaled_width=160, frame_height=72, bit_depth=10, noise_amp=32,
               superres_denom=11, enable_fg=1, fg_num_y_points=8, fg_ar_lag=2, fg_chroma_points=4)
    hi = (1 << 10) + 4
    oc = oracle_mod.calls()
    if stage in ("lf", "cdef"):
        pl = [p.copy() for p in f.planes]
        pl[0][9, 13] = hi
        pl[1][4, 5] = -3
        fn = "loop_filter_frame" if stage == "lf" else "cdef_frame"
        _eq(getattr(gpu_calls, fn)(f, pl), getattr(oc, fn)(f, pl), stage)
    elif stage == "sr":
        pl = [p.copy() for p in f.planes]
        pl[0][2, 20] = hi
        shapes = [f.out_shape(p) for p in range(f.num_planes)]
        st_g, out_g = gpu_calls.upscale_frame(f.sr_ctx(), pl, shapes)
        st_o, out_o = oc.upscale_frame(f.sr_ctx(), pl, shapes)
        assert st_g == st_o == 0
        _eq(out_g, out_o, "sr")
    elif stage == "lr":
        shapes = [f.out_shape(p) for p in range(f.num_planes)]
        rng = np.random.default_rng(3)
        cur = [rng.integers(0, 1 << 10, size=sh).astype(np.int32) for sh in shapes]
        cdf = [rng.integers(0, 1 << 10, size=sh).astype(np.int32) for sh in shapes]
        cdf[0][40, 40] = hi
        cur[2][10, 11] = -3
        _eq(gpu_calls.loop_restoration_frame(f, cur, cdf, lr_init=cdf),
            oc.loop_restoration_frame(f, cur, cdf, lr_init=cdf), "lr")
    else:
        d = f.desc
        shapes = [f.out_shape(p) for p in range(f.num_planes)]
        rng = np.random.default_rng(4)
        pl = [rng.integers(0, 1 << 10, size=sh).astype(np.int32) for sh in shapes]
        pl[0][7, 8] = hi
        pl[1][3, 3] = -3
        args = (d.fg, d.bit_depth, d.num_planes, d.subsampling_x, d.subsampling_y, d.matrix_coefficients)
        _eq(gpu_calls.film_grain_synthesis(*args, pl), oc.film_grain_synthesis(*args, pl), "fg")
