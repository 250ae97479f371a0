"""GPU parity of the whole chain (engine, u8/u16 HBM planes) against the
oracle's restatement of frame_driver.rs:790-1116, for the plain
(one launch per stage per frame), batched persistent and hipGraph paths.
"""
import numpy as np
import pytest

from conftest import gen_config

pytestmark = pytest.mark.gpu


def _check(dev_frames, oracle_mod, what):
    for i, df in enumerate(dev_frames):
        ref_g, out_g = df.result()
        ref_o, out_o, _ = oracle_mod.chain(df.frame)
        for p in range(df.frame.num_planes):
            for name, a, b in (("ref", ref_g[p], ref_o[p]), ("out", out_g[p], out_o[p])):
                if not np.array_equal(a, b):
                    d = np.argwhere(a != b)
                    y, x = d[0]
                    raise AssertionError(f"{what} frame {i} plane {p} {name}: {len(d)} differ, first "
                                         f"({y},{x}) gpu {a[y, x]} oracle {b[y, x]}")


def _frames(cfgs):
    from oxav1 import Frame
    from oxav1.engine import DeviceFrame
    return [DeviceFrame(Frame.generate(c, s)) for c, s in cfgs]


SMALL = [
    (dict(), 1),
    (dict(bit_depth=10, noise_amp=32, enable_fg=1, fg_num_y_points=8, fg_ar_lag=3,
          fg_chroma_points=4, fg_overlap=1), 2),
    (dict(bit_depth=12, subsampling_x=1, subsampling_y=0, noise_amp=128, upscaled_width=130,
          frame_height=66), 3),
    (dict(bit_depth=8, subsampling_x=0, subsampling_y=0, superres_denom=13, upscaled_width=160,
          frame_height=72, enable_fg=1, fg_num_y_points=5, fg_ar_lag=1, fg_chroma_points=3), 4),
    (dict(mono=1, bit_depth=10, noise_amp=32, superres_denom=10, upscaled_width=200,
          frame_height=90), 5),
    (dict(sb128=1, lr_unit_shift=1, lr_uv_shift=1, superres_denom=16, upscaled_width=256,
          frame_height=136, exercise_segments=1, exercise_delta_lf=1, cdef_skip_units=1), 6),
]


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_chain_small_mixed(mode, oracle_mod):
    from oxav1.engine import Engine
    frames = _frames([(gen_config(**kw), s) for kw, s in SMALL])
    eng = Engine(max_frames=len(frames))
    if mode == 0:
        for df in frames:
            eng.prepare([df])
            eng.run(0)
            eng.sync()
    else:
        eng.prepare(frames)
        eng.run(mode)
        eng.sync()
        if mode == 2:  # replays: same inputs, same outputs, every time
            for _ in range(3):
                _check(frames, oracle_mod, "graph replay")
                for df in frames:
                    df.clear_outputs()
                eng.run(2)
                eng.sync()
    _check(frames, oracle_mod, f"mode {mode}")
    eng.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_chain_padded_pitch(mode, oracle_mod):
    """Caller planes with a row pitch of the width rounded up to four
    samples plus four, the padding at the largest sample value: deblocking
    loads its input four samples per lane from such rows (padded planes are
    a multiple of four samples wide), and the cropped outputs, often of odd
    width, are written into them; nothing may read or write the padding."""
    from oxav1 import Frame
    from oxav1.engine import DeviceFrame, Engine
    cfgs = [(gen_config(**kw), s) for kw, s in SMALL]
    cfgs.append((gen_config(upscaled_width=226, frame_height=118, noise_amp=4), 7))  # chroma 114 wide
    cfgs.append((gen_config(bit_depth=10, upscaled_width=198, frame_height=70, subsampling_x=0,
                            subsampling_y=0, noise_amp=8), 8))
    frames = [DeviceFrame(Frame.generate(c, s), pad_pitch=True) for c, s in cfgs]
    assert any(df.frame.out_shape(p)[1] % 4 for df in frames for p in range(df.frame.num_planes))
    eng = Engine(max_frames=len(frames))
    if mode == 0:
        for df in frames:
            eng.prepare([df])
            eng.run(0)
            eng.sync()
    else:
        eng.prepare(frames)
        eng.run(mode)
        eng.sync()
    _check(frames, oracle_mod, f"padded pitch mode {mode}")
    eng.close()


GEO_FORMATS = [  # (bit_depth, sx, sy, mono)
    (8, 1, 1, 0), (10, 1, 1, 0), (12, 1, 1, 0), (8, 1, 0, 0), (10, 0, 0, 0), (12, 1, 0, 0),
    (8, 1, 1, 1), (10, 0, 0, 0),
]


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_chain_decoder_geometry(mode, oracle_mod):
    """The decoder's block and transform geometry (4:1 and A/B partitions,
    transforms up to 64 samples, intra tx_depth, var-tx trees) in all eight
    formats, 64x64 and 128x128 superblocks: plain, batched and graph
    replays equal the oracle."""
    from oxav1.engine import Engine
    cfgs = []
    for i, (bd, sx, sy, mono) in enumerate(GEO_FORMATS):
        cfgs.append((gen_config(upscaled_width=[200, 136, 256, 168][i % 4],
                                frame_height=[130, 96, 144, 72][i % 4], bit_depth=bd,
                                subsampling_x=sx, subsampling_y=sy, mono=mono,
                                noise_amp=8 << (bd - 8), decoder_geometry=1, sb128=i % 2,
                                superres_denom=8 + (i % 3) * 2, exercise_segments=i % 3 == 1,
                                enable_fg=i % 2, fg_num_y_points=6, fg_ar_lag=i % 4,
                                fg_chroma_points=3 * (1 - mono)), 0x6E0 + i))
    frames = _frames(cfgs)
    eng = Engine(max_frames=len(frames))
    if mode == 0:
        for df in frames:
            eng.prepare([df])
            eng.run(0)
            eng.sync()
    else:
        eng.prepare(frames)
        eng.run(mode)
        eng.sync()
        if mode == 2:
            for _ in range(2):
                _check(frames, oracle_mod, "decoder geometry, graph replay")
                for df in frames:
                    df.clear_outputs()
                eng.run(2)
                eng.sync()
    _check(frames, oracle_mod, f"decoder geometry, mode {mode}")
    eng.close()


@pytest.mark.parametrize("sizes", [[(16, 16)], [(24, 16), (64, 40), (16, 24)], [(40, 24)] * 11])
def test_chain_batches_smaller_than_the_claim_partitions(sizes, oracle_mod):
    """Batched launches whose stages have fewer work items than there are
    claim partitions (empty partitions, items stolen across partitions)
    and more frames than partitions: every frame equals the oracle."""
    from oxav1.engine import Engine
    frames = _frames([(gen_config(upscaled_width=w, frame_height=h, superres_denom=8 + (i % 3),
                                  enable_fg=i % 2, fg_num_y_points=4), 0x7A00 + i)
                      for i, (w, h) in enumerate(sizes)])
    eng = Engine(max_frames=len(frames))
    eng.prepare(frames)
    for _ in range(2):
        for df in frames:
            df.clear_outputs()
        eng.run(2)
        eng.sync()
        _check(frames, oracle_mod, f"{len(frames)} small frames")
    eng.close()


@pytest.mark.parametrize("cfg", [0, 1])
def test_chain_config(cfg, oracle_mod):
    from oxav1 import Frame
    from oxav1.engine import DeviceFrame, Engine
    df = DeviceFrame(Frame.generate(cfg, 0x0A1F0000 + cfg))
    eng = Engine(max_frames=1)
    eng.prepare([df])
    eng.run(1)
    eng.sync()
    _check([df], oracle_mod, f"cfg {cfg}")
    eng.close()


def test_chain_4k_10bit_superres(oracle_mod):
    """3840x2160 10-bit, superres 12 (coded 2560), LR 128, film grain overlap."""
    from oxav1 import Frame
    from oxav1.engine import DeviceFrame, Engine
    df = DeviceFrame(Frame.generate(2, 0x0A1F0002))
    assert df.desc.frame_width == 2560 and df.desc.upscaled_width == 3840
    eng = Engine(max_frames=1)
    eng.prepare([df])
    eng.run(2)
    eng.sync()
    _check([df], oracle_mod, "cfg 2")
    eng.close()


@pytest.mark.parametrize("knobs", [
    dict(blocks=1, sr_rows=4, lr_w=32, waves=0, sub=0, claim=0),
    dict(blocks=2, sr_rows=16, lr_w=64, waves=6, sub=2, claim=3),
    dict(blocks=0, sr_rows=32, lr_w=32, waves=7, sub=3, claim=255),
    dict(blocks=3, sr_rows=5, lr_w=0, waves=8, sub=4, claim=16),
    dict(blocks=0, sr_rows=0, lr_w=0, waves=8, sub=2, claim=2),
])
def test_chain_tuned_launch_shapes(knobs, oracle_mod):
    """The autotuner's knobs (register-budget variant, persistent workgroups
    per CU, items per claim, superres rows per item, loop-restoration tile
    width, sub-batches on their own streams) leave every output byte as the oracle's, for 8-bit
    and 10/12-bit frames in one batch, on the batched path, its graph
    replays and the plain path."""
    from oxav1 import abi
    from oxav1.engine import Engine
    frames = _frames([(gen_config(**kw), s) for kw, s in SMALL])
    t = abi.Tuning()
    for s in range(5):
        for k in range(2):
            t.blocks_per_cu[s][k] = knobs["blocks"]
            t.waves_per_eu[s][k] = knobs["waves"]
            t.items_per_claim[s][k] = knobs["claim"]
    for k in range(2):
        t.sr_rows_per_item[k] = knobs["sr_rows"]
        t.lr_tile_w[k] = knobs["lr_w"]
        t.sub_batches[k] = knobs["sub"]
    eng = Engine(max_frames=len(frames), tuning=t)
    eng.prepare(frames)
    eng.run(1)
    eng.sync()
    _check(frames, oracle_mod, f"tuning {knobs}")
    for _ in range(3):  # graph replays (one graph per sub-batch), outputs cleared between
        for df in frames:
            df.clear_outputs()
        eng.run(2)
        eng.sync()
        _check(frames, oracle_mod, f"tuning {knobs}, graph replay")
    for df in frames:
        df.clear_outputs()
        eng.prepare([df])
        eng.run(0)
        eng.sync()
    _check(frames, oracle_mod, f"tuning {knobs}, plain")
    eng.close()
