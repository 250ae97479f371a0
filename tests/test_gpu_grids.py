"""The engine's side information built on the GPU from the block walker's
per-mi grids (oxav1_pf_build_grids; cdf.rs:20320-20678, loopfilter_tx_size_at
20559-20593): packed per-mi words and per-plane loop-filter transform sizes,
byte-identical to the host-side sampling, and a chain fed with them
byte-identical to the oracle."""
import ctypes as C

import numpy as np
import pytest

from conftest import gen_config

pytestmark = pytest.mark.gpu

FORMATS = [(8, 1, 1, 0), (10, 1, 0, 0), (12, 0, 0, 0), (8, 1, 1, 1)]


def _build(frame_like, mi_rows, mi_cols, np_, sx, sy):
    """Run oxav1_pf_build_grids on walker arrays (dict of uint8 arrays)."""
    from oxav1 import abi
    from oxav1.engine import DevBuffer
    lib = abi.load()
    keep = {k: DevBuffer(v.shape, np.uint8, v) for k, v in frame_like.items()}
    g = abi.WalkerGrids(**{k: b.ptr for k, b in keep.items()})
    lf_mi = DevBuffer((mi_rows, mi_cols), np.uint32)
    shapes = [(mi_rows, mi_cols)] + [((mi_rows + sy) >> sy, (mi_cols + sx) >> sx)] * 2
    tx = [DevBuffer(shapes[p], np.uint8) for p in range(3)]
    ptrs = (C.c_void_p * 3)(*[t.ptr for t in tx])
    abi.check(lib.oxav1_pf_build_grids(C.byref(g), mi_rows, mi_cols, np_, sx, sy, lf_mi.ptr, ptrs, None),
              "build_grids")
    out = lf_mi.download(), [t.download() for t in tx]
    for b in list(keep.values()) + [lf_mi] + tx:
        b.free()
    return out


def _walker_from_frame(f, rng=None):
    """Walker arrays of a generated frame; with rng, the unused one of
    TxSizes / InterTxSizes holds noise so the selection is exercised."""
    w = dict(is_inter=(f.is_intra == 0).astype(np.uint8), skip=f.skip, ref_frame=f.ref_frame, mode=f.mode,
             segment_id=f.segment_id, mi_size=f.mi_size)
    luma = f.lf_tx_size[0]
    if rng is None:
        w["tx_size"] = luma.copy()
        w["inter_tx_size"] = luma.copy()
    else:
        noise = rng.integers(0, 19, luma.shape).astype(np.uint8)
        w["tx_size"] = np.where(w["is_inter"] == 0, luma, noise).astype(np.uint8)
        w["inter_tx_size"] = np.where(w["is_inter"] == 1, luma, noise).astype(np.uint8)
    return {k: np.ascontiguousarray(v, dtype=np.uint8) for k, v in w.items()}


@pytest.mark.parametrize("fmt", FORMATS)
def test_build_grids_equals_host_sampling(fmt):
    from oxav1 import Frame
    bd, sx, sy, mono = fmt
    rng = np.random.default_rng(bd + 2 * sx + sy)
    for size, seed in (((96, 80), 1), ((130, 66), 2), ((200, 136), 3)):
        f = Frame.generate(gen_config(upscaled_width=size[0], frame_height=size[1], bit_depth=bd,
                                      subsampling_x=sx, subsampling_y=sy, mono=mono), seed)
        d = f.desc
        lf_mi, tx = _build(_walker_from_frame(f, rng), d.mi_rows, d.mi_cols, d.num_planes, sx, sy)
        assert np.array_equal(lf_mi, f.lf_mi), (fmt, size)
        for p in range(d.num_planes):
            assert np.array_equal(tx[p], f.lf_tx_size[p]), (fmt, size, p)


def test_build_grids_chroma_tx_vs_oracle(oracle_mod):
    """Random MiSize (including BLOCK_INVALID) for every subsampling: chroma
    sizes equal the oracle's get_tx_size (None -> TX_4X4)."""
    oo = oracle_mod.load()
    rng = np.random.default_rng(77)
    for sx, sy in ((1, 1), (1, 0), (0, 0)):
        mr, mc = 23, 37
        w = {k: rng.integers(0, 2, (mr, mc)).astype(np.uint8) for k in ("is_inter", "skip")}
        w.update(ref_frame=rng.integers(0, 8, (mr, mc)).astype(np.uint8),
                 mode=rng.integers(0, 25, (mr, mc)).astype(np.uint8),
                 segment_id=rng.integers(0, 8, (mr, mc)).astype(np.uint8),
                 mi_size=rng.integers(0, 23, (mr, mc)).astype(np.uint8),
                 tx_size=rng.integers(0, 19, (mr, mc)).astype(np.uint8),
                 inter_tx_size=rng.integers(0, 19, (mr, mc)).astype(np.uint8))
        lf_mi, tx = _build(w, mr, mc, 3, sx, sy)
        luma = np.where(w["is_inter"] == 1, w["inter_tx_size"], w["tx_size"])
        assert np.array_equal(tx[0], luma)
        assert np.array_equal(lf_mi & 31, w["mi_size"].astype(np.uint32))
        assert np.array_equal((lf_mi >> 6) & 1, 1 - w["is_inter"].astype(np.uint32))
        for p in (1, 2):
            for r in range(tx[p].shape[0]):
                for c in range(tx[p].shape[1]):
                    lr, lc = r << sy, c << sx
                    want = oo.oo_get_tx_size(p, int(luma[lr, lc]), int(w["mi_size"][lr, lc]), sx, sy)
                    assert tx[p][r, c] == (want if want >= 0 else 0), (sx, sy, p, r, c)


def test_chain_with_device_built_grids(oracle_mod):
    """A chain whose lf_mi and transform grids come from the device builder
    equals the oracle's chain of the frame."""
    from oxav1 import Frame
    from oxav1.engine import DevBuffer, DeviceFrame, Engine
    f = Frame.generate(gen_config(bit_depth=10, noise_amp=32, upscaled_width=160, frame_height=96,
                                  superres_denom=12, enable_fg=1, fg_num_y_points=6, fg_ar_lag=2,
                                  fg_chroma_points=3), 11)
    d = f.desc
    lf_mi, tx = _build(_walker_from_frame(f, np.random.default_rng(5)), d.mi_rows, d.mi_cols, d.num_planes,
                       d.subsampling_x, d.subsampling_y)
    df = DeviceFrame(f)
    bufs = [DevBuffer(lf_mi.shape, np.uint32, lf_mi)] + [DevBuffer(t.shape, np.uint8, t) for t in tx]
    df.dev.lf_mi = bufs[0].ptr
    for p in range(3):
        df.dev.lf_tx_size[p] = bufs[1 + p].ptr
    eng = Engine(max_frames=1)
    eng.prepare([df])
    eng.run(1)
    eng.sync()
    ref_g, out_g = df.result()
    ref_o, out_o, _ = oracle_mod.chain(f)
    for p in range(d.num_planes):
        assert np.array_equal(ref_g[p], ref_o[p]) and np.array_equal(out_g[p], out_o[p]), p
    eng.close()
    for b in bufs:
        b.free()
