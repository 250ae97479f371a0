"""GPU: the allocation-free entry point oxav1_pf_run_frame_noalloc against
oxav1_pf_engine_run_frame, both through the C ABI: two calls in a row (a
3840x2160 10-bit frame coded 2560 wide with superres, LR and film grain,
then a smaller 10-bit frame with grain), each byte-identical to the
existing entry point's output; and a frame whose intermediates exceed the
scratch is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from conftest import gen_config

pytestmark = pytest.mark.gpu


def _existing(lib, df):
    from oxav1 import abi
    e = C.c_void_p()
    abi.check(lib.oxav1_pf_engine_create(C.byref(e), None, 1), "engine_create")
    try:
        abi.check(lib.oxav1_pf_engine_run_frame(e, C.byref(df.desc), C.byref(df.dev)), "engine_run_frame")
        abi.check(lib.oxav1_pf_engine_sync(e), "engine_sync")
        return df.result()
    finally:
        lib.oxav1_pf_engine_destroy(e)


def test_noalloc_equals_run_frame_twice():
    from oxav1 import Frame, abi
    from oxav1.engine import DeviceFrame
    lib = abi.load()
    abi.check(lib.oxav1_pf_init(), "init")
    big = Frame.generate(2, 0x0A1F0002)
    small = Frame.generate(gen_config(upscaled_width=200, frame_height=120, bit_depth=10, noise_amp=32,
                                      superres_denom=11, enable_fg=1, fg_num_y_points=8, fg_ar_lag=3,
                                      fg_chroma_points=4, fg_overlap=1), 0x0A1F0001)
    for name, f in (("4K 10-bit superres", big), ("200x120 10-bit", small)):
        print(f"noalloc: {name}", flush=True)
        df = DeviceFrame(f)
        try:
            ref_w, out_w = _existing(lib, df)
            df.clear_outputs()
            abi.check(lib.oxav1_pf_run_frame_noalloc(C.byref(df.desc), C.byref(df.dev), None),
                      "run_frame_noalloc")
            ref_g, out_g = df.result()  # the download waits for the default stream
            for p in range(f.num_planes):
                assert np.array_equal(ref_g[p], ref_w[p]), f"{name}: plane {p} pre-grain differs"
                assert np.array_equal(out_g[p], out_w[p]), f"{name}: plane {p} grain output differs"
        finally:
            df.free()


def test_noalloc_refuses_frames_larger_than_its_scratch():
    from oxav1 import Frame, abi
    from oxav1.engine import DeviceFrame
    lib = abi.load()
    f = Frame.generate(gen_config(upscaled_width=3840, frame_height=2160, bit_depth=10, subsampling_x=0,
                                  subsampling_y=0, noise_amp=32), 0x5C7A)
    df = DeviceFrame(f)
    try:
        df.clear_outputs()
        st = lib.oxav1_pf_run_frame_noalloc(C.byref(df.desc), C.byref(df.dev), None)
        assert st == abi.ERR_SCRATCH_TOO_SMALL
        ref, _ = df.result()
        assert all(not r.any() for r in ref), "a refused call wrote output"
    finally:
        df.free()
