import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oxideav-av1_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


def gen_config(**kw):
    """GenConfig with small defaults; keyword overrides by field name."""
    from oxav1 import abi
    c = abi.GenConfig()
    defaults = dict(upscaled_width=96, frame_height=80, bit_depth=8, subsampling_x=1,
                    subsampling_y=1, mono=0, sb128=0, superres_denom=8, lr_unit_shift=0,
                    lr_uv_shift=0, enable_lf=1, enable_cdef=1, enable_lr=1, enable_fg=0,
                    fg_overlap=0, fg_num_y_points=0, fg_ar_lag=0, fg_chroma_points=0,
                    noise_amp=8, rect_partitions=1, exercise_segments=0, exercise_delta_lf=0,
                    cdef_skip_units=0, lr_types=7, decoder_geometry=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(c, k, v)
    return c


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gpu_calls():
    from oxav1 import abi, dropin
    lib = abi.load()
    st = lib.oxav1_pf_init()
    if st != 0:
        raise RuntimeError(f"oxav1_pf_init failed: {lib.oxav1_pf_strerror(st).decode()}")
    return dropin.gpu()
