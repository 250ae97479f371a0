"""The workload generator's decoder geometry (CPU only): 4:1 and A/B
partitions, transforms up to Max_Tx_Size_Rect (the 64-sample family), intra
tx_depth splits and var-tx split trees inside non-skip inter blocks, each
tiled legally (a transform tile is aligned to its own size and lies inside
one block).  The deblocking equivalence of the sequential walk and the
all-edges-at-once walk on this geometry is in test_oracle_selfcheck.py."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT, gen_config

FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "spec_tables.json")))
TXW, TXH = np.array(FIX["TX_WIDTH"]), np.array(FIX["TX_HEIGHT"])
BW, BH = np.array(FIX["NUM_4X4_BLOCKS_WIDE"]), np.array(FIX["NUM_4X4_BLOCKS_HIGH"])
FOUR_TO_ONE = {16, 17, 18, 19, 20, 21}      # BLOCK_4X16 .. BLOCK_64X16
TX64 = {4, 11, 12, 17, 18}                  # TX_64X64, 32X64, 64X32, 16X64, 64X16


def _frames():
    from oxav1 import Frame
    out = []
    for i, (w, h, sb) in enumerate(((512, 256, 1), (320, 200, 0), (264, 136, 1))):
        out.append(Frame.generate(gen_config(upscaled_width=w, frame_height=h, sb128=sb,
                                             decoder_geometry=1), 0xD0 + i))
    return out


def test_decoder_geometry_is_emitted():
    ms_all, tx_all, var = set(), set(), 0
    for f in _frames():
        R, Cc = f.desc.mi_rows, f.desc.mi_cols
        ms = f.mi_size.reshape(R, Cc)
        tx = f.lf_tx_size[0].reshape(R, Cc)
        ms_all |= set(np.unique(ms).tolist())
        tx_all |= set(np.unique(tx).tolist())
        inter = (f.is_intra.reshape(R, Cc) == 0) & (f.skip.reshape(R, Cc) == 0)
        # a var-tx block: one block (same mi_size run) with several transform sizes
        for r in range(0, R, 16):
            for c in range(0, Cc, 16):
                b = ms[r, c]
                h4, w4 = BH[b], BW[b]
                if r % h4 or c % w4 or not inter[r, c]:
                    continue
                if len(np.unique(tx[r:r + h4, c:c + w4])) > 1:
                    var += 1
    assert ms_all & FOUR_TO_ONE, "no 4:1 blocks"
    assert tx_all & TX64, "no 64-sample transforms"
    assert var > 0, "no var-tx block"


@pytest.mark.parametrize("k", range(3))
def test_transform_tiles_are_aligned_and_inside_their_block(k):
    f = _frames()[k]
    R, Cc = f.desc.mi_rows, f.desc.mi_cols
    ms = f.mi_size.reshape(R, Cc)
    tx = f.lf_tx_size[0].reshape(R, Cc)
    for r in range(R):
        for c in range(Cc):
            t, b = tx[r, c], ms[r, c]
            h4, w4 = TXH[t] // 4, TXW[t] // 4
            assert h4 <= BH[b] and w4 <= BW[b], (r, c, t, b)
            r0, c0 = r - r % h4, c - c % w4
            tile = tx[r0:r0 + h4, c0:c0 + w4]
            assert (tile == t).all(), (r, c, t)
            assert (ms[r0:r0 + h4, c0:c0 + w4] == b).all(), (r, c, t, b)
