"""Algorithmic HBM bytes per stage of one frame.

Every plane sample a stage reads is counted once and every sample it writes
is counted once (1 byte per sample at 8-bit, 2 at 10/12-bit), plus the side
information the stage reads.  Superres and loop restoration read only the
stripe-boundary rows of the deblocked frame (4 rows per 64 luma rows /
per (64 >> subsampling_y) chroma rows), which is all loop restoration uses.
"""


def _planes(desc):
    d = desc
    out = []
    for p in range(d.num_planes):
        sx = d.subsampling_x if p else 0
        sy = d.subsampling_y if p else 0
        pad = ((d.mi_rows * 4) >> sy) * ((d.mi_cols * 4) >> sx)
        outn = ((d.frame_height + sy) >> sy) * ((d.upscaled_width + sx) >> sx)
        coded = ((d.frame_height + sy) >> sy) * ((d.frame_width + sx) >> sx)
        out.append((pad, outn, coded, 4.0 / (64 >> sy)))
    return out


def stage_bytes(desc):
    """dict stage -> algorithmic bytes for one frame (0 for a skipped stage)."""
    d = desc
    B = 1 if d.bit_depth == 8 else 2
    mi = d.mi_rows * d.mi_cols
    planes = _planes(d)
    tx = sum(((d.mi_rows + (d.subsampling_y if p else 0)) >> (d.subsampling_y if p else 0))
             * ((d.mi_cols + (d.subsampling_x if p else 0)) >> (d.subsampling_x if p else 0))
             for p in range(d.num_planes))
    use_sr = d.use_superres and d.upscaled_width > d.frame_width
    uses_lr = d.lr.uses_lr
    res = {}
    res["lf"] = (sum(2 * pad * B for pad, _, _, _ in planes) + 4 * mi + tx) if d.apply_lf else 0
    res["cdef"] = (sum(2 * pad * B for pad, _, _, _ in planes) + 4 * mi
                   + ((d.mi_rows + 15) // 16) * ((d.mi_cols + 15) // 16)) if d.apply_cdef else 0
    sr = 0
    if use_sr:
        for pad, outn, coded, frac in planes:
            sr += coded * B + outn * B
            if uses_lr:
                sr += frac * (coded + outn) * B
    res["sr"] = int(sr)
    lr = 0
    for pad, outn, coded, frac in planes:
        lr += 2 * outn * B + (frac * outn * B if uses_lr else 0)
    res["lr"] = int(lr)
    res["fg"] = sum(2 * outn * B for _, outn, _, _ in planes) if d.fg.apply_grain else 0
    return res


def chain_bytes(desc):
    return sum(stage_bytes(desc).values())
