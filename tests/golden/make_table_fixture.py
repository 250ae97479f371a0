#!/usr/bin/env python3
"""Write tests/golden/spec_tables.json: the numeric values of the constant
tables the filter path reads, parsed from the reference's Rust sources.

The tables are spec data (AV1 "Additional tables"); the fixture holds only
their values so the CPU oracle's and the HIP library's own copies can be
checked against the reference without the reference being present (it is
not present on the GPU machines).

Sources parsed (read as text, nothing is compiled or run):
  src/film_grain_tables.rs:13-132   GAUSSIAN_SEQUENCE
  src/superres.rs:144-209           UPSCALE_FILTER
  src/loop_restoration.rs:165-182   SGR_PARAMS
  src/loop_restoration.rs:134-152   WIENER_TAPS_*/SGRPROJ_XQD_*
  src/cdef.rs:66-102                CDEF_PRI_TAPS, CDEF_SEC_TAPS,
                                    CDEF_DIRECTIONS, DIV_TABLE, CDEF_UV_DIR
  src/cdf.rs:1064-1080              NUM_4X4_BLOCKS_WIDE/HIGH
  src/cdf.rs:12547-12555            TX_WIDTH, TX_HEIGHT
  src/cdf.rs:1176-1221              SUBSAMPLED_SIZE (symbolic -> ordinals)
  src/cdf.rs:12870-12893            MAX_TX_SIZE_RECT (symbolic -> ordinals)

Usage: python tests/golden/make_table_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _read(rel):
    with open(os.path.join(REF, rel), encoding="utf-8") as f:
        return f.read()


def _const_body(text, name):
    """Return the bracketed initialiser text of `pub const NAME: ... = [ ... ];`."""
    m = re.search(r"pub const %s\s*:[^=]*=\s*" % re.escape(name), text)
    if not m:
        raise KeyError(name)
    i = text.index("[", m.end())
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "[":
            depth += 1
        elif text[j] == "]":
            depth -= 1
            if depth == 0:
                return text[i:j + 1]
    raise ValueError(name)


def _strip_comments(s):
    return re.sub(r"//[^\n]*", "", s)


def _ints(body):
    return [int(v) for v in re.findall(r"-?\d+", _strip_comments(body))]


def _symbols(body):
    return re.findall(r"[A-Z][A-Z0-9_]+", _strip_comments(body))


def main():
    fgt = _read("src/film_grain_tables.rs")
    sr = _read("src/superres.rs")
    lr = _read("src/loop_restoration.rs")
    cd = _read("src/cdef.rs")
    cdf = _read("src/cdf.rs")

    block_ord = {}
    for m in re.finditer(r"pub const (BLOCK_[0-9X]+|BLOCK_INVALID): usize = (\d+);", cdf):
        block_ord[m.group(1)] = int(m.group(2))
    tx_ord = {}
    for m in re.finditer(r"pub const (TX_[0-9X]+): usize = (\d+);", cdf):
        tx_ord[m.group(1)] = int(m.group(2))

    out = {
        "GAUSSIAN_SEQUENCE": _ints(_const_body(fgt, "GAUSSIAN_SEQUENCE")),
        "UPSCALE_FILTER": _ints(_const_body(sr, "UPSCALE_FILTER")),
        "SGR_PARAMS": _ints(_const_body(lr, "SGR_PARAMS")),
        "WIENER_TAPS_MIN": _ints(_const_body(lr, "WIENER_TAPS_MIN")),
        "WIENER_TAPS_MAX": _ints(_const_body(lr, "WIENER_TAPS_MAX")),
        "SGRPROJ_XQD_MIN": _ints(_const_body(lr, "SGRPROJ_XQD_MIN")),
        "SGRPROJ_XQD_MAX": _ints(_const_body(lr, "SGRPROJ_XQD_MAX")),
        "CDEF_PRI_TAPS": _ints(_const_body(cd, "CDEF_PRI_TAPS")),
        "CDEF_SEC_TAPS": _ints(_const_body(cd, "CDEF_SEC_TAPS")),
        "CDEF_DIRECTIONS": _ints(_const_body(cd, "CDEF_DIRECTIONS")),
        "DIV_TABLE": _ints(_const_body(cd, "DIV_TABLE")),
        "CDEF_UV_DIR": _ints(_const_body(cd, "CDEF_UV_DIR")),
        "NUM_4X4_BLOCKS_WIDE": _ints(_const_body(cdf, "NUM_4X4_BLOCKS_WIDE")),
        "NUM_4X4_BLOCKS_HIGH": _ints(_const_body(cdf, "NUM_4X4_BLOCKS_HIGH")),
        "TX_WIDTH": _ints(_const_body(cdf, "TX_WIDTH")),
        "TX_HEIGHT": _ints(_const_body(cdf, "TX_HEIGHT")),
        "SUBSAMPLED_SIZE": [block_ord[s] for s in _symbols(_const_body(cdf, "SUBSAMPLED_SIZE"))],
        "MAX_TX_SIZE_RECT": [tx_ord[s] for s in _symbols(_const_body(cdf, "MAX_TX_SIZE_RECT"))],
    }
    expect_len = {
        "GAUSSIAN_SEQUENCE": 2048, "UPSCALE_FILTER": 512, "SGR_PARAMS": 64,
        "CDEF_DIRECTIONS": 32, "DIV_TABLE": 9, "CDEF_UV_DIR": 32,
        "NUM_4X4_BLOCKS_WIDE": 22, "NUM_4X4_BLOCKS_HIGH": 22, "TX_WIDTH": 19,
        "TX_HEIGHT": 19, "SUBSAMPLED_SIZE": 88, "MAX_TX_SIZE_RECT": 22,
    }
    for k, n in expect_len.items():
        assert len(out[k]) == n, (k, len(out[k]))
    dst = os.path.join(HERE, "spec_tables.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")
    print("wrote", dst)


if __name__ == "__main__":
    main()
