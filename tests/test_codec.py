"""Compression level 1 (the difference codec), CPU side: kmst_encode_host's
stream decodes, under a restatement of the reference's decoder, to the input
records; byte counts follow the reference's size rule; LEB128 known answers."""
import numpy as np
import pytest

import kamsta_amd as k


def decode_value(buf, pos):
    """decode_value, include/datastructures/variable_length_encoding.hpp:131-160."""
    v = 0
    shift = 0
    while True:
        b = int(buf[pos])
        pos += 1
        v |= (b & 0x7F) << shift
        if not b & 0x80:
            return v, pos
        shift += 7


def decode(buf, off, first, m, chunk):
    """get_WEdges + decode_edge, include/datastructures/compression/
    difference_encoded_graph.hpp:92-104, 201-212 (src kept mod 2^32)."""
    s = np.zeros(m, dtype=np.uint32)
    d = np.zeros(m, dtype=np.uint32)
    w = np.zeros(m, dtype=np.uint32)
    for c in range((m + chunk - 1) // chunk):
        lo, hi = c * chunk, min(m, (c + 1) * chunk)
        s[lo], d[lo], w[lo] = first[3 * c], first[3 * c + 1], first[3 * c + 2]
        pos = int(off[c])
        prev = int(s[lo])
        for i in range(lo + 1, hi):
            ds, pos = decode_value(buf, pos)
            prev = (prev + ds) & 0xFFFFFFFF
            s[i] = prev
            d[i], pos = decode_value(buf, pos)
            w[i], pos = decode_value(buf, pos)
        assert pos == off[c + 1]
    return s, d, w


def size_rule(v):
    """num_bytes_for_encoding, difference_encoded_graph.hpp:236-242."""
    bits = int(v).bit_length()
    return max((bits + 6) // 7, 1)


@pytest.mark.parametrize("chunk", [1, 3, 32, 1000])
@pytest.mark.parametrize("case", ["gnm", "wide", "unsorted"])
def test_roundtrip(case, chunk):
    rng = np.random.default_rng(7)
    if case == "gnm":
        s, d, w = k.gen_gnm_host(1 << 10, 1 << 13, 254, 3)
    else:
        m = 2001
        s = np.sort(rng.integers(0, 1 << 32 - 1, m, dtype=np.uint64)).astype(np.uint32)
        d = rng.integers(0, (1 << 32) - 1, m, dtype=np.uint64).astype(np.uint32)
        w = rng.integers(0, 1 << 32, m, dtype=np.uint64).astype(np.uint32)
        if case == "unsorted":  # deltas wrap mod 2^32
            rng.shuffle(s)
    buf, off, first = k.encode_host(s, d, w, chunk)
    m = len(s)
    assert len(off) == (m + chunk - 1) // chunk + 1 and off[0] == 0 and off[-1] == len(buf)
    gs, gd, gw = decode(buf, off, first, m, chunk)
    assert np.array_equal(gs, s) and np.array_equal(gd, d) and np.array_equal(gw, w)
    expect = sum(size_rule((int(s[i]) - int(s[i - 1])) & 0xFFFFFFFF) + size_rule(d[i]) +
                 size_rule(w[i]) for i in range(1, m) if i % chunk)
    assert len(buf) == expect


def test_known_answers():
    # 300 = 0b10_0101100: LEB128 bytes 0xAC 0x02; 2^28 takes five bytes
    s = np.array([0, 5, 5], dtype=np.uint32)
    d = np.array([1, 300, 1 << 28], dtype=np.uint32)
    w = np.array([300, 1, 127], dtype=np.uint32)
    buf, off, first = k.encode_host(s, d, w, 8)
    assert list(first) == [0, 1, 300]
    assert list(buf) == [0x05, 0xAC, 0x02, 0x01,  0x00, 0x80, 0x80, 0x80, 0x80, 0x01, 0x7F]
    assert list(off) == [0, 11]


def test_empty_and_single():
    buf, off, first = k.encode_host(np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                                    np.zeros(0, np.uint32), 32)
    assert len(buf) == 0 and list(off) == [0] and len(first) == 0
    buf, off, first = k.encode_host(np.array([3], np.uint32), np.array([4], np.uint32),
                                    np.array([5], np.uint32), 32)
    assert len(buf) == 0 and list(off) == [0, 0] and list(first) == [3, 4, 5]
