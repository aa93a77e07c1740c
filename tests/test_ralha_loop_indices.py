"""CPU audit of the buffer indices of the one-launch bidiagonalization loop (ralha.hip
`k_ralha_loop`): every index expression of the kernel is restated here, evaluated for every
workgroup, wavefront, lane and step, and checked against the sizes the library itself reports for
the workspace, the LDS, the boxes and the link (`exc_ralha_loop_workspace`, `exc_ralha_loop_shape`,
`exc_ralha_box_bytes`). Covers the host-link, device-box and one-owner (sentinel) paths, with one,
two and four columns per wavefront: the four-column form once faulted on the GPU, and no index of
its addressing leaves its buffer here either. Needs the library, not a GPU."""
import ctypes

import numpy as np
import pytest

LINK_HEADER = 1280      # LoopLink: 64-byte header and done[256]; the message follows
BOX_HEADER = 256        # a box's counter; two message slots follow
SYNC_BYTES = 256 + 1024  # device words, then one barrier flag per workgroup (at most 256)
MAX_WG = 256


@pytest.fixture(scope="module")
def lib():
    from excalibur_amd import hip
    return hip.lib


def loop_ld(m):
    return (m + 3 + 31) // 32 * 32


def r256(b):
    return (b + 255) // 256 * 256


def layout(m, c):
    """Byte offsets of the loop workspace regions (exc_ralha_loop_start)."""
    g = r256((m + 3) * 8)
    gmax = -(-c // 4)
    part = r256(gmax * (m + 3) * 8)
    return {"gdev": SYNC_BYTES, "partial": SYNC_BYTES + g, "solo": SYNC_BYTES + g + part,
            "end": SYNC_BYTES + g + part + (2 * gmax + 3) * loop_ld(m) * 8}


def grid(m, c, cpw):
    return -(-c // (4 * cpw))


def step_slices(m, G, j):
    """Slice [r0, r1) of message j owned by each workgroup, and the polled bound (valid)."""
    length = m + 3 - j
    valid = (m - 1 - j) + (4 if j > 0 else 0)
    per = -(-length // G)
    w = np.arange(G)
    r0 = w * per
    r1 = np.minimum(length, r0 + per)
    return length, valid, r0, r1


@pytest.mark.parametrize("m,c,cpw", [(12, 12, 1), (60, 30, 2), (101, 34, 1), (600, 301, 2), (1000, 500, 1),
                                     (1000, 500, 2), (1000, 500, 4), (1000, 1000, 2), (1000, 1000, 4),
                                     (1100, 550, 1), (2048, 400, 1), (3000, 100, 1)])
def test_loop_indices_stay_in_their_buffers(lib, m, c, cpw):
    ws = lib.exc_ralha_loop_workspace(m, c)
    lay = layout(m, c)
    assert lay["end"] <= ws, "workspace layout larger than exc_ralha_loop_workspace"
    G = grid(m, c, cpw)
    assert G <= MAX_WG
    if cpw in (1, 2):
        nwg = ctypes.c_int()
        lds = ctypes.c_size_t()
        assert lib.exc_ralha_loop_shape(m, c, cpw, ctypes.byref(nwg), ctypes.byref(lds)) == 0
        if nwg.value:
            assert nwg.value == G
            lds_doubles = lds.value // 8
        else:
            lds_doubles = 6 * m + 19
    else:
        lds_doubles = 6 * m + 19  # the four-column form would stage the same LDS
    km = next(k for k in (4, 8, 16, 32, 64) if m <= 64 * k)
    k_per = (64 * km + 3 + 255) // 256
    L = loop_ld(m)
    box = lib.exc_ralha_box_bytes(m)
    # LDS: hbuf [0, m), the four wavefronts' shares [m, m + 4 (m + 4)), the message [.., + m + 3)
    hbuf, mp, gl = 0, m, m + 4 * (m + 4)
    assert gl + m + 3 <= lds_doubles
    # registers: lane + 64 q covers every row, and the receive's k_per loads cover every entry
    assert 64 * km >= m and 256 * k_per >= m + 3
    # columns: (w * 4 + wave) * cpw + cc, every access guarded by col < c
    cols = (np.arange(G)[:, None] * 4 + np.arange(4)[None, :]) * cpw
    assert cols.min() == 0 and cols.max() + cpw - 1 >= c - 1  # all columns owned
    for j in range(0, m - 2):
        length, valid, r0, r1 = step_slices(m, G, j)
        assert r1.max() == length and (r0[1:] >= r1[:-1]).all()  # slices cover [0, length)
        cnt = valid
        # reg_share: LDS rows of each wavefront, then the partial row (cnt entries)
        nrows = m - 1 - j
        assert mp + 3 * (m + 4) + max(m - j - 2, nrows + 3) < gl
        assert cnt <= m + 3 <= L
        # host and box paths: partial row w at w (m + 3); loads v (m + 3) + e, e < r1 <= length
        assert (G - 1) * (m + 3) + cnt <= lay["solo"] // 8 - lay["partial"] // 8
        assert (G - 1) * (m + 3) + length - 1 < (lay["solo"] - lay["partial"]) // 8
        # the message: link, gdev (m + 3 doubles) and a box slot ((j & 1) (m + 3) + e)
        assert length <= m + 3
        assert BOX_HEADER + ((j & 1) * (m + 3) + length) * 8 <= box
        assert LINK_HEADER + length * 8 <= LINK_HEADER + (m + 3) * 8
        # one owner: partial ((j & 1) G + v) L + e, message (2 G + j % 3) L + e, resets of slot
        # (j - 1) % 3 up to entry length (the last slice)
        solo_doubles = (lay["end"] - lay["solo"]) // 8
        assert ((j & 1) * G + G - 1) * L + max(cnt, length) <= solo_doubles
        assert (2 * G + j % 3) * L + length <= solo_doubles
        if j >= 1:
            p1 = np.where(r1 == length, length + 1, r1)
            assert p1.max() <= m + 3 <= L
            assert (2 * G + (j - 1) % 3) * L + p1.max() <= solo_doubles
        # reg_apply: h at row - i - 1 < m - i - 1, house column i + 1 rows i + 1 .. m - 1
        i = j
        assert m - i - 1 <= m - hbuf
        assert (m - 1) * m + (i + 1) < m * m
    # barrier flags and the link's done words: one per workgroup
    assert SYNC_BYTES - 256 >= 4 * G


def test_loop_workspace_covers_every_grid(lib):
    """The workspace is sized for the largest grid (one column per wavefront), so the one-owner
    buffers (2 G + 3) L of any grid the loop can choose fit."""
    for m in (12, 100, 1000, 2000, 3000):
        for c in (1, 7, 64, 500, 1000):
            lay = layout(m, c)
            assert lay["end"] <= lib.exc_ralha_loop_workspace(m, c)
            for cpw in (1, 2, 4):
                G = grid(m, c, cpw)
                assert lay["solo"] + (2 * G + 3) * loop_ld(m) * 8 <= lay["end"]
