"""Checkpoint file layout, host side (no GPU): a file written byte by byte
from the layout documented with cuco_checkpoint_header is accepted by
cuco_checkpoint_info with its size, and every malformed variant is refused."""
import ctypes as C

import numpy as np
import pytest

from cuco_amd import engine as E


def header(n, nboxes=1, timeline=1, pp=E.FP32, pm=E.FP64):
    h = E.CheckpointHeader()
    h.magic = b"CUCOCKPT"
    h.version = 1
    h.nboxes = nboxes
    h.n = n
    h.timeline = timeline
    h.pp_precision = pp
    h.pm_precision = pm
    h.params.time_begin = 0.2
    h.params.time_max = 0.3
    h.state.ti_current = 1 << 20
    h.state.pm_ti_endstep = 1 << 21
    h.state.steps = 7
    h.state.time = 0.21
    return h


def body(h, counts):
    """The arrays after the header, in the documented order and sizes."""
    n = h.n
    parts = [bytes(E.Params(24, 0, 30.0, 1.0, 1.5, 6.75, 0.1)) for _ in counts]
    parts.append(np.asarray(counts, np.int64).tobytes())
    parts.append(np.zeros((n, 4)).tobytes())                  # x, y, z, mass
    parts.append(np.zeros((n, 4)).tobytes())                  # velocity, 0
    if h.timeline:
        parts += [np.zeros(n, np.int32).tobytes(), np.zeros(n, np.int32).tobytes()]
    for p in (h.pp_precision, h.pm_precision):
        if p == E.FP64:
            parts.append(np.zeros((n, 4)).tobytes())
        elif p == E.FP32:
            parts.append(np.zeros((n, 4), np.float32).tobytes())
    return b"".join(parts)


def write(path, h, counts, cut=0, extra=b""):
    data = bytes(h) + body(h, counts)
    path.write_bytes((data[:-cut] if cut else data) + extra)
    return len(data)


def test_header_is_168_bytes():
    assert C.sizeof(E.CheckpointHeader) == 168
    assert C.sizeof(E.Timeline) == 80 and C.sizeof(E.TimelineState) == 48


@pytest.mark.parametrize("timeline,pp,pm", [(1, E.FP32, E.FP64), (0, -1, -1),
                                            (0, E.FP64, E.FP64), (1, -1, E.FP32)])
def test_documented_layout_is_accepted(tmp_path, timeline, pp, pm):
    h = header(37, nboxes=2, timeline=timeline, pp=pp, pm=pm)
    size = write(tmp_path / "c.ckpt", h, [20, 17])
    acc = {E.FP64: 32, E.FP32: 16, -1: 0}
    assert C.sizeof(E.Params) == 48
    assert size == 168 + 2 * (48 + 8) + 37 * (64 + 8 * timeline + acc[pp] + acc[pm])
    info = E.checkpoint_info(str(tmp_path / "c.ckpt"))
    assert info["bytes"] == size
    assert (info["n"], info["nboxes"], info["timeline"]) == (37, 2, timeline)
    assert (info["pp_precision"], info["pm_precision"]) == (pp, pm)
    assert info["state"]["ti_current"] == 1 << 20 and info["state"]["steps"] == 7
    assert info["params"]["time_max"] == 0.3


def test_malformed_files_are_refused(tmp_path):
    p = tmp_path / "bad.ckpt"
    cases = []
    write(p, header(5), [5], cut=1)
    cases.append("short")
    for name, (h, extra) in {
        "trailing byte": (header(5), b"\0"),
        "magic": (header(5), b""),
        "version": (header(5), b""),
        "timeline flag": (header(5), b""),
        "precision": (header(5), b""),
        "no particles": (header(5), b""),
    }.items():
        if name == "magic":
            h.magic = b"CUCOCKPX"
        elif name == "version":
            h.version = 2
        elif name == "timeline flag":
            h.timeline = 2
        elif name == "precision":
            h.pp_precision = 3
        elif name == "no particles":
            h.n = 0
        cases.append(name)
        q = tmp_path / (name.replace(" ", "_") + ".ckpt")
        q.write_bytes(bytes(h) + (body(h, [max(h.n, 1)]) if h.n > 0 else b"") + extra)
        with pytest.raises(E.CucoError) as err:
            E.checkpoint_info(str(q))
        assert err.value.status == 1, name
    with pytest.raises(E.CucoError):
        E.checkpoint_info(str(p))
    with pytest.raises(E.CucoError):
        E.checkpoint_info(str(tmp_path / "missing.ckpt"))
    assert len(cases) == 7


def test_size_of_out_of_range_header_is_negative():
    h = header(5, timeline=3)
    assert E.lib().cuco_checkpoint_bytes(C.byref(h)) == -1
    h = header(5)
    assert E.lib().cuco_checkpoint_bytes(C.byref(h)) == 168 + 56 + 5 * (64 + 8 + 16 + 32)
