"""Checkpoint files are validated completely (magic, checksum, header
consistency, size) before a scheduler is created, so damaged files are
rejected without touching a device."""
import struct

import pytest

import cudakpz_amd as K


def fnv1a64(b):
    h = 1469598103934665603
    for x in b:
        h ^= x
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def state_bytes(kind=0, lx=4, ly=4, streams=2, words=None, tags=0, master=3076, tail=None, has_map=0,
                version=2, body_map=None):
    words = (1 << (lx + ly - 4)) if words is None else words
    head = b"KPZSTAT1" + struct.pack("<I", version)
    head += struct.pack("<iiiqiiQIIdd", kind, lx, ly, streams, 2, 2, 30721, 7, 0, 1.0, 0.0)
    head += struct.pack("<6dI", 0.5, 0.5, 0.5, 0.1, 0.1, 0.1, 0b10)
    head += struct.pack("<IQQII", master, streams, words, tags, has_map)
    body_map = (has_map == 1) if body_map is None else body_map
    body = bytes(master) + bytes(8 * streams) + bytes(4 * words) * (1 + int(body_map))
    data = head + body + (tail or b"")
    return data + struct.pack("<Q", fnv1a64(data))


def load_err(tmp_path, data):
    p = tmp_path / "s.kpzstate"
    p.write_bytes(data)
    with pytest.raises(K.KpzError) as e:
        K.load_state(str(p))
    return str(e.value)


def test_missing_file(tmp_path):
    with pytest.raises(K.KpzError, match="cannot open"):
        K.load_state(str(tmp_path / "nope"))


def test_not_a_state_file(tmp_path):
    assert "not a state file" in load_err(tmp_path, b"hello world, not a state file at all")


def test_checksum_mismatch(tmp_path):
    d = bytearray(state_bytes())
    d[100] ^= 1
    assert "checksum" in load_err(tmp_path, bytes(d))


def test_inconsistent_header(tmp_path):
    assert "inconsistent header" in load_err(tmp_path, state_bytes(kind=9))
    assert "inconsistent header" in load_err(tmp_path, state_bytes(words=17))
    assert "inconsistent header" in load_err(tmp_path, state_bytes(master=100))
    assert "inconsistent header" in load_err(tmp_path, state_bytes(has_map=2))


def test_unsupported_version(tmp_path):
    assert "unsupported version" in load_err(tmp_path, state_bytes(version=1))


def test_size_mismatch(tmp_path):
    assert "size does not match" in load_err(tmp_path, state_bytes(tail=b"\0" * 4))
    assert "size does not match" in load_err(tmp_path, state_bytes(tags=1))
    assert "size does not match" in load_err(tmp_path, state_bytes(has_map=1, body_map=False))
    assert "size does not match" in load_err(tmp_path, state_bytes(has_map=0, body_map=True))
