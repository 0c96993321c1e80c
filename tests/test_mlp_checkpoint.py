"""MLP weights checkpoint / resume (ps_mlp_checkpoint_save / _load): host file I/O, no device.

Bars: a saved checkpoint loads back bit for bit (with and without the Adam moments); the bytes
follow the documented layout, checked field by field here with `struct` and an FNV-1a 64
computed here; every malformed file (magic, version, layer sizes, parameter count, flags,
size, checksum) is rejected with PS_E_FORMAT and a missing file with PS_E_IO; a save replaces
an existing file and leaves no temporary file behind."""
import os
import struct

import numpy as np
import pytest

import pure_sift as ps

N = ps.MLP_PARAMS


def _fnv1a(b: bytes) -> int:
    h = 1469598103934665603
    for x in b:
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _arrays(seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(N).astype(np.float32)
    p[:3] = [np.inf, -0.0, np.float32(1e-45)]    # special values survive the round trip bit for bit
    return p, rng.standard_normal(N).astype(np.float32), np.abs(rng.standard_normal(N)).astype(np.float32)


def test_round_trip_with_moments(tmp_path):
    p, m, v = _arrays(1)
    path = str(tmp_path / "mlp.ckpt")
    ps.save_mlp_checkpoint(path, p, m, v, step=123)
    got = ps.load_mlp_checkpoint(path)
    assert got["step"] == 123
    for a, b in ((got["params"], p), (got["m"], m), (got["v"], v)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert not os.path.exists(path + ".tmp")


def test_round_trip_params_only(tmp_path):
    p, _, _ = _arrays(2)
    path = str(tmp_path / "mlp.ckpt")
    ps.save_mlp_checkpoint(path, p)
    got = ps.load_mlp_checkpoint(path)
    assert got["m"] is None and got["v"] is None and got["step"] == 0
    assert np.array_equal(got["params"].view(np.uint32), p.view(np.uint32))


def test_layout(tmp_path):
    p, m, v = _arrays(3)
    path = str(tmp_path / "mlp.ckpt")
    ps.save_mlp_checkpoint(path, p, m, v, step=7)
    b = open(path, "rb").read()
    assert len(b) == 48 + 3 * 4 * N + 8
    assert b[:8] == b"PSMLPCKP"
    version, nl, d0, d1, d2, d3, n, step, flags, zero = struct.unpack_from("<10I", b, 8)
    assert (version, nl, (d0, d1, d2, d3), n, step, flags, zero) == (1, 4, (8, 64, 64, 1), N, 7, 1, 0)
    assert np.array_equal(np.frombuffer(b, "<f4", N, 48).view(np.uint32), p.view(np.uint32))
    assert np.array_equal(np.frombuffer(b, "<f4", N, 48 + 4 * N), m)
    assert np.array_equal(np.frombuffer(b, "<f4", N, 48 + 8 * N), v)
    assert struct.unpack_from("<Q", b, len(b) - 8)[0] == _fnv1a(b[:-8])


def _corrupt(path, edit):
    b = bytearray(open(path, "rb").read())
    edit(b)
    if len(b) >= 56:   # keep the checksum valid so that the header checks themselves are hit
        b[-8:] = struct.pack("<Q", _fnv1a(bytes(b[:-8])))
    open(path, "wb").write(bytes(b))


@pytest.mark.parametrize("edit,msg", [
    (lambda b: b.__setitem__(slice(0, 8), b"PSIFTCAT"), "magic"),
    (lambda b: struct.pack_into("<I", b, 8, 2), "version 2"),
    (lambda b: struct.pack_into("<I", b, 20, 32), "layer sizes"),
    (lambda b: struct.pack_into("<I", b, 32, 4800), "4800 parameters"),
    (lambda b: struct.pack_into("<I", b, 40, 3), "flags"),
    (lambda b: b.__delitem__(slice(100, 104)), "size"),
])
def test_malformed_files_rejected(tmp_path, edit, msg):
    p, m, v = _arrays(4)
    path = str(tmp_path / "mlp.ckpt")
    ps.save_mlp_checkpoint(path, p, m, v, step=1)
    _corrupt(path, edit)
    with pytest.raises(ValueError, match=msg):
        ps.load_mlp_checkpoint(path)


def test_checksum_and_missing_file(tmp_path):
    p, _, _ = _arrays(5)
    path = str(tmp_path / "mlp.ckpt")
    ps.save_mlp_checkpoint(path, p)
    b = bytearray(open(path, "rb").read())
    b[60] ^= 1    # one parameter bit, checksum left as written
    open(path, "wb").write(bytes(b))
    with pytest.raises(ValueError, match="checksum"):
        ps.load_mlp_checkpoint(path)
    open(path, "wb").write(b"PSMLP")
    with pytest.raises(ValueError, match="Truncated"):
        ps.load_mlp_checkpoint(path)
    with pytest.raises(OSError):
        ps.load_mlp_checkpoint(str(tmp_path / "nope.ckpt"))
    with pytest.raises(OSError):
        ps.save_mlp_checkpoint(str(tmp_path / "no_dir" / "x.ckpt"), p)


def test_save_replaces_existing_file(tmp_path):
    p1, _, _ = _arrays(6)
    p2, m2, v2 = _arrays(7)
    path = str(tmp_path / "mlp.ckpt")
    ps.save_mlp_checkpoint(path, p1)
    ps.save_mlp_checkpoint(path, p2, m2, v2, step=9)
    got = ps.load_mlp_checkpoint(path)
    assert got["step"] == 9 and np.array_equal(got["params"].view(np.uint32), p2.view(np.uint32))
    assert sorted(os.listdir(tmp_path)) == ["mlp.ckpt"]
