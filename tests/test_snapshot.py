"""The snapshot writer's bigfile directory layout (reference: fastpm/core.py:65-89,
through bigfile.FileMPI), on the host: header lines, data files, checksums
and attribute encoding, and the reader's round trip.  bigfile is not
importable here; the checksum is pinned against coreutils' System V `sum -s`
and the attribute bytes against numpy's own encodings."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from fastpm.snapshot import Block, byte_sum, read_snapshot, write_snapshot


def test_header_and_data_file(tmp_path):
    a = (np.arange(12, dtype="<f4") * 1.5 - 3).reshape(4, 3)
    Block(str(tmp_path), "1/Position").write(a)
    d = tmp_path / "1" / "Position"
    raw = (d / "000000").read_bytes()
    assert raw == a.tobytes()
    s = sum(raw) % 2 ** 32
    r = (s & 0xFFFF) + (s >> 16)
    fold = (r & 0xFFFF) + (r >> 16)
    assert (d / "header").read_text() == "DTYPE: <f4\nNMEMB: 3\nNFILE: 1\n000000: 4 : %d : %d\n" % (s, fold)
    assert byte_sum(raw) == s


@pytest.mark.skipif(shutil.which("sum") is None, reason="coreutils sum not present")
def test_folded_checksum_is_system_v_sum(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((5000, 3)).astype("<f8")  # large byte sums: the fold matters
    Block(str(tmp_path), "blk").write(a)
    _, _, files = Block(str(tmp_path), "blk").header()
    out = subprocess.run(["sum", "-s", str(tmp_path / "blk" / "000000")], capture_output=True,
                         text=True, check=True).stdout.split()
    assert files[0][2] == int(out[0])
    assert files[0][1] > 0xFFFF


def test_attributes_hex(tmp_path):
    Block(str(tmp_path), "Header").write(None, {"Om0": 0.3, "name": "abc", "ids": np.array([1, 2])})
    lines = (tmp_path / "Header" / "attr-v2").read_text().splitlines()
    assert lines[0] == "Om0 <f8 1 " + np.float64(0.3).tobytes().hex().upper()
    assert lines[0].endswith("333333333333D33F")
    assert lines[1] == "name S1 3 616263"
    assert lines[2] == "ids <i8 2 01000000000000000200000000000000"
    assert (tmp_path / "Header" / "header").read_text() == "DTYPE: <i8\nNMEMB: 1\nNFILE: 0\n"
    got = Block(str(tmp_path), "Header").attrs()
    assert got["Om0"][0] == 0.3 and got["name"] == "abc" and list(got["ids"]) == [1, 2]


@pytest.mark.parametrize("dtype,shape,nfile", [("<f4", (1000, 3), 1), ("<f8", (1001,), 3),
                                               ("<i8", (7, 2), 4), ("<f4", (0, 3), 1)])
def test_round_trip(tmp_path, dtype, shape, nfile):
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(shape) * 100).astype(dtype)
    write_snapshot(str(tmp_path), {"Time": 0.5}, {"1/X": a}, nfile=nfile)
    head, blocks = read_snapshot(str(tmp_path), ["1/X"])
    assert head["Time"][0] == 0.5
    assert blocks["1/X"].dtype == np.dtype(dtype) and np.array_equal(blocks["1/X"], a)
    _, _, files = Block(str(tmp_path), "1/X").header()
    assert len(files) == nfile and sum(f[0] for f in files) == shape[0]


def test_corruption_detected(tmp_path):
    a = np.arange(30, dtype="<f4").reshape(10, 3)
    Block(str(tmp_path), "b").write(a)
    f = tmp_path / "b" / "000000"
    raw = bytearray(f.read_bytes())
    raw[5] ^= 0x10
    f.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        Block(str(tmp_path), "b").read()
    f.write_bytes(bytes(raw[:-4]))
    with pytest.raises(ValueError, match="bytes"):
        Block(str(tmp_path), "b").read()
