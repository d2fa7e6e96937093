"""Savefile formats on the CPU (no GPU call): the CRC matches zlib's CRC-32, files written
by the engine are byte-for-byte the reference's PRP v13 / LL v13 layout
(src/Saver.cpp:17-30, src/File.h:213-231), and the engine reads the reference's older
PRP v12 / LL v1 files. Headers are built here from the reference's format strings."""
import os
import random
import struct
import zlib

import pytest

import gpuowl_amd as owl


def words_bytes(words):
    return struct.pack(f"<{len(words)}I", *words)


def random_words(E, seed):
    rng = random.Random(seed)
    w = [rng.getrandbits(32) for _ in range(owl.n_words(E))]
    w[-1] &= (1 << ((E - 1) % 32 + 1)) - 1
    return w


def test_crc32_is_the_ieee_crc():
    assert owl.crc32(b"123456789") == 0xCBF43926  # the CRC-32 check value
    assert owl.crc32(b"") == 0
    rng = random.Random(5)
    for n in (1, 3, 4, 1000, 4097):
        data = bytes(rng.getrandbits(8) for _ in range(n))
        assert owl.crc32(data) == zlib.crc32(data)


def test_prp_v13_written_as_the_reference_lays_it_out(tmp_path):
    E, k, B, res64, nerr, t = 1_257_787, 120_000, 1000, 0x0123456789ABCDEF, 2, 3456.5
    check = random_words(E, 1)
    path = tmp_path / "1257787-000120000.prp"
    owl.savefile_write_prp(path, E, k, B, res64, nerr, t, check)
    header = ("OWL PRP 13 N=1*2^%u-1 k=%u block=%u res64=%016x err=%u time=%f\n" % (E, k, B, res64, nerr, t)).encode()
    body = words_bytes(check)
    assert path.read_bytes() == header + struct.pack("<I", zlib.crc32(body)) + body
    assert not os.path.exists(str(path) + ".new")
    got = owl.savefile_read_prp(path)
    assert got == {"E": E, "k": k, "block_size": B, "res64": res64, "n_errors": nerr, "elapsed": t, "check": check}


def test_ll_v13_written_as_the_reference_lays_it_out(tmp_path):
    E, k, t = 2_976_221, 50_000, 12.25
    data = random_words(E, 2)
    path = tmp_path / "2976221-000050000.ll"
    owl.savefile_write_ll(path, E, k, t, data)
    header = ("OWL LL 13 N=1*2^%u-1 k=%u time=%f\n" % (E, k, t)).encode()
    body = words_bytes(data)
    assert path.read_bytes() == header + struct.pack("<I", zlib.crc32(body)) + body
    assert owl.savefile_read_ll(path) == {"E": E, "k": k, "elapsed": t, "data": data}


def test_reads_the_previous_formats(tmp_path):
    E = 1_000_003
    check = random_words(E, 3)
    body = words_bytes(check)
    p12 = tmp_path / "v12.prp"
    p12.write_bytes(("OWL PRP 12 %u %u %u %016x %u %u\n" % (E, 2000, 500, 0xFEDCBA9876543210, 1, zlib.crc32(body))).encode()
                    + body)
    got = owl.savefile_read_prp(p12)
    assert (got["E"], got["k"], got["block_size"], got["res64"], got["n_errors"], got["elapsed"]) == \
        (E, 2000, 500, 0xFEDCBA9876543210, 1, 0.0)
    assert got["check"] == check
    p1 = tmp_path / "v1.ll"
    p1.write_bytes(("OWL LL 1 E=%u k=%u CRC=%u\n" % (E, 700, zlib.crc32(body))).encode() + body)
    got = owl.savefile_read_ll(p1)
    assert (got["E"], got["k"], got["data"]) == (E, 700, check)


def test_damaged_files_are_rejected(tmp_path):
    E = 1_000_003
    check = random_words(E, 4)
    good = tmp_path / "good.prp"
    owl.savefile_write_prp(good, E, 1000, 100, 5, 0, 0.0, check)
    raw = good.read_bytes()
    cases = {
        "flipped_word": raw[:-7] + bytes([raw[-7] ^ 0x10]) + raw[-6:],
        "short": raw[:-4],
        "bad_header": raw.replace(b"OWL PRP 13", b"OWL PRP 99", 1),
        "no_newline": raw[:40],
    }
    for name, data in cases.items():
        p = tmp_path / f"{name}.prp"
        p.write_bytes(data)
        with pytest.raises(owl.OwlError):
            owl.savefile_read_prp(p)
    with pytest.raises(owl.OwlError):
        owl.savefile_read_prp(tmp_path / "missing.prp")
    with pytest.raises(owl.OwlError):   # the word count must be the exponent's
        owl.savefile_write_prp(tmp_path / "x.prp", E, 1000, 100, 5, 0, 0.0, check[:-1])
