"""Resuming from savefiles (the reference's Saver / loadPRP / isPrimeLL reload,
src/Gpu.cpp:1169-1191, 1545-1551, 1595-1646): a run interrupted at a verified check point
and resumed from its savefile gives GMP's residues, a damaged or inconsistent most recent
file is set aside and the one before it is used, and a PRP v12 file written from the same
state resumes too."""
import struct
import sys
import zlib

import pytest

import gpuowl_amd as owl
import oracle

pytestmark = pytest.mark.gpu

E = 1_000_003     # 256:2:256, N = 262,144 (3.8 bits per word)
B = 100           # block size: 100 = 4 * 25, so the reload ladder both doubles and steps


def announce(*a):
    print("CASE", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def gmp_prp():
    return oracle.gmp_res64(E, 1000, 200)    # x_k's res64 at k = 200, 400, ..., 1000


def savefiles(d, kind="prp"):
    sub = d / (str(E) if kind == "prp" else f"ll-{E}")
    return sorted(p.name for p in sub.iterdir() if p.suffix == "." + kind)


def test_prp_resume_gives_gmp_residues(tmp_path, gmp_prp):
    announce("prp resume", E, B)
    with owl.Gpu(E) as g:
        out, res = g.prp_saved(tmp_path, 600, block_size=B, check_step=200, res_step=200)
    assert res == gmp_prp[:3] and out.res64 == gmp_prp[2] and out.nErrors == 0
    assert savefiles(tmp_path) == [f"{E}-000000200.prp", f"{E}-000000400.prp", f"{E}-000000600.prp"]
    st = owl.savefile_read_prp(tmp_path / str(E) / f"{E}-000000600.prp")
    assert (st["E"], st["k"], st["block_size"], st["res64"]) == (E, 600, B, gmp_prp[2])
    # the saved check is the product of the block-start residues before x_k: at k = 200,
    # x_0 * x_100 = 3 * x_100 (mod 2^E - 1), from GMP
    st = owl.savefile_read_prp(tmp_path / str(E) / f"{E}-000000200.prp")
    x100 = owl.words_to_int(oracle.gmp_residue(E, 100))
    assert owl.words_to_int(st["check"]) == 3 * x100 % ((1 << E) - 1)

    with owl.Gpu(E) as g:   # a new engine: the state comes from the 600 file only
        out, res = g.prp_saved(tmp_path, 1000, block_size=B, check_step=200, res_step=200)
    assert res == [0, 0, 0] + gmp_prp[3:5]
    assert out.res64 == gmp_prp[4] and out.k == 1000 and out.nErrors == 0 and out.nChecks == 2


def test_prp_damaged_latest_file_falls_back(tmp_path, gmp_prp):
    announce("prp damaged savefile", E)
    with owl.Gpu(E) as g:
        g.prp_saved(tmp_path, 600, block_size=B, check_step=200)
    latest = tmp_path / str(E) / f"{E}-000000600.prp"
    raw = bytearray(latest.read_bytes())
    raw[-5] ^= 0x01                                   # CRC mismatch
    latest.write_bytes(bytes(raw))
    with owl.Gpu(E) as g:
        out, res = g.prp_saved(tmp_path, 800, block_size=B, check_step=200, res_step=200)
    assert (tmp_path / str(E) / f"{E}-000000600.prp.bad").exists()
    assert res[2:] == gmp_prp[2:4] and out.res64 == gmp_prp[3]

    # a well-formed 800 file whose res64 does not match its check is set aside too, and the
    # 600 file (written again by the last run) is used
    st = owl.savefile_read_prp(tmp_path / str(E) / f"{E}-000000800.prp")
    owl.savefile_write_prp(tmp_path / str(E) / f"{E}-000000800.prp", E, 800, B, st["res64"] ^ 1, 0, 0.0, st["check"])
    with owl.Gpu(E) as g:
        out, _ = g.prp_saved(tmp_path, 1000, block_size=B, check_step=200)
    assert out.res64 == gmp_prp[4]


def test_prp_resume_from_v12_file(tmp_path, gmp_prp):
    announce("prp v12 resume", E)
    with owl.Gpu(E) as g:
        g.prp_saved(tmp_path, 400, block_size=B, check_step=200)
    st = owl.savefile_read_prp(tmp_path / str(E) / f"{E}-000000400.prp")
    body = struct.pack(f"<{len(st['check'])}I", *st["check"])
    d = tmp_path / "v12" / str(E)
    d.mkdir(parents=True)
    (d / f"{E}-000000400.prp").write_bytes(
        ("OWL PRP 12 %u %u %u %016x %u %u\n" % (E, 400, B, st["res64"], 0, zlib.crc32(body))).encode() + body)
    with owl.Gpu(E) as g:
        out, _ = g.prp_saved(tmp_path / "v12", 1000, block_size=B, check_step=200)
    assert out.res64 == gmp_prp[4]


def test_ll_resume_gives_gmp_residues(tmp_path):
    announce("ll resume", E)
    want = oracle.gmp_res64(E, 1000, 250, ll=True)
    with owl.Gpu(E) as g:
        out, res = g.ll_saved(tmp_path, 500, save_step=250, res_step=250)
    assert res == want[:2] and out.res64 == want[1]
    assert savefiles(tmp_path, "ll") == [f"{E}-000000250.ll", f"{E}-000000500.ll"]
    with owl.Gpu(E) as g:
        out, res = g.ll_saved(tmp_path, 1000, save_step=250, res_step=250)
    assert res == [0, 0] + want[2:] and out.res64 == want[3]
