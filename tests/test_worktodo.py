"""Serving front-end formats on the CPU: worktodo lines parsed as the reference does
(src/Worktodo.cpp:25-106; the first three lines are the examples its source gives), the
PRP-3 type-1 residue x * 9^-1 mod 2^E - 1 against Python integers, and results lines in
the reference's JSON layout (src/Task.cpp:86-170)."""
import json
import random

import pytest

import gpuowl_amd as owl

AID = "FEEE9DCD59A0855711265C1165C4C693"


def test_reference_examples():
    assert owl.worktodo_parse(f"PRP={AID},1,2,124647911,-1,77,0") == \
        {"kind": "prp", "exponent": 124647911, "aid": AID, "squarings": 0}
    assert owl.worktodo_parse("DoubleCheck=E0F583710728343C61643028FBDBA0FB,70198703,75,1") == \
        {"kind": "ll", "exponent": 70198703, "aid": "E0F583710728343C61643028FBDBA0FB", "squarings": 0}
    assert owl.worktodo_parse("Cert=B2EE67DC0A514753E488794C9DD6F6BD,1,2,82997591,-1,162105") == \
        {"kind": "cert", "exponent": 82997591, "aid": "B2EE67DC0A514753E488794C9DD6F6BD", "squarings": 162105}


def test_line_variants():
    p = owl.worktodo_parse
    assert p("PRP=N/A,1,2,100000007,-1,76,0")["exponent"] == 100000007
    assert p("PRP=,1,2,100000007,-1,76,0") == {"kind": "prp", "exponent": 100000007, "aid": "", "squarings": 0}
    assert p("PRP=1,2,100000007,-1")["exponent"] == 100000007
    assert p("PRPDC=" + AID + ",1,2,80000023,-1,76,0")["kind"] == "prp"
    assert p("Test=57885161,74,1") == {"kind": "ll", "exponent": 57885161, "aid": "", "squarings": 0}
    assert p("Test=57885161\r\n")["exponent"] == 57885161
    for bad in ("", "# PRP=1,2,100000007,-1", "PRP=1,2,999,-1", "Factor=" + AID + ",100000007,70,71", "PRP=abc",
                "PRP=1,2,100000007,-1=5", "Cert=" + AID + ",1,2,82997591,-1,50", "Cert=1,2,82997591,-1,162105"):
        assert p(bad) is None, bad


def test_best_task():
    lines = ["Test=57885161", "# comment", "PRP=1,2,100000007,-1", "DoubleCheck=40000003"]
    assert owl.worktodo_best(lines) == 3                       # smallest exponent
    lines.append(f"Cert={AID},1,2,90000001,-1,20000")
    lines.append(f"Cert={AID},1,2,80000021,-1,20000")
    assert owl.worktodo_best(lines) == 5                       # a CERT first, the smallest one
    assert owl.worktodo_best(["# nothing", "garbage"]) == -1


def to_words(x, E):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(owl.n_words(E))]


def test_type1_residue_is_x_over_9():
    rng = random.Random(9)
    for E in (61, 89, 107, 521, 607, 1279, 2203, 4253):
        M = (1 << E) - 1
        inv9 = pow(9, -1, M)
        for x in [9, 0, 1, M - 1] + [rng.randrange(M) for _ in range(20)]:
            q = x * inv9 % M
            r64, r2048 = owl.prp_type1(to_words(x, E), E)
            assert r64 == q & ((1 << 64) - 1), (E, x)
            assert r2048 == "%0512x" % (q & ((1 << 2048) - 1)), (E, x)
    assert owl.prp_type1(to_words(9, 1279), 1279)[0] == 1      # a prime: 3^(M-1) = 1


def test_result_lines():
    prp = json.loads(owl.result_json("prp", 124647911, False, 0x0123456789ABCDEF, 7340032, res2048="0" * 512,
                                     n_errors=1, aid=AID, user="someone"))
    assert list(prp)[:8] == ["status", "exponent", "worktype", "res64", "res2048", "residue-type", "errors",
                             "fft-length"]
    assert (prp["status"], prp["exponent"], prp["worktype"], prp["res64"], prp["residue-type"], prp["errors"],
            prp["fft-length"]) == ("C", 124647911, "PRP-3", "0123456789abcdef", 1, {"gerbicz": 1}, 7340032)
    assert list(prp)[8:] == ["program", "user", "aid", "timestamp"]
    assert set(prp["program"]) == {"name", "version", "port", "os"} and prp["aid"] == AID
    ll = json.loads(owl.result_json("ll", 57885161, True, 0, 3145728, program_name="prpll", program_version="0.15"))
    assert list(ll) == ["status", "exponent", "worktype", "res64", "fft-length", "shift-count", "error-code",
                        "program", "timestamp"]
    assert (ll["status"], ll["worktype"], ll["res64"], ll["shift-count"], ll["error-code"]) == \
        ("P", "LL", "0000000000000000", 0, "00000000")
    assert ll["program"]["name"] == "prpll"
    assert json.loads(owl.result_json("ll", 57885161, False, 1, 3145728, user='a "quoted" name'))["user"] == \
        'a "quoted" name'


def test_result_line_with_proof_object(tmp_path):
    # a PRP run that made a proof: {"version":1, "power":p, "hashsize":64, "md5":<file md5>} after
    # fft-length (src/Task.cpp:151-163)
    import hashlib
    proof = tmp_path / "x.proof"
    proof.write_bytes(b"PRP PROOF\nVERSION=2\n" + bytes(range(256)) * 300)
    md5 = owl.proof_md5(proof)
    assert md5 == hashlib.md5(proof.read_bytes()).hexdigest()
    prp = json.loads(owl.result_json("prp", 124647911, False, 5, 7340032, res2048="0" * 512, proof_power=8,
                                     proof_md5=md5))
    assert list(prp)[:9] == ["status", "exponent", "worktype", "res64", "res2048", "residue-type", "errors",
                             "fft-length", "proof"]
    assert prp["proof"] == {"version": 1, "power": 8, "hashsize": 64, "md5": md5}
    # no proof, no object
    assert "proof" not in json.loads(owl.result_json("prp", 124647911, False, 5, 7340032, res2048="0" * 512))


def test_md5_matches_hashlib(tmp_path):
    # RFC 1321 test suite plus the 55/56/64-byte padding boundaries
    for data in (b"", b"a", b"abc", b"message digest", b"abcdefghijklmnopqrstuvwxyz", b"x" * 55, b"y" * 56,
                 b"z" * 64, b"q" * 119, bytes(range(256)) * 517):
        f = tmp_path / "m"
        f.write_bytes(data)
        assert owl.proof_md5(f) == hashlib_md5(data), len(data)
    f.write_bytes(b"abc")
    assert owl.proof_md5(f) == "900150983cd24fb0d6963f7d28e17f72"


def hashlib_md5(data):
    import hashlib
    return hashlib.md5(data).hexdigest()


def test_cert_result_line():
    # Task::writeResultCERT (src/Task.cpp:182-197): no status; worktype, exponent, sha3-hash,
    # squarings, fft-length, shift-count, error-code, then the program block and the rest
    h = "0123456789abcdef" * 4
    cert = json.loads(owl.result_json("cert", 120000023, fft_length=6815744, sha3_hash=h, squarings=12345, aid=AID))
    assert list(cert) == ["worktype", "exponent", "sha3-hash", "squarings", "fft-length", "shift-count",
                          "error-code", "program", "aid", "timestamp"]
    assert (cert["worktype"], cert["exponent"], cert["sha3-hash"], cert["squarings"], cert["fft-length"],
            cert["shift-count"], cert["error-code"]) == ("Cert", 120000023, h, 12345, 6815744, 0, "00000000")
    with pytest.raises(owl.OwlError):
        owl.result_json("cert", 120000023, fft_length=6815744, sha3_hash="abc", squarings=1)
