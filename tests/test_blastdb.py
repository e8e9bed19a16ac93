"""BLAST protein database reader (blz_blastdb_*), on the CPU.

The known-answer case is a three-sequence database written byte by byte from NCBI's published
layout of makeblastdb's files (see the field comments); format version 5 adds the volume number
and the LMDB file name.  Sequences in NCBIstdaa: M 12, K 10, V 19, A 1, C 3, W 20."""
import struct

import numpy as np
import pytest

import blz


def be32(*v):
    return b"".join(struct.pack(">I", x) for x in v)


def lstr(s: bytes):
    return be32(len(s)) + s


SEQS = [bytes([12, 10, 19]), bytes([1, 3]), bytes([20] * 5)]          # MKV, AC, WWWWW
HDRS = [b"\x30\x80\x00\x00", b"\x30\x80\xa0\x80\x00", b""]             # header bytes, stored as given


def kat_files(version: int):
    psq = b"\x00" + b"".join(s + b"\x00" for s in SEQS)               # NUL first, NUL after each
    phr = b"".join(HDRS)
    pin = (be32(version)                      # format version
           + be32(1)                          # sequence type: protein
           + (be32(7) if version == 5 else b"")  # version 5: volume number
           + lstr(b"kat")                     # title
           + (lstr(b"kat.pdb") if version == 5 else b"")  # version 5: LMDB file name
           + lstr(b"Dec 4, 2026  0:00 AM\x00\x00\x00\x00")  # date (NUL padding kept by its length)
           + be32(3)                          # number of sequences
           + struct.pack("<Q", 10)            # residues, little-endian int64
           + be32(5)                          # longest sequence
           + be32(0, 4, 9, 9)                 # header offsets into .phr
           + be32(1, 5, 8, 14))               # sequence offsets into .psq
    return pin, psq, phr


def write_db(path, pin, psq, phr):
    (path.parent / (path.name + ".pin")).write_bytes(pin)
    (path.parent / (path.name + ".psq")).write_bytes(psq)
    (path.parent / (path.name + ".phr")).write_bytes(phr)


@pytest.mark.parametrize("version", [4, 5])
def test_known_answer_database(tmp_path, version):
    write_db(tmp_path / "kat", *kat_files(version))
    with blz.BlastDB(tmp_path / "kat") as db:
        assert (db.num_seqs, db.num_residues, db.max_length, db.title) == (3, 10, 5, "kat")
        for oid in range(3):
            assert bytes(db.sequence(oid)) == SEQS[oid]
            assert db.header(oid) == HDRS[oid]
        res, offs, perm = db.gather(sort_by_length=False)
        assert bytes(res) == b"".join(SEQS) and list(offs) == [0, 3, 5, 10] and list(perm) == [0, 1, 2]
        res, offs, perm = db.gather(sort_by_length=True)            # lengths 3, 2, 5
        assert bytes(res) == SEQS[1] + SEQS[0] + SEQS[2] and list(offs) == [0, 2, 5, 10]
        assert list(perm) == [1, 0, 2]


def test_alias_file_joins_volumes(tmp_path):
    write_db(tmp_path / "a", *kat_files(4))
    write_db(tmp_path / "b", *kat_files(5))
    (tmp_path / "both.pal").write_text('TITLE two volumes\nDBLIST a "b"\n')
    with blz.BlastDB(tmp_path / "both") as db:
        assert (db.num_seqs, db.num_residues) == (6, 20)
        assert [bytes(db.sequence(i)) for i in range(6)] == SEQS + SEQS
        res, offs, perm = db.gather(sort_by_length=True)
        assert list(perm) == [1, 4, 0, 3, 2, 5]                      # stable within a length


@pytest.mark.parametrize("what", ["version", "type", "total", "offset", "longest", "lead", "missing"])
def test_malformed_files_rejected(tmp_path, what):
    pin, psq, phr = kat_files(4)
    if what == "version":
        pin = be32(3) + pin[4:]
    elif what == "type":
        pin = pin[:4] + be32(0) + pin[8:]
    elif what == "total":
        i = pin.index(struct.pack("<Q", 10))
        pin = pin[:i] + struct.pack("<Q", 11) + pin[i + 8:]
    elif what == "offset":
        pin = pin[:-4] + be32(99)
    elif what == "longest":
        i = pin.index(struct.pack("<Q", 10)) + 8
        pin = pin[:i] + be32(4) + pin[i + 4:]
    elif what == "lead":
        psq = b"\x01" + psq[1:]
    write_db(tmp_path / "bad", pin, psq, phr)
    if what == "missing":
        (tmp_path / "bad.phr").unlink()
    with pytest.raises(blz.BlzError) as e:
        blz.BlastDB(tmp_path / "bad")
    assert e.value.status == 1


def write_synthetic(path, res, offs, version=4):
    """The same layout for a larger database (headers empty)."""
    n = len(offs) - 1
    seqs = [bytes(res[int(offs[i]):int(offs[i + 1])]) for i in range(n)]
    soff = [1]
    for s in seqs:
        soff.append(soff[-1] + len(s) + 1)
    pin = (be32(version, 1) + (be32(0) if version == 5 else b"") + lstr(b"synthetic")
           + (lstr(b"x.pdb") if version == 5 else b"") + lstr(b"today") + be32(n)
           + struct.pack("<Q", int(offs[-1])) + be32(max(map(len, seqs))) + be32(*([0] * (n + 1)))
           + be32(*soff))
    write_db(path, pin, b"\x00" + b"".join(s + b"\x00" for s in seqs), b"")


def test_round_trip_synthetic(tmp_path):
    q = blz.synth_query(3, 200)
    spec = blz.synth_spec(seed=3, n_subjects=5000, sort_by_length=0)
    res, offs, _ = blz.synth_db(spec, [q])
    write_synthetic(tmp_path / "syn", res, offs)
    with blz.BlastDB(tmp_path / "syn") as db:
        r, o, perm = db.gather(sort_by_length=False)
        assert np.array_equal(r, res) and np.array_equal(o, offs)
        r, o, perm = db.gather(sort_by_length=True)
        lens = np.diff(offs.astype(np.int64))
        assert np.array_equal(np.diff(o.astype(np.int64)), lens[perm])
        assert (np.diff(lens[perm]) >= 0).all()
        for i in (0, 17, 4999):
            assert bytes(r[int(o[i]):int(o[i + 1])]) == bytes(res[int(offs[perm[i]]):int(offs[perm[i] + 1])])
