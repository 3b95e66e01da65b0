"""ILDG/LIME gauge files (host logic, no GPU): the library's reader and writer
against an independent implementation of the LIME record format written
here with struct (144-byte big-endian header: magic 0x456789ab, version 1,
MB/ME flags in bits 15/14, payload length, 128-byte type; payload padded to
8 bytes), round trips, 32-bit payloads, foreign records, and the errors the
reader must raise.  The reference's own reader/writer need c-lime, which is
absent, so they are not run; the writer's XML text was compared with the
reference's format string when it was written."""
import struct

import numpy as np
import pytest

import qkxtm_mpi_amd as q

MAGIC = 0x456789AB


def record(rtype, payload, mb=False, me=False):
    flags = (0x8000 if mb else 0) | (0x4000 if me else 0)
    head = struct.pack(">IHHQ", MAGIC, 1, flags, len(payload)) + rtype.encode().ljust(128, b"\0")
    return head + payload + b"\0" * ((8 - len(payload) % 8) % 8)


def parse(blob):
    out, o = [], 0
    while o < len(blob):
        magic, ver, flags, n = struct.unpack(">IHHQ", blob[o:o + 16])
        assert magic == MAGIC and ver == 1
        rtype = blob[o + 16:o + 144].rstrip(b"\0").decode()
        payload = blob[o + 144:o + 144 + n]
        out.append((rtype, bool(flags & 0x8000), bool(flags & 0x4000), payload))
        o += 144 + (n + 7) // 8 * 8
    return out


def ildg_xml(L, prec=64):
    return (f"<?xml version=\"1.0\"?>\n<ildgFormat>\n <field>su3gauge</field>\n <precision>{prec}</precision>\n"
            f" <lx>{L[0]}</lx> <ly>{L[1]}</ly> <lz>{L[2]}</lz> <lt>{L[3]}</lt>\n</ildgFormat>").encode()


def links(L, seed=5):
    rng = np.random.default_rng(seed)
    V = int(np.prod(L))
    return rng.standard_normal((V, 4, 3, 3)) + 1j * rng.standard_normal((V, 4, 3, 3))


L = (4, 2, 6, 4)


def test_reader_on_independent_file(tmp_path):
    U = links(L)
    blob = (record("ildg-format", ildg_xml(L), mb=True) + record("scidac-private-file-xml", b"<x>odd</x>") +
            record("ildg-binary-data", U.astype(">c16").tobytes(), me=True) +
            record("ildg-data-lfn", b"lfn://example/conf.0001"))
    f = tmp_path / "conf.lime"
    f.write_bytes(blob)
    assert np.array_equal(q.read_lime(f, L), U)


def test_reader_32_bit_payload(tmp_path):
    U = links(L, 7)
    blob = record("ildg-format", ildg_xml(L, 32), mb=True) + record("ildg-binary-data", U.astype(">c8").tobytes(),
                                                                      me=True)
    f = tmp_path / "conf32.lime"
    f.write_bytes(blob)
    assert np.array_equal(q.read_lime(f, L), U.astype(np.complex64).astype(np.complex128))


def test_writer_records_and_round_trip(tmp_path):
    U = links(L, 9)
    f = tmp_path / "out.lime"
    q.write_lime(f, L, U, "plaquette 0.5 seed 9")
    recs = parse(f.read_bytes())
    assert [(r[0], r[1], r[2]) for r in recs] == [("ildg-format", True, False), ("xlf-info", False, False),
                                                  ("ildg-binary-data", False, True)]
    xml = recs[0][3].decode()
    for tag, v in (("precision", 64), ("lx", L[0]), ("ly", L[1]), ("lz", L[2]), ("lt", L[3])):
        assert f"<{tag}> {v} </{tag}>" in xml
    assert "<field> su3gauge </field>" in xml and xml.startswith("<?xml version=\"1.0\" encoding=\"UTF-8\"?>")
    assert recs[1][3] == b"plaquette 0.5 seed 9"
    assert np.array_equal(np.frombuffer(recs[2][3], dtype=">c16").reshape(U.shape), U)
    assert len(f.read_bytes()) % 8 == 0
    assert np.array_equal(q.read_lime(f, L), U)


@pytest.mark.parametrize("case", ["dims", "size", "magic", "precision", "truncated", "nodata"])
def test_reader_errors(tmp_path, case):
    U = links(L, 11)
    data = U.astype(">c16").tobytes()
    fmt = record("ildg-format", ildg_xml(L), mb=True)
    blob = {
        "dims": record("ildg-format", ildg_xml((4, 2, 6, 2)), mb=True) + record("ildg-binary-data", data, me=True),
        "size": fmt + record("ildg-binary-data", data[:-16], me=True),
        "magic": b"\0" * 4 + (fmt + record("ildg-binary-data", data))[4:],
        "precision": record("ildg-format", ildg_xml(L, 16), mb=True) + record("ildg-binary-data", data, me=True),
        "truncated": (fmt + record("ildg-binary-data", data, me=True))[:-100],
        "nodata": fmt + record("xlf-info", b"no links here"),
    }[case]
    f = tmp_path / "bad.lime"
    f.write_bytes(blob)
    with pytest.raises(q.QkxError):
        q.read_lime(f, L)
