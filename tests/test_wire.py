"""Byte-string format for H-matrices and instruction programs (CPU).

  * encode -> decode -> encode is the identity on bytes, and the decoded
    structure / program answers every query as the original did;
  * the committed fixtures decode and re-encode to the same bytes (format
    stability);
  * a decoded factor program, executed by the oracle, factors the matrix
    exactly as the original program's streams do;
  * corrupted, truncated and inconsistent strings are rejected.
"""
import os
import struct

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fnv1a64(b):
    h = 1469598103934665603
    for x in b:
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def sections(data):
    """[(tag, payload offset, length)] of a string (checked independently of
    the library: header, lengths, padding, FNV-1a 64 checksums)."""
    magic, ver, ns, zero = struct.unpack_from("<4sIII", data, 0)
    assert magic == b"PPHL" and ver == 1 and zero == 0
    at, out = 16, []
    for _ in range(ns):
        tag, z, n = struct.unpack_from("<4sIQ", data, at)
        assert z == 0
        at += 16
        payload = data[at:at + n]
        pad = (n + 7) // 8 * 8
        assert data[at + n:at + pad] == b"\0" * (pad - n)
        (sum_,) = struct.unpack_from("<Q", data, at + pad)
        assert sum_ == fnv1a64(payload)
        out.append((tag.decode(), at, n))
        at += pad + 8
    assert at == len(data)
    return out


def reseal(data):
    """Recomputes every section checksum (for crafted strings)."""
    data = bytearray(data)
    magic, ver, ns, zero = struct.unpack_from("<4sIII", data, 0)
    at = 16
    for _ in range(ns):
        tag, z, n = struct.unpack_from("<4sIQ", data, at)
        at += 16
        pad = (n + 7) // 8 * 8
        struct.pack_into("<Q", data, at + pad, fnv1a64(bytes(data[at:at + n])))
        at += pad + 8
    return bytes(data)


def program_view(prog):
    info = prog.info()
    keep = {k: info[k] for k in ("n_insts", "workers", "n_slots", "n_edges", "flops", "temp_elems", "scratch_elems",
                                 "total_words", "comm_waits", "dag_density")}
    n = info["n_insts"]
    return (keep, [list(map(int, s)) for s in prog.streams()], [prog.inst_words(i).tolist() for i in range(n)],
            [prog.inst_flops(i) for i in range(n)], [prog.preds(i) for i in range(n)],
            [prog.schedule(w) for w in range(info["workers"])])


@pytest.mark.parametrize("kind,kw", [("hodlr", dict(levels=3, rank=8)), ("dense", dict(leaf=64))])
def test_hmatrix_round_trip(pp, kind, kw):
    hm = pp.HMatrix(kind, 512, **kw)
    arena = np.random.default_rng(5).standard_normal(hm.storage)
    data = hm.to_bytes(arena)
    assert [s[0] for s in sections(data)] == ["TREE", "ARNA"]
    hm2, arena2 = pp.HMatrix.from_bytes(data)
    assert hm2.n == 512 and hm2.storage == hm.storage
    assert hm2.leaves() == hm.leaves()
    assert np.array_equal(arena2, arena)
    assert hm2.to_bytes(arena2) == data
    # structure only
    s = hm.to_bytes()
    assert [x[0] for x in sections(s)] == ["TREE"]
    hm3, none = pp.HMatrix.from_bytes(s)
    assert none is None and hm3.leaves() == hm.leaves()


@pytest.mark.parametrize("which", ["factor", "solve", "factor_inv"])
def test_program_round_trip(pp, which):
    hm = pp.HMatrix("hodlr", 1024, levels=3, rank=16)
    if which == "factor":
        prog = pp.Program.factor(hm, 7)
    elif which == "solve":
        prog = pp.Program.solve(hm, 4, 5)
    else:
        prog = pp.Program.factor_inv_batch([hm, hm], 9)
    data = prog.to_bytes()
    assert [s[0] for s in sections(data)] == ["PROG"]
    dec = pp.Program.from_bytes(data)
    assert program_view(dec) == program_view(prog)
    assert dec.temp_elems(0) == prog.temp_elems(0)
    assert dec.to_bytes() == data


def test_fixtures_are_stable(pp):
    import sys
    sys.path.insert(0, GOLDEN)
    try:
        import make_wire_fixtures as mk
    finally:
        sys.path.pop(0)
    hm, arena = mk.fixture_matrix()
    with open(os.path.join(GOLDEN, "wire_hodlr_n64_l2_r2.pphl"), "rb") as f:
        golden_h = f.read()
    with open(os.path.join(GOLDEN, "wire_factor_prog_w2.pphl"), "rb") as f:
        golden_p = f.read()
    sections(golden_h)
    sections(golden_p)
    assert hm.to_bytes(arena) == golden_h
    assert pp.Program.factor(hm, 2).to_bytes() == golden_p
    hm2, arena2 = pp.HMatrix.from_bytes(golden_h)
    assert hm2.leaves() == hm.leaves() and np.array_equal(arena2, arena)
    # the TREE payload spelled out: ptr_id 0, storage, node count, then the
    # root (hierarchical 64 x 64 split 32 / 32)
    (tag, at, n) = sections(golden_h)[0]
    ptr_id, zero, storage, nodes = struct.unpack_from("<iiqq", golden_h, at)
    assert (ptr_id, zero, storage) == (0, 0, hm.storage)
    assert nodes == 1 + 4 + 2 * 4  # root, its 4 children, 4 children of each of the 2 diagonal ones
    root = struct.unpack_from("<6i2i2i", golden_h, at + 24)
    assert root == (3, 64, 64, 0, 2, 2, 32, 32, 32, 32)


def test_decoded_program_factors_like_the_original(pp, orc):
    hm, kern, arena = pp.build_config(1, 1024, levels=2, rank=16)
    prog = pp.Program.factor(hm, 4)
    dec = pp.Program.from_bytes(prog.to_bytes())
    info = dec.info()
    rnd = pp.sketch(pp.SKETCH_SEED, 1 << 16)
    out = []
    for p in (prog, dec):
        fac = arena.copy()
        tmp = np.zeros(max(info["temp_elems"], 1))
        orc.run_streams(p.streams(), [fac, tmp, None], info["n_insts"], info["scratch_elems"], rnd=rnd, threads=1)
        out.append(fac)
    assert np.array_equal(out[0], out[1])
    # and the factors are an H-LU of the matrix, up to the low-rank
    # recompression of the updates (rank 16 here)
    a = pp.densify(hm, arena)
    lu = pp.densify(hm, out[1])
    l, u = pp.lu_factors(lu)
    assert np.linalg.norm(a - l @ u) / np.linalg.norm(a) < 1e-3


def test_rejects_bad_strings(pp):
    hm = pp.HMatrix("hodlr", 256, levels=2, rank=4)
    arena = np.arange(hm.storage, dtype=np.float64)
    data = hm.to_bytes(arena)
    secs = sections(data)

    def rejected(b, what):
        with pytest.raises(pp.PastelError) as e:
            pp.HMatrix.from_bytes(b)
        assert what in str(e.value), str(e.value)

    rejected(b"XXXX" + data[4:], "magic")
    rejected(data[:4] + struct.pack("<I", 2) + data[8:], "version")
    rejected(data[:-1], "")
    rejected(data[:20], "")
    rejected(data + b"\0" * 8, "trailing")
    # a flipped payload byte in each section fails its checksum
    for tag, at, n in secs:
        bad = bytearray(data)
        bad[at + n // 2] ^= 0x40
        rejected(bytes(bad), "checksum")
    # resealed strings with inconsistent contents fail the structure checks
    tat = secs[0][1]
    bad = bytearray(data)
    struct.pack_into("<q", bad, tat + 8, hm.storage + 32)  # declared arena size
    rejected(reseal(bad), "arena size")
    bad = bytearray(data)
    struct.pack_into("<i", bad, tat + 24 + 24, 100)  # root row split no longer covers 256
    rejected(reseal(bad), "partition")
    bad = bytearray(data)
    struct.pack_into("<i", bad, tat + 24, 7)  # unknown block type
    rejected(reseal(bad), "block type")
    # a string without a TREE section
    prog = pp.Program.factor(hm, 2)
    rejected(prog.to_bytes(), "TREE")
    # arena buffer demanded from a structure-only string
    with pytest.raises(pp.PastelError):
        s = hm.to_bytes()
        import ctypes as C
        L = pp.lib()
        a = np.empty(hm.storage)
        h = C.c_void_p()
        pp.check(L.pp_hmatrix_decode(s, len(s), a.ctypes.data_as(C.POINTER(C.c_double)), a.size, None, C.byref(h)))


def test_rejects_inconsistent_program(pp):
    hm = pp.HMatrix("hodlr", 256, levels=2, rank=4)
    prog = pp.Program.factor(hm, 3)
    data = prog.to_bytes()
    (tag, at, n) = sections(data)[0]
    words = [w for s in prog.streams() for w in s]
    # the streams are the last part of the payload: change one operand word of
    # the first execute entry of worker 0 (its words no longer match the
    # instruction's)
    streams = prog.streams()
    tail = sum(4 + 4 * len(s) for s in streams)
    w0 = at + n - tail + 4  # first word of worker 0's stream
    assert struct.unpack_from("<i", data, w0)[0] in (1, 2)
    off = 0
    s0 = streams[0]
    while s0[off] == 2:  # skip waits
        off += 3
    assert s0[off] == 1 and len(words) > 0
    bad = bytearray(data)
    struct.pack_into("<i", bad, w0 + 4 * (off + 3), s0[off + 3] + 1)
    with pytest.raises(pp.PastelError) as e:
        pp.Program.from_bytes(reseal(bad))
    assert "differs" in str(e.value)
    # a stream cut before its finish word
    bad = bytearray(data)
    last = at + n - 4  # finish word of the last worker
    struct.pack_into("<i", bad, last, 2)
    with pytest.raises(pp.PastelError):
        pp.Program.from_bytes(reseal(bad))


def _instruction_at(data, at, words):
    """Byte offset of the instruction list entry whose words (from the
    operation on) are `words` (the first match: instructions precede the
    streams in the payload)."""
    pat = struct.pack(f"<{len(words) - 2}i", *[int(x) for x in words[2:]])
    i = data.find(pat, at)
    assert i > 0
    return i - 8  # back to words[0] (execute, id)


def test_rejects_bad_operands_and_slots(pp):
    """A resealed PROG string whose instruction words address a slot outside
    the pointer table, carry a pitch narrower than the block or a negative
    size, or whose header has fewer slots than its members use, is rejected
    before anything is uploaded."""
    hm = pp.HMatrix("hodlr", 256, levels=2, rank=4)
    prog = pp.Program.factor(hm, 3)
    data = prog.to_bytes()
    (tag, at, n) = sections(data)[0]
    w0 = prog.inst_words(0)
    assert w0[2] == 1  # getrf [ex, id, getrf, P, off, nx, ny, ld]
    pos = _instruction_at(data, at, w0)
    for word, value, why in ((3, 3, "bad getrf operands"),      # slot 3 of a 3-slot program
                             (3, -1, "bad getrf operands"),
                             (7, int(w0[5]) - 1, "bad getrf operands"),  # pitch below the row width
                             (6, -5, "bad getrf operands")):
        bad = bytearray(data)
        struct.pack_into("<i", bad, pos + 4 * word, value)
        with pytest.raises(pp.PastelError) as e:
            pp.Program.from_bytes(reseal(bad))
        assert why in str(e.value), (word, value, str(e.value))
    # header: n_slots = 1 while member 0's temp slot is 1
    assert struct.unpack_from("<ii", data, at) == (3, 3)
    bad = bytearray(data)
    struct.pack_into("<i", bad, at + 4, 1)
    with pytest.raises(pp.PastelError) as e:
        pp.Program.from_bytes(reseal(bad))
    assert "bad program header" in str(e.value)


def test_rejects_scratch_beyond_program(pp):
    """An accum whose scratch offsets reach past the program's per-worker
    scratch, or make its temporaries overlap, is rejected."""
    hm = pp.HMatrix("hodlr", 1024, levels=3, rank=8)
    prog = pp.Program.factor(hm, 4)
    data = prog.to_bytes()
    (tag, at, n) = sections(data)[0]
    acc = next(i for i in range(prog.info()["n_insts"]) if prog.inst_words(i)[2] == 8)
    w = prog.inst_words(acc)
    pos = _instruction_at(data, at, w)
    scratch = prog.info()["scratch_elems"]
    for word, value, why in ((20, scratch, "needs more scratch"),
                             (19, 0, "accum temporaries overlap"),
                             (13, 65, "accum rank above 64")):
        bad = bytearray(data)
        struct.pack_into("<i", bad, pos + 4 * word, value)
        with pytest.raises(pp.PastelError) as e:
            pp.Program.from_bytes(reseal(bad))
        assert why in str(e.value), (word, value, str(e.value))
