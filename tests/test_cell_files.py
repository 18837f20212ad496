"""The reference's cell-memory initialisation files (.mif / Intel HEX) read and written by
libmdrl (CPU), and the 64-cell demo's initial state from its files on the GPU.

Pins: every cell file shipped with the reference for the LJArgon sets decodes to the cell
lists the oracle regenerates from the LJArgon-20k positions (shift to the origin, floor
binning, input order within a cell), bit for bit, from .hex and .mif alike; the writer
reproduces the shipped .mif and .hex files byte for byte (including the generator's
unpadded %X count and the CR LF line ends); the ApoA1 290-word memories and the boundary
memories round-trip byte for byte; a truncated debug cell agrees across .hex, .mif and the
generator's decimal/hex .txt listing; malformed files are rejected.
"""
import gzip
import os
import tarfile

import numpy as np
import pytest

from oracle import oracle as O
from sc2019_md_amd import mdrl

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def extract(archive, dest):
    with tarfile.open(os.path.join(GOLDEN, archive), "r:gz") as t:
        for m in t.getmembers():
            assert m.isfile() and "/" not in m.name
            with open(os.path.join(dest, m.name), "wb") as f:
                f.write(t.extractfile(m).read())
    return dest


@pytest.fixture(scope="module")
def ljargon_lists():
    text = gzip.open(os.path.join(GOLDEN, "input_positions_ljargon_20000_box_58_49_49.txt.gz"), "rt").read()
    raw = np.array([[float(v) for v in line.split()[1:4]] for line in text.splitlines() if line.strip()])
    pos = O.generator_positions(raw)
    lists, dropped = O.generator_cell_lists(pos, 8.5, (7, 6, 6))
    assert dropped == 0
    return pos, lists


@pytest.fixture(scope="module")
def demo64(tmp_path_factory):
    return extract("cells_LJArgon_64.tar.gz", str(tmp_path_factory.mktemp("demo64")))


@pytest.fixture(scope="module")
def meminit(tmp_path_factory):
    return extract("cells_LJArgon_Mem_Ini.tar.gz", str(tmp_path_factory.mktemp("meminit")))


def cell_names(d, ext):
    return sorted(f for f in os.listdir(d) if f.startswith("cell_ini_file_") and f.endswith(ext)
                  and "debug" not in f)


def xyz_of(name):
    return tuple(int(t) for t in name.rsplit(".", 1)[0].split("_")[-3:])


def test_shipped_cells_decode_to_regenerated_lists(ljargon_lists, demo64, meminit):
    _, lists = ljargon_lists
    seen = 0
    for d, exts in ((demo64, (".hex",)), (meminit, (".hex", ".mif"))):
        for ext in exts:
            for name in cell_names(d, ext):
                got = mdrl.read_cell_file(os.path.join(d, name))
                want = lists[xyz_of(name)]
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name
                seen += 1
    assert seen == 64 + 14 + 14


def test_writer_reproduces_shipped_files(ljargon_lists, demo64, meminit, tmp_path):
    _, lists = ljargon_lists
    for d, exts in ((demo64, (".hex",)), (meminit, (".hex", ".mif"))):
        for ext in exts:
            for name in cell_names(d, ext):
                out = str(tmp_path / name)
                mdrl.write_cell_file(out, lists[xyz_of(name)], depth=100)
                assert open(out, "rb").read() == open(os.path.join(d, name), "rb").read(), name


def test_cell_dir_round_trip(ljargon_lists, demo64, tmp_path):
    """write_cell_dir bins and writes the whole 64-cell demo from the shifted positions
    exactly as the shipped files; read_cell_dir returns them in the generator's order."""
    pos, lists = ljargon_lists
    sel = np.all(pos < np.float32(34.0), axis=1)
    # positions beyond cell 4 fall outside the written range: 20000 - 5564 of them
    left = mdrl.write_cell_dir(str(tmp_path), pos, 8.5, (1, 1, 1), (4, 4, 4), depth=100)
    assert left == 20000 - 5564
    for name in cell_names(demo64, ".hex"):
        assert open(str(tmp_path / name), "rb").read() == open(os.path.join(demo64, name), "rb").read(), name
    part, cells = mdrl.read_cell_dir(demo64, (1, 1, 1), (4, 4, 4))
    assert part.shape == (5564, 11) and np.all(part[:, 3:] == 0)
    want = np.concatenate([lists[(x, y, z)] for x in range(1, 5) for y in range(1, 5) for z in range(1, 5)])
    assert np.array_equal(part[:, :3].view(np.uint32), want.view(np.uint32))
    assert np.array_equal(np.sort(part[:, :3].view(np.uint32), axis=0), np.sort(pos[sel].view(np.uint32), axis=0))
    assert cells[0].tolist() == [1, 1, 1] and cells[-1].tolist() == [4, 4, 4]
    assert np.all(np.diff(cells[:, 0] * 100 + cells[:, 1] * 10 + cells[:, 2]) >= 0)


def test_boundary_files(demo64, meminit, tmp_path):
    for d, name, edge, depth, crlf in ((meminit, "cell_ini_boundary_value.mif", 8.5, 8, True),
                                       (meminit, "cell_ini_boundary_value.hex", 8.5, 8, True),
                                       (demo64, "cell_ini_boundary_value_64_Cells.mif", 8.5, 5, True),
                                       (demo64, "cell_ini_boundary_value_64_Cells.hex", 8.5, 5, False)):
        b = mdrl.read_boundary_file(os.path.join(d, name))
        assert b.tolist() == [k * edge for k in range(depth)], name
        out = str(tmp_path / name)
        mdrl.write_boundary_file(out, edge, depth, crlf)
        assert open(out, "rb").read() == open(os.path.join(d, name), "rb").read(), name


def test_apoa1_memories_round_trip(tmp_path):
    d = extract("cells_ApoA1.tar.gz", str(tmp_path))
    for name in ("cell_ini_file_2_2_2.hex", "cell_ini_file_3_3_3.hex"):
        xyz = mdrl.read_cell_file(os.path.join(d, name))
        assert 100 < xyz.shape[0] < 290
        assert np.all(xyz >= 0) and np.all(xyz < 40)
        # ceil binning of 12 A cells: cell k spans (12 (k - 1), 12 k]
        k = np.array(xyz_of(name), np.float32)
        assert np.all(mdrl.bin_cells(xyz, 12.0, rule="ceil") == k)
        out = str(tmp_path / ("out_" + name))
        mdrl.write_cell_file(out, xyz, depth=290)
        assert open(out, "rb").read() == open(os.path.join(d, name), "rb").read(), name
    b = mdrl.read_boundary_file(os.path.join(d, "cell_ini_boundary_value.hex"))
    assert b.tolist() == [12.0 * k for k in range(10)]


def test_debug_cell_formats_agree(meminit, ljargon_lists):
    """The debug cell's .mif has a 23-digit word 0 (count 13 printed as "D"); .mif, .hex
    and the generator's text listing hold the same words, and they are the first 13
    particles of the full cell."""
    base = os.path.join(meminit, "cell_ini_file_debug_1_1_1")
    a = mdrl.read_cell_file(base + ".mif")
    b = mdrl.read_cell_file(base + ".hex")
    assert a.shape == (13, 3) and np.array_equal(a.view(np.uint32), b.view(np.uint32))
    first = open(base + ".mif").read().splitlines()[6]
    assert first.split(":")[1].strip() == "0000000000000000000000D;"
    hexes = [line.split(":")[1].split() for line in open(base + ".txt").read().splitlines()
             if line.startswith("HEX")]
    txt = np.array([[int(h, 16) for h in row[::-1]] for row in hexes], np.uint32)   # z, y, x -> x, y, z
    assert np.array_equal(a.view(np.uint32), txt)
    _, lists = ljargon_lists
    assert np.array_equal(a.view(np.uint32), lists[(1, 1, 1)][:13].view(np.uint32))


def test_large_cell_mif_round_trip(tmp_path):
    """Counts of 256 and more: the generator's %X word 0 grows to 25 hex digits (a value
    still equal to the count); the reader takes it right-aligned."""
    rng = np.random.default_rng(4)
    xyz = rng.uniform(0, 12, (300, 3)).astype(np.float32)
    for ext in (".mif", ".hex"):
        path = str(tmp_path / ("cell_ini_file_1_1_1" + ext))
        mdrl.write_cell_file(path, xyz, depth=400)
        assert np.array_equal(mdrl.read_cell_file(path).view(np.uint32), xyz.view(np.uint32))
    word0 = open(str(tmp_path / "cell_ini_file_1_1_1.mif")).read().splitlines()[6]
    assert word0 == "0 : 000000000000000000000012C;"


def test_malformed_files_rejected(demo64, tmp_path):
    src = open(os.path.join(demo64, "cell_ini_file_1_1_1.hex")).read()
    bad = tmp_path / "cell_ini_file_9_9_9.hex"
    bad.write_text(src.replace(":0C00010040848A63", ":0C00010040848A64", 1))    # checksum
    with pytest.raises(mdrl.MdrlError) as e:
        mdrl.read_cell_file(str(bad))
    assert e.value.code == -3
    bad.write_text(src.replace(":00000001FF", ""))                           # no EOF record
    with pytest.raises(mdrl.MdrlError):
        mdrl.read_cell_file(str(bad))
    with pytest.raises(mdrl.MdrlError):                                     # room for too few
        mdrl.read_cell_file(os.path.join(demo64, "cell_ini_file_1_1_1.hex"), cap=10)
    with pytest.raises(mdrl.MdrlError) as e:                                # boundary file is no cell file
        mdrl.read_cell_file(os.path.join(demo64, "cell_ini_boundary_value_64_Cells.mif"))
    assert e.value.code == -3
    with pytest.raises(mdrl.MdrlError):                                     # missing cell in range
        mdrl.read_cell_dir(demo64, (1, 1, 1), (5, 1, 1))
    with pytest.raises(mdrl.MdrlError):                                     # depth must exceed count
        mdrl.write_cell_file(str(tmp_path / "x.hex"), np.zeros((5, 3), np.float32), depth=5)
    mif = tmp_path / "cell_ini_file_8_8_8.mif"
    mif.write_text("DEPTH = 4;\nWIDTH = 96;\nADDRESS_RADIX = DEC;\nDATA_RADIX = HEX;\nCONTENT\nBEGIN\n"
                   "0 : 0000000000000000000000005;\nEND;\n")                  # count beyond DEPTH
    with pytest.raises(mdrl.MdrlError):
        mdrl.read_cell_file(str(mif))


@pytest.mark.gpu
def test_demo64_forces_bit_exact(demo64):
    """The reference's 64-cell demo state (5,564 LJArgon atoms from its cell memories, a
    34 A periodic box of 4^3 cells, rc 8.5 A; sigma 2.1 A and epsilon 3.984 kJ/mol of the
    LJArgon tables) on the GPU: int64 forces and pair count equal the oracle's."""
    part, _ = mdrl.read_cell_dir(demo64, (1, 1, 1), (4, 4, 4))
    print("demo64: %d atoms" % part.shape[0], flush=True)
    n = part.shape[0]
    box = np.array([34.0, 34.0, 34.0], np.float32)
    eps, sig, mass = [3.984], [2.1], [39.948]
    g = mdrl.System(n, box, 8.5, eps, sig, mass, dt=2.0)
    g.set_types(np.zeros(n, np.int32))
    g.set_particles(part)
    got = g.forces_fixed()
    st = g.stats()
    g.close()
    t = mdrl.Table.load()
    apoa1 = {k: t.get(k) for k in range(6)}
    from tests.test_oracle_golden import apoa1_tab6
    rl = O.RL(part[:, :3], box, [4, 4, 4], 8.5, np.zeros(n, np.int32), eps, sig, apoa1_tab6(apoa1))
    want, npairs = rl.forces_fix()
    assert st.pair_evals == npairs
    assert np.array_equal(got, want)


def flips(src):
    """Every single-byte corruption (XOR 0x5a) of a file's bytes."""
    for i in range(len(src)):
        b = bytearray(src)
        b[i] ^= 0x5a
        yield i, bytes(b)


def accepts(read, path, data):
    with open(path, "wb") as f:
        f.write(data)
    try:
        return read(path)
    except mdrl.MdrlError:
        return None


def test_damaged_hex_files_rejected(meminit, tmp_path):
    """Intel HEX carries a checksum per record: every single-byte corruption of a cell file
    and of a table file, and every truncation before the end-of-file record, is an error
    (not a different memory)."""
    cases = [(os.path.join(meminit, "cell_ini_file_3_2_3.hex"), mdrl.read_cell_file, 1),
             (os.path.join(GOLDEN, "c0_14.hex"), mdrl.read_table_file, 13)]
    for path, read, stride in cases:
        src = open(path, "rb").read()
        tmp = str(tmp_path / ("damaged" + path[-4:]))
        ok = [i for i, b in flips(src) if i % stride == 0 and accepts(read, tmp, b) is not None]
        assert ok == [], (path, ok[:5])
        end = src.rindex(b":00000001FF") + 11
        ok = [c for c in range(0, end, stride) if accepts(read, tmp, src[:c]) is not None]
        assert ok == [], (path, ok[:5])


def test_damaged_mif_files_rejected(meminit, tmp_path):
    """A .mif has no checksum: a corruption that turns a value digit into another hex digit
    reads as that one changed word, and every other single-byte corruption and every
    truncation before "END;" is an error."""
    path = os.path.join(meminit, "cell_ini_file_3_2_3.mif")
    src = open(path, "rb").read()
    want = mdrl.read_cell_file(path)
    tmp = str(tmp_path / "damaged.mif")
    hexdigits = b"0123456789ABCDEFabcdef"
    for i, b in flips(src):
        got = accepts(mdrl.read_cell_file, tmp, b)
        if got is None:
            continue
        line_start = src.rindex(b"\n", 0, i) + 1
        in_value = b":" in src[line_start:i] and b";" in src[i:src.index(b"\n", i)]
        assert in_value and src[i] in hexdigits and b[i] in hexdigits, (i, src[i:i + 1])
        assert got.shape == want.shape and np.count_nonzero(got.view(np.uint32) != want.view(np.uint32)) <= 1, i
    end = src.rindex(b"END;") + 4
    ok = [c for c in range(end) if accepts(mdrl.read_cell_file, tmp, src[:c]) is not None]
    assert ok == [], ok[:5]
    tab = os.path.join(os.path.dirname(mdrl.__file__), "tables", "apoa1", "c0_3.mif")
    src = open(tab, "rb").read()
    assert accepts(mdrl.read_table_file, tmp, src).shape == (3584,)
    lines = src.split(b"\n")
    assert lines[100].startswith(b"94 : ")
    semi = src.index(b";", src.index(b"\n0 : "))
    for bad in (b"\n".join(lines[:100] + lines[101:]),                        # entry 94 missing
                src.replace(b"DATA_RADIX = HEX", b"DATA_RADIX = DEC", 1),   # radix the reader lacks
                src.replace(b"BEGIN", b"BEGINS", 1),
                src[:semi] + src[semi + 1:],                                 # word 0 without ';'
                src.replace(b"\n1 : ", b"\n1x : ", 1)):
        assert accepts(mdrl.read_table_file, tmp, bad) is None


def test_unreadable_table_term_is_an_error(tmp_path):
    """A table directory whose r^-3 file is present but damaged fails to load; only a
    missing file leaves a term out."""
    src = os.path.join(os.path.dirname(mdrl.__file__), "tables", "apoa1")
    for name in os.listdir(src):
        data = open(os.path.join(src, name), "rb").read()
        if name == "c0_3.mif":
            data = data.replace(b"5 : ", b"5 ; ", 1)
        (tmp_path / name).write_bytes(data)
    with pytest.raises(mdrl.MdrlError) as e:
        mdrl.Table.load(str(tmp_path))
    assert b"c0_3.mif" in mdrl.lib().md_rl_last_error() or "c0_3" in str(e.value)
    for k in (0, 1):
        os.remove(str(tmp_path / ("c%d_3.mif" % k)))
    t = mdrl.Table.load(str(tmp_path))
    assert t.get(2)[0] is None and t.get(0)[0] is not None
