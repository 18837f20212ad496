"""The host-only sources of libmdrl (mdrl_io.cpp, mdrl_cells.cpp, mdrl_gen.cpp: the .mif /
Intel HEX / .txt readers and writers and the synthetic-system generators) built with
AddressSanitizer and UndefinedBehaviorSanitizer (g++, -fno-sanitize-recover=all, leak
checking on) and driven over a corpus of damaged and hostile inputs: byte flips and
truncations of the reference's cell, boundary and table files, over-long lines and values,
oversized DEPTH, reversed and out-of-range address ranges, unknown HEX record types, NUL
bytes, empty and binary files, non-numeric and out-of-range .txt entries, damaged table
directories, out-of-range generator arguments and writer limits.

Bars: the sanitized process ends normally (any sanitizer report aborts it), and every
command's status and output words (hashed) equal those of the regular library on the same
input, so the sanitized build checks the code that ships.
"""
import os
import random
import shutil
import subprocess
import tarfile

import numpy as np
import pytest

from sc2019_md_amd import mdrl

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
CSRC = os.path.join(REPO, "sc2019_md_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
SOURCES = ["mdrl_io.cpp", "mdrl_cells.cpp", "mdrl_gen.cpp"]
P = np.uint64(0x100000001b3)


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.exists(os.path.join(ROCM, "include", "hip", "hip_runtime.h")):
        pytest.skip("g++ or the HIP headers are not installed")
    out = str(tmp_path_factory.mktemp("asan") / "host_driver")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
           "-I" + os.path.join(REPO, "include"), "-I" + CSRC]
    cmd += [os.path.join(CSRC, s) for s in SOURCES] + [os.path.join(HERE, "sanitize", "host_driver.cpp")]
    cmd += ["-o", out, "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"),
            "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def phash(arrays):
    """The driver's hash: sum over the 32-bit words w_i of (w_i + 1) P^i mod 2^64."""
    words = [np.ascontiguousarray(a).view(np.uint32).ravel() for a in arrays if a is not None and a.size]
    if not words:
        return 0
    w = np.concatenate(words).astype(np.uint64)
    pw = np.ones(w.size, np.uint64)
    if w.size > 1:
        pw[1:] = np.cumprod(np.full(w.size - 1, P, np.uint64))
    with np.errstate(over="ignore"):
        return int(np.sum((w + np.uint64(1)) * pw, dtype=np.uint64))


def regular(cmd):
    """The same command on the regular library: (status, hash)."""
    L = mdrl.lib()
    f = cmd.split()
    fp = lambda a: a.ctypes.data_as(mdrl.ctypes.POINTER(mdrl.ctypes.c_float))
    if f[0] in ("table", "cells", "bounds"):
        cap = int(f[2])
        per = 3 if f[0] == "cells" else 1
        out = np.zeros(max(cap, 1) * per, np.float32)
        o = fp(out) if cap > 0 else None
        fn = {"table": L.md_rl_table_read_file, "cells": L.md_rl_cell_file_read,
              "bounds": L.md_rl_boundary_file_read}[f[0]]
        rc = fn(f[1].encode(), o, cap)
        return rc, phash([out[:rc * per]]) if rc > 0 else 0
    if f[0] == "celldir":
        lo = (mdrl.ctypes.c_int * 3)(*map(int, f[2:5]))
        hi = (mdrl.ctypes.c_int * 3)(*map(int, f[5:8]))
        cap = int(f[8])
        part = np.zeros((max(cap, 1), 11), np.float32)
        cell = np.zeros((max(cap, 1), 3), np.int32)
        rc = L.md_rl_cell_dir_read(f[1].encode(), b"cell_ini_file_", lo, hi, fp(part) if cap > 0 else None,
                                   cell.ctypes.data_as(mdrl.ctypes.POINTER(mdrl.ctypes.c_int32)), cap)
        return rc, phash([part[:rc], cell[:rc]]) if rc > 0 else 0
    if f[0] == "tabledir":
        try:
            t = mdrl.Table.load(f[1], int(f[2]), int(f[3]))
        except mdrl.MdrlError as e:
            return e.code, 0
        arrs = []
        for k in range(6):
            c0, c1 = t.get(k)
            if c0 is not None:
                arrs += [np.asarray(c0, np.float32), np.asarray(c1, np.float32)]
        return 0, phash(arrs)
    if f[0] == "gen":
        try:
            s = mdrl.SyntheticSystem(int(f[1]), int(f[2]), int(f[3]))
        except mdrl.MdrlError as e:
            return e.code, 0
        arrs = [s.particle, s.type, s.charge]
        if s.nterms:
            arrs += [s.term_atoms, s.term_params]
        arrs.append(s.excl)
        return 0, phash(arrs)
    if f[0] == "wcells":
        count, depth, flags = int(f[2]), int(f[3]), int(f[4])
        xyz = (np.arange(max(count, 1) * 3) // 3 + 0.25 * (np.arange(max(count, 1) * 3) % 3 + 1)).astype(np.float32)
        return L.md_rl_cell_file_write(f[1].encode(), fp(xyz) if count > 0 else None, count, depth, flags), 0
    if f[0] == "wbounds":
        return L.md_rl_boundary_file_write(f[1].encode(), float(f[2]), int(f[3]), int(f[4])), 0
    raise ValueError(cmd)


def extract(archive, dest, names):
    with tarfile.open(os.path.join(GOLDEN, archive), "r:gz") as t:
        for n in names:
            m = t.getmember(n)
            assert m.isfile() and "/" not in m.name
            with open(os.path.join(dest, n), "wb") as f:
                f.write(t.extractfile(m).read())


MIF_HEAD = "DEPTH = {d};\nWIDTH = {w};\nADDRESS_RADIX = DEC;\nDATA_RADIX = HEX;\nCONTENT\nBEGIN\n"


def hostile_files(d):
    """(reader, file name, cap, bytes) of hand-made malformed inputs."""
    rnd = random.Random(5)
    h = MIF_HEAD
    out = [
        ("cells", "empty.mif", 64, b""),
        ("cells", "empty.hex", 64, b""),
        ("table", "empty.txt", 64, b""),
        ("cells", "binary.mif", 64, bytes(rnd.randrange(256) for _ in range(4096))),
        ("cells", "binary.hex", 64, b":" + bytes(rnd.randrange(256) for _ in range(4096))),
        ("table", "binary.txt", 64, bytes(rnd.randrange(256) for _ in range(4096))),
        ("cells", "longline.mif", 64, (h.format(d=4, w=96) + "0 : " + "0" * 5000 + "1;\nEND;\n").encode()),
        ("cells", "longline2.mif", 64, ("-- " + "x" * 3000 + "\n" + h.format(d=4, w=96) + "END;\n").encode()),
        ("cells", "wide.mif", 64, (h.format(d=4, w=96) + "0 : 1" + "0" * 30 + ";\nEND;\n").encode()),
        ("cells", "depth_huge.mif", 64, (h.format(d=999999999, w=96) + "END;\n").encode()),
        ("cells", "depth_max.mif", 64, (h.format(d=1 << 20, w=96) + "[0..1048575] : 0;\nEND;\n").encode()),
        ("cells", "depth_zero.mif", 64, (h.format(d=0, w=96) + "END;\n").encode()),
        ("cells", "width64.mif", 64, (h.format(d=4, w=64) + "END;\n").encode()),
        ("cells", "neg_addr.mif", 64, (h.format(d=4, w=96) + "-1 : 0;\nEND;\n").encode()),
        ("cells", "rev_range.mif", 64, (h.format(d=8, w=96) + "[5..2] : 0;\nEND;\n").encode()),
        ("cells", "range_past.mif", 64, (h.format(d=8, w=96) + "[0..8] : 0;\nEND;\n").encode()),
        ("cells", "addr_past.mif", 64, (h.format(d=8, w=96) + "999999999 : 0;\nEND;\n").encode()),
        ("cells", "addr_10digits.mif", 64, (h.format(d=8, w=96) + "1234567890 : 0;\nEND;\n").encode()),
        ("cells", "count_past.mif", 64, (h.format(d=4, w=96) + "0 : FFFFFFFF;\nEND;\n").encode()),
        ("cells", "radix_hex.mif", 64, h.format(d=4, w=96).replace("= DEC", "= HEX").encode() + b"END;\n"),
        ("cells", "no_content.mif", 64, b"DEPTH = 4;\nWIDTH = 96;\nBEGIN\nEND;\n"),
        ("cells", "after_end.mif", 64, (h.format(d=4, w=96) + "END;\n0 : 0;\n").encode()),
        ("cells", "nul.mif", 64, (h.format(d=4, w=96) + "0 : 0\x00;\nEND;\n").encode()),
        ("cells", "nul_head.mif", 64, b"DEPTH\x00 = 4;\n"),
        ("cells", "no_end.mif", 64, (h.format(d=4, w=96) + "0 : 1;\n").encode()),
        ("cells", "crlf_ok.mif", 64, (h.format(d=4, w=96) + "0 : 1;\n1 : 3F80000040000000" "40400000;\nEND;\n")
         .replace("\n", "\r\n").encode()),
        ("bounds", "bounds_ok.mif", 64, (h.format(d=3, w=32) + "0 : 0;\n1 : 41080000;\n2 : 41880000;\nEND;\n")
         .encode()),
        ("bounds", "bounds_cap.mif", 2, (h.format(d=3, w=32) + "[0..2] : 0;\nEND;\n").encode()),
        ("table", "table_gap.mif", 64, (h.format(d=3, w=32) + "0 : 0;\n2 : 0;\nEND;\n").encode()),
        ("cells", "hex_cnt_ff.hex", 64, b":FF0000000000\n:00000001FF\n"),
        ("cells", "hex_short.hex", 64, b":0C000000\n"),
        ("cells", "hex_type4.hex", 64, b":020000040001F9\n:00000001FF\n"),
        ("cells", "hex_type5.hex", 64, b":04000005000000CD2A\n:00000001FF\n"),
        ("cells", "hex_mixed.hex", 64, b":0C0000000000000000000000000000" b"01F3\n:040001000000000FB\n"
         b":00000001FF\n"),
        ("cells", "hex_addr_top.hex", 64, b":0CFFFF00000000000000000000000000F6\n:00000001FF\n"),
        ("cells", "hex_lower.hex", 64, b":0c0000000000000000000000000000" b"00f4\n:00000001ff\n"),
        ("cells", "hex_trailing.hex", 64, b":0C000000000000000000000000000000F4 x\n:00000001FF\n"),
        ("cells", "hex_nul.hex", 64, b":0C00000000\x0000000000000000000000F4\n:00000001FF\n"),
        ("cells", "hex_eof_only.hex", 64, b":00000001FF\n"),
        ("bounds", "hex_eof_only_b.hex", 64, b":00000001FF\n"),
        ("table", "hex_eof_only_t.hex", 64, b":00000001FF\n"),
        ("table", "txt_specials.txt", 64, b"nan\ninf\n-inf\n1e999\n0x1p3\n-0\n  7.5  \n\n"),
        ("table", "txt_word.txt", 64, b"1.0\n2.0x\n"),
        ("table", "txt_long.txt", 64, b"1" * 3000 + b"\n"),
        ("table", "txt_cap.txt", 3, b"1\n2\n3\n4\n"),
        ("table", "txt_nul.txt", 64, b"1.0\x00\n2.0\n"),
        ("table", "suffix.bin", 64, b"1\n"),
    ]
    return out


def corpus(d):
    """Script lines for the driver (and the files they read) in directory d."""
    cells = ["cell_ini_file_3_2_3.hex", "cell_ini_file_3_2_3.mif", "cell_ini_boundary_value.hex",
             "cell_ini_boundary_value.mif"]
    extract("cells_LJArgon_Mem_Ini.tar.gz", d, cells)
    shutil.copy(os.path.join(GOLDEN, "c0_14.hex"), os.path.join(d, "c0_14.hex"))
    tab_mif = os.path.join(REPO, "sc2019_md_amd", "tables", "apoa1", "c0_3.mif")
    shutil.copy(tab_mif, os.path.join(d, "c0_3.mif"))
    lines = []
    seeds = [("cells", "cell_ini_file_3_2_3.hex", 4096, 3), ("cells", "cell_ini_file_3_2_3.mif", 4096, 3),
             ("bounds", "cell_ini_boundary_value.hex", 4096, 1), ("bounds", "cell_ini_boundary_value.mif", 4096, 1),
             ("table", "c0_14.hex", 1 << 14, 97), ("table", "c0_3.mif", 1 << 14, 89)]
    n = 0
    for kind, name, cap, stride in seeds:
        src = open(os.path.join(d, name), "rb").read()
        ext = name.rsplit(".", 1)[1]
        lines.append("%s %s %d" % (kind, os.path.join(d, name), cap))
        lines.append("%s %s %d" % (kind, os.path.join(d, name), 3))          # too little room
        for i in range(0, len(src), stride):                               # one byte flipped
            for x in (0x5A, 0x01):
                b = bytearray(src)
                b[i] ^= x
                p = os.path.join(d, "m%05d.%s" % (n, ext))
                n += 1
                open(p, "wb").write(bytes(b))
                lines.append("%s %s %d" % (kind, p, cap))
        for c in range(0, len(src), max(1, stride * 3)):                   # truncated
            p = os.path.join(d, "m%05d.%s" % (n, ext))
            n += 1
            open(p, "wb").write(src[:c])
            lines.append("%s %s %d" % (kind, p, cap))
    for kind, name, cap, data in hostile_files(d):
        p = os.path.join(d, name)
        open(p, "wb").write(data)
        lines.append("%s %s %d" % (kind, p, cap))
    lines.append("cells %s 0" % os.path.join(d, "cell_ini_file_3_2_3.hex"))   # no room at all
    lines.append("table %s 0" % os.path.join(d, "c0_14.hex"))
    lines.append("cells %s 64" % os.path.join(d, "missing.hex"))

    # cell directories: whole, a missing cell, a damaged cell, zero room
    cd = os.path.join(d, "celldir")
    os.mkdir(cd)
    names = ["cell_ini_file_%d_%d_%d.hex" % (x, y, z) for x in (2, 3) for y in (2, 3) for z in (2, 3)]
    extract("cells_LJArgon_Mem_Ini.tar.gz", cd, names)
    lines += ["celldir %s 2 2 2 3 3 3 100000" % cd, "celldir %s 2 2 2 3 3 4 100000" % cd,
              "celldir %s 2 2 2 3 3 3 10" % cd, "celldir %s 2 2 2 3 3 3 0" % cd,
              "celldir %s 0 2 2 3 3 3 100000" % cd, "celldir %s 3 2 2 2 3 3 100000" % cd,
              "celldir %s 2 2 2 1000 3 3 100000" % cd]
    bd = os.path.join(d, "celldir_bad")
    shutil.copytree(cd, bd)
    with open(os.path.join(bd, "cell_ini_file_3_3_3.hex"), "r+b") as f:
        f.seek(40)
        f.write(b"Z")
    lines.append("celldir %s 2 2 2 3 3 3 100000" % bd)

    # table directories: the shipped one, a damaged term, a term with too few entries,
    # only one of c0/c1 of a term, bad geometry
    td = os.path.join(REPO, "sc2019_md_amd", "tables", "apoa1")
    lines += ["tabledir %s 121 14" % td, "tabledir %s 121 13" % td, "tabledir %s 121 0" % td,
              "tabledir %s 121 -5" % td, "tabledir %s 250 14" % td, "tabledir %s 0 14" % td,
              "tabledir %s 121 100000000" % td, "tabledir %s 121 14" % os.path.join(d, "nodir")]
    for case in ("damaged", "short", "half"):
        t = os.path.join(d, "tab_" + case)
        shutil.copytree(td, t)
        p = os.path.join(t, "c0_3.mif")
        src = open(p, "rb").read()
        if case == "damaged":
            open(p, "wb").write(src.replace(b"5 : ", b"5 ; ", 1))
        elif case == "short":
            open(p, "wb").write(src.replace(b"DEPTH = 3584", b"DEPTH = 3583", 1).replace(
                src[src.index(b"\n3583 : ") + 1:src.index(b"END;")], b""))
        else:
            os.remove(os.path.join(t, "c1_3.mif"))
        lines.append("tabledir %s 121 14" % t)

    # generators: every workload, edge replicas, bad arguments
    for which, index in [(1, 0), (2, 0), (3, 0), (4, 0), (4, 511), (4, 100000), (5, 0), (5, 127), (4, -1),
                         (5, -1), (0, 0), (6, 0), (-1, 0)]:
        lines.append("gen %d %d %d" % (which, index, 2019))

    # writers: limits of depth and count, line ends, bad edges
    wd = os.path.join(d, "w")
    os.mkdir(wd)
    for i, (count, depth, flags) in enumerate([(0, 1, 0), (5, 6, 1), (5, 5, 0), (-1, 4, 0), (100, 65536, 1),
                                               (10, 65537, 0), (3, 64, 0)]):
        for ext in ("mif", "hex"):
            lines.append("wcells %s %d %d %d" % (os.path.join(wd, "c%d.%s" % (i, ext)), count, depth, flags))
    lines.append("wcells %s 3 8 0" % os.path.join(wd, "c.txt"))
    for i, (edge, depth, flags) in enumerate([(8.5, 4, 1), (1e30, 65536, 0), (0.0, 4, 0), (-1.0, 4, 0),
                                              (float("nan"), 4, 0), (8.5, 0, 0), (8.5, 65537, 0)]):
        for ext in ("mif", "hex"):
            lines.append("wbounds %s %r %d %d" % (os.path.join(wd, "b%d.%s" % (i, ext)), edge, depth, flags))
    lines.append("wcells %s 3 8 0" % os.path.join(d, "nodir", "c.mif"))
    return lines


def test_host_sources_clean_under_asan_ubsan(driver, tmp_path):
    d = str(tmp_path / "corpus")
    os.mkdir(d)
    lines = corpus(d)
    script = tmp_path / "script.txt"
    script.write_text("\n".join(lines) + "\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=99",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=98")
    r = subprocess.run([driver, str(script)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    got = r.stdout.split("\n")[:-1]
    assert len(got) == len(lines)

    # the regular library on the same commands (writers into a parallel directory)
    wd, wd2 = os.path.join(d, "w"), os.path.join(d, "w_regular")
    os.mkdir(wd2)
    mism = []
    statuses = set()
    for line, out in zip(lines, got):
        rc, h = (int(out.split()[0]), int(out.split()[1], 16))
        statuses.add(rc if rc <= 0 else 1)
        cmd = line.replace(wd + os.sep, wd2 + os.sep) if line.startswith("w") else line
        want = regular(cmd)
        if (rc, h) != (want[0], want[1]):
            mism.append((line, (rc, h), want))
    assert mism == [], mism[:5]
    # the writers wrote the same bytes
    for name in sorted(os.listdir(wd)):
        assert open(os.path.join(wd, name), "rb").read() == open(os.path.join(wd2, name), "rb").read(), name
    assert sorted(os.listdir(wd)) == sorted(os.listdir(wd2))
    # the corpus reached successes (counts > 0, MDRL_OK) and both kinds of rejection
    # (-1 bad argument, -3 malformed or missing file)
    assert {1, 0, -1, -3} <= statuses, statuses
