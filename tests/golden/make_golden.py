#!/usr/bin/env python3
"""Copy the reference's own data files that pin the range-limited path into the
repository as fixtures (data only, no source), with SHA-256 provenance.

Run in a container where the reference tree is mounted at /root/reference:

    python tests/golden/make_golden.py

Outputs
  tests/golden/VERIFICATION_REFERENCE_OUTPUT.txt.gz
        10,000 ordered pair records (ApoA1 atoms 1..100 x 1..100, unit-coefficient
        r^-14 / r^-8 tables, cutoff 14 A), written by
        Range_Limited_demo/SourceCode/MatlabScripts/Simulation_Verification_Raw_ApoA1_LJ_no_smooth.m:113-200
  tests/golden/input_positions_ljargon_20000_box_58_49_49.txt.gz
        LJArgon-20k positions, the input of
        Range_Limited_demo/SourceCode/MatlabScripts/Simulation_Verification_LJArgon.m:90-127
  tests/golden/VERIFICATION_PARTICLE_PAIR_INPUT.sha256
        SHA-256 and line count of the 3.8 MB half-shell pair stream that script
        writes (:518-572); the head of the file is kept as
  tests/golden/VERIFICATION_PARTICLE_PAIR_INPUT.head.txt.gz
        (first 720 lines = 4 reference particles) for readable diffs.
  tests/golden/input_positions_ApoA1_first100.txt
        First 100 rows of MatlabScripts/input_positions_ApoA1.txt (the only rows the
        pair golden file uses).
  tests/golden/c0_14.hex
        Intel-HEX copy of the ApoA1 c0 r^-14 table
        (Range_Limited_demo/SourceCode/OpenCL_HDL_Lib/HDL_Source/c0_14.hex), to pin the
        Intel-HEX reader against the .mif reader.
  sc2019_md_amd/tables/apoa1/c{0,1}_{14,8,3,12,6,1}.mif
        The ApoA1-layout unit-coefficient tables
        (Range_Limited_demo/SourceCode/CellMemoryModules/ApoA1_Mem_Ini_File/), which the
        library loads at run time.
  tests/golden/cells_LJArgon_64.tar.gz
        The 64-cell demo's cell-memory initialisation files (4 x 4 x 4 cells of 8.5 A):
        CellMemoryModules/LJArgon_v_File_64_Cells/cell_ini_file_X_Y_Z.hex and
        cell_ini_boundary_value_64_Cells.{mif,hex}.
  tests/golden/cells_LJArgon_Mem_Ini.tar.gz
        CellMemoryModules/LJArgon_Mem_Ini_File/cell_ini_file_X_Y_Z.{mif,hex},
        cell_ini_boundary_value.{mif,hex}, and the .mif/.hex/.txt of one truncated debug
        cell (LJArgon_debug_ini_File_64_Cells/cell_ini_file_debug_1_1_1.*).
  tests/golden/cells_ApoA1.tar.gz
        Two ApoA1 cells (CellMemoryModules/ApoA1_Mem_Ini_File/cell_ini_file_{2_2_2,3_3_3}.hex,
        290-word memories) and cell_ini_boundary_value.hex.
  tests/golden/PROVENANCE.json
        source path, SHA-256 and size of every file above (tar members as
        <archive>:<member>).
"""
import gzip
import hashlib
import json
import os
import shutil
import sys
import tarfile

REF = "/root/reference"
D = os.path.join(REF, "Range_Limited_demo", "SourceCode")
M = os.path.join(D, "MatlabScripts")
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
TABLE_DIR = os.path.join(REPO, "sc2019_md_amd", "tables", "apoa1")


def sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def gz_copy(src, dst):
    with open(src, "rb") as fi, gzip.GzipFile(dst, "wb", mtime=0) as fo:
        shutil.copyfileobj(fi, fo)


def tar_copy(srcs, dst):
    """A reproducible .tar.gz of files (flat names, fixed owner and times)."""
    with gzip.GzipFile(dst, "wb", mtime=0) as gz:
        with tarfile.open(fileobj=gz, mode="w", format=tarfile.USTAR_FORMAT) as tar:
            for src in sorted(srcs, key=os.path.basename):
                info = tarfile.TarInfo(os.path.basename(src))
                info.size = os.path.getsize(src)
                info.mode = 0o644
                info.mtime = 0
                with open(src, "rb") as f:
                    tar.addfile(info, f)


def main():
    if not os.path.isdir(D):
        sys.exit("reference tree not mounted at %s" % REF)
    prov = {}

    def record(src, dst):
        prov[os.path.relpath(dst, REPO)] = {
            "source": os.path.relpath(src, REF),
            "sha256": sha256(src),
            "bytes": os.path.getsize(src),
        }

    src = os.path.join(M, "VERIFICATION_REFERENCE_OUTPUT.txt")
    dst = os.path.join(HERE, "VERIFICATION_REFERENCE_OUTPUT.txt.gz")
    gz_copy(src, dst)
    record(src, dst)

    src = os.path.join(M, "input_positions_ljargon_20000_box_58_49_49.txt")
    dst = os.path.join(HERE, "input_positions_ljargon_20000_box_58_49_49.txt.gz")
    gz_copy(src, dst)
    record(src, dst)

    src = os.path.join(M, "VERIFICATION_PARTICLE_PAIR_INPUT.txt")
    with open(src, "rb") as f:
        lines = f.read().split(b"\n")
    dst = os.path.join(HERE, "VERIFICATION_PARTICLE_PAIR_INPUT.sha256")
    with open(dst, "w") as f:
        f.write("%s  %d\n" % (sha256(src), len(lines)))
    record(src, dst)
    dst = os.path.join(HERE, "VERIFICATION_PARTICLE_PAIR_INPUT.head.txt.gz")
    with gzip.GzipFile(dst, "wb", mtime=0) as fo:
        fo.write(b"\n".join(lines[:721]) + b"\n")
    record(src, dst)

    src = os.path.join(M, "input_positions_ApoA1.txt")
    dst = os.path.join(HERE, "input_positions_ApoA1_first100.txt")
    with open(src, "rb") as f:
        head = f.read().split(b"\n")[:100]
    with open(dst, "wb") as f:
        f.write(b"\n".join(head) + b"\n")
    record(src, dst)

    src = os.path.join(D, "OpenCL_HDL_Lib", "HDL_Source", "c0_14.hex")
    dst = os.path.join(HERE, "c0_14.hex")
    shutil.copyfile(src, dst)
    record(src, dst)

    os.makedirs(TABLE_DIR, exist_ok=True)
    for c in (0, 1):
        for k in (14, 8, 3, 12, 6, 1):
            name = "c%d_%d.mif" % (c, k)
            src = os.path.join(D, "CellMemoryModules", "ApoA1_Mem_Ini_File", name)
            dst = os.path.join(TABLE_DIR, name)
            shutil.copyfile(src, dst)
            record(src, dst)

    cm = os.path.join(D, "CellMemoryModules")

    def cells(archive, srcs):
        dst = os.path.join(HERE, archive)
        tar_copy(srcs, dst)
        for src in srcs:
            prov["%s:%s" % (os.path.relpath(dst, REPO), os.path.basename(src))] = {
                "source": os.path.relpath(src, REF), "sha256": sha256(src),
                "bytes": os.path.getsize(src)}

    v64 = os.path.join(cm, "LJArgon_v_File_64_Cells")
    cells("cells_LJArgon_64.tar.gz",
          [os.path.join(v64, "cell_ini_file_%d_%d_%d.hex" % (x, y, z))
           for x in range(1, 5) for y in range(1, 5) for z in range(1, 5)] +
          [os.path.join(v64, "cell_ini_boundary_value_64_Cells." + e) for e in ("mif", "hex")])
    mem = os.path.join(cm, "LJArgon_Mem_Ini_File")
    dbg = os.path.join(cm, "LJArgon_debug_ini_File_64_Cells")
    cells("cells_LJArgon_Mem_Ini.tar.gz",
          sorted(os.path.join(mem, f) for f in os.listdir(mem)
                 if f.startswith("cell_ini_") and f.endswith((".mif", ".hex"))) +
          [os.path.join(dbg, "cell_ini_file_debug_1_1_1." + e) for e in ("mif", "hex", "txt")])
    apo = os.path.join(cm, "ApoA1_Mem_Ini_File")
    cells("cells_ApoA1.tar.gz",
          [os.path.join(apo, f) for f in ("cell_ini_file_2_2_2.hex", "cell_ini_file_3_3_3.hex",
                                          "cell_ini_boundary_value.hex")])

    with open(os.path.join(HERE, "PROVENANCE.json"), "w") as f:
        json.dump(prov, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote %d fixtures" % len(prov))


if __name__ == "__main__":
    main()
