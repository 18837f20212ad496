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
  tests/golden/PROVENANCE.json
        source path, SHA-256 and size of every file above.
"""
import gzip
import hashlib
import json
import os
import shutil
import sys

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

    with open(os.path.join(HERE, "PROVENANCE.json"), "w") as f:
        json.dump(prov, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote %d fixtures" % len(prov))


if __name__ == "__main__":
    main()
