// host_driver.cpp -- runs libmdrl's file readers, writers and system generators (the
// host-only sources mdrl_io.cpp, mdrl_cells.cpp and mdrl_gen.cpp) on the commands of a
// script, for a build of those sources with AddressSanitizer and UndefinedBehaviorSanitizer.
//
// One command per line:
//   table PATH CAP                      md_rl_table_read_file
//   cells PATH CAP                      md_rl_cell_file_read
//   bounds PATH CAP                     md_rl_boundary_file_read
//   celldir DIR LX LY LZ HX HY HZ CAP   md_rl_cell_dir_read (prefix cell_ini_file_)
//   tabledir DIR OFFSET NSEG            md_rl_table_load
//   gen WHICH INDEX SEED                md_rl_make_system / md_rl_free_system
//   wcells PATH COUNT DEPTH FLAGS       md_rl_cell_file_write (row k: k + 0.25, k + 0.5, k + 0.75)
//   wbounds PATH EDGE DEPTH FLAGS       md_rl_boundary_file_write
// For each it prints "<status> <hash>": the entry point's return value and a hash of the
// 32-bit words it produced (the values read, the generated arrays; 0 when nothing), so the
// output can be compared with the regular library's on the same inputs.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "mdrl.h"

namespace {

// Polynomial hash of a sequence of 32-bit words: sum over i of (w_i + 1) P^i mod 2^64
// (numpy computes the same sum vectorised).
struct Hash {
    uint64_t h = 0, p = 1;
    void add(const void *d, size_t bytes)
    {
        const unsigned char *b = static_cast<const unsigned char *>(d);
        for (size_t i = 0; i + 4 <= bytes; i += 4) {
            uint32_t w;
            memcpy(&w, b + i, 4);
            h += ((uint64_t)w + 1) * p;
            p *= 0x100000001b3ull;
        }
    }
};

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s SCRIPT\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "r");
    if (!f) {
        perror(argv[1]);
        return 2;
    }
    char line[8192], cmd[32], path[4096];
    while (fgets(line, sizeof line, f)) {
        long rc = 0;
        Hash h;
        int cap = 0;
        if (sscanf(line, "%31s", cmd) != 1) continue;
        if (!strcmp(cmd, "table") || !strcmp(cmd, "cells") || !strcmp(cmd, "bounds")) {
            if (sscanf(line, "%*s %4095s %d", path, &cap) != 2) return 3;
            const int per = !strcmp(cmd, "cells") ? 3 : 1;
            std::vector<float> out((size_t)(cap > 0 ? cap : 1) * per);
            float *o = cap > 0 ? out.data() : nullptr;
            if (!strcmp(cmd, "table")) rc = md_rl_table_read_file(path, o, cap);
            else if (!strcmp(cmd, "cells")) rc = md_rl_cell_file_read(path, o, cap);
            else rc = md_rl_boundary_file_read(path, o, cap);
            if (rc > 0) h.add(out.data(), sizeof(float) * (size_t)rc * per);
        } else if (!strcmp(cmd, "celldir")) {
            int lo[3], hi[3];
            if (sscanf(line, "%*s %4095s %d %d %d %d %d %d %d", path, &lo[0], &lo[1], &lo[2], &hi[0], &hi[1],
                       &hi[2], &cap) != 8)
                return 3;
            std::vector<float> part((size_t)(cap > 0 ? cap : 1) * 11);
            std::vector<int32_t> cell((size_t)(cap > 0 ? cap : 1) * 3);
            rc = md_rl_cell_dir_read(path, "cell_ini_file_", lo, hi, cap > 0 ? part.data() : nullptr, cell.data(), cap);
            if (rc > 0) {
                h.add(part.data(), sizeof(float) * 11 * (size_t)rc);
                h.add(cell.data(), sizeof(int32_t) * 3 * (size_t)rc);
            }
        } else if (!strcmp(cmd, "tabledir")) {
            int off = 0, nseg = 0;
            if (sscanf(line, "%*s %4095s %d %d", path, &off, &nseg) != 3) return 3;
            md_rl_table *t = nullptr;
            rc = md_rl_table_load(&t, path, off, nseg);
            if (rc == MDRL_OK) {
                for (int k = 0; k < MDRL_NTERMS; ++k) {
                    const float *c0 = nullptr, *c1 = nullptr;
                    int nent = 0;
                    md_rl_table_get(t, k, &c0, &c1, &nent, nullptr, nullptr);
                    if (c0) {
                        h.add(c0, sizeof(float) * nent);
                        h.add(c1, sizeof(float) * nent);
                    }
                }
                md_rl_table_free(t);
            }
        } else if (!strcmp(cmd, "gen")) {
            int which = 0, index = 0;
            unsigned long long seed = 0;
            if (sscanf(line, "%*s %d %d %llu", &which, &index, &seed) != 3) return 3;
            md_rl_system s;
            rc = md_rl_make_system(&s, which, index, seed);
            if (rc == MDRL_OK) {
                h.add(s.particle, sizeof(float) * 11 * (size_t)s.n);
                h.add(s.type, sizeof(int32_t) * (size_t)s.n);
                h.add(s.charge, sizeof(float) * (size_t)s.n);
                if (s.nterms) {
                    h.add(s.term_atoms, sizeof(int32_t) * 4 * (size_t)s.nterms);
                    h.add(s.term_params, sizeof(float) * 4 * (size_t)s.nterms);
                }
                if (s.excl) h.add(s.excl, sizeof(uint32_t) * (size_t)s.n);
                md_rl_free_system(&s);
            }
        } else if (!strcmp(cmd, "wcells")) {
            int count = 0, depth = 0, flags = 0;
            if (sscanf(line, "%*s %4095s %d %d %d", path, &count, &depth, &flags) != 4) return 3;
            std::vector<float> xyz((size_t)(count > 0 ? count : 1) * 3);
            for (size_t k = 0; k < xyz.size(); ++k) xyz[k] = (float)(k / 3) + 0.25f * (float)(k % 3 + 1);
            rc = md_rl_cell_file_write(path, count > 0 ? xyz.data() : nullptr, count, depth, flags);
        } else if (!strcmp(cmd, "wbounds")) {
            int depth = 0, flags = 0;
            float edge = 0.0f;
            if (sscanf(line, "%*s %4095s %f %d %d", path, &edge, &depth, &flags) != 4) return 3;
            rc = md_rl_boundary_file_write(path, edge, depth, flags);
        } else {
            fprintf(stderr, "unknown command %s\n", cmd);
            return 3;
        }
        printf("%ld %016llx\n", rc, (unsigned long long)h.h);
    }
    fclose(f);
    return 0;
}
