// mdrl_io.cpp -- the reference's on-disk table formats and the error channel.
//
// Table files (all big-endian IEEE-754 single words or decimals, one entry per address):
//   .mif  Quartus memory init: DEPTH/WIDTH/ADDRESS_RADIX=DEC/DATA_RADIX=HEX header,
//         "addr : XXXXXXXX;" lines (D/CellMemoryModules/ApoA1_Mem_Ini_File/c0_14.mif)
//   .hex  Intel HEX, 4-byte data records ":04AAAA00DDDDDDDDCC", ":00000001FF" EOF
//         (O/HDL_Source/c0_14.hex)
//   .txt  one decimal per line (M/c0_14.txt, written by
//         M/LJ_no_smooth_poly_interpolation_function.m:248-258)
// .mif and .hex go through the strict memory-image reader of mdrl_cells.cpp; a table must
// assign every entry, and a damaged file is an error, never a shorter or absent table.
#include <atomic>
#include <dlfcn.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "mdrl_internal.h"

namespace mdrl {

static thread_local std::string g_err;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int hip_fail(hipError_t e, const char *what)
{
    set_error("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
    return MDRL_EHIP;
}

}  // namespace mdrl

using mdrl::set_error;

extern "C" const char *md_rl_last_error(void) { return mdrl::g_err.c_str(); }
extern "C" int md_rl_version(void) { return MDRL_VERSION; }

static float word_to_float(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// One decimal per line; blank lines are skipped, anything else is an error.
static int read_txt(FILE *f, const char *path, float *out, int cap)
{
    char line[512];
    int n = 0;
    while (fgets(line, sizeof line, f)) {
        const char *p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == 0 || *p == '\r' || *p == '\n') continue;
        char *end;
        const double v = strtod(p, &end);
        while (end != p && (*end == ' ' || *end == '\t' || *end == '\r' || *end == '\n')) ++end;
        if (end == p || *end != 0) { set_error("%s: line %d is not a number", path, n + 1); return MDRL_EIO; }
        if (n >= cap) { set_error("%s: more than %d entries", path, cap); return MDRL_EIO; }
        out[n++] = (float)v;
    }
    return n;
}

extern "C" int md_rl_table_read_file(const char *path, float *out, int cap)
{
    if (!path || !out || cap <= 0) { set_error("md_rl_table_read_file: bad argument"); return MDRL_EINVAL; }
    const char *dot = strrchr(path, '.');
    if (dot && (strcmp(dot, ".mif") == 0 || strcmp(dot, ".hex") == 0)) {
        mdrl::WordFile w;
        const int rc = mdrl::read_words(path, w);
        if (rc) return rc;
        if (w.width != 32) { set_error("%s: %d-bit words, a table has 32", path, w.width); return MDRL_EIO; }
        if (w.depth > cap) { set_error("%s: %d entries, room for %d", path, w.depth, cap); return MDRL_EIO; }
        for (int a = 0; a < w.depth; ++a) {
            if (!w.given[(size_t)a]) { set_error("%s: entry %d missing", path, a); return MDRL_EIO; }
            out[a] = word_to_float(w.parts[(size_t)a]);
        }
        return w.depth;
    }
    FILE *f = fopen(path, "r");
    if (!f) { set_error("cannot open %s", path); return MDRL_EIO; }
    const int n = read_txt(f, path, out, cap);
    fclose(f);
    return n;
}

static const int kTermPow[MDRL_NTERMS] = {14, 8, 3, 12, 6, 1};

extern "C" int md_rl_table_from_arrays(md_rl_table **out, int exp_offset, int nseg,
                                       const float *const c0[MDRL_NTERMS],
                                       const float *const c1[MDRL_NTERMS])
{
    if (!out || nseg <= 0 || nseg > 32 || exp_offset <= 0 || exp_offset + nseg > 255) {
        set_error("md_rl_table_from_arrays: bad geometry (offset %d, nseg %d)", exp_offset, nseg);
        return MDRL_EINVAL;
    }
    if (!c0 || !c1 || !c0[MDRL_TERM_14] || !c1[MDRL_TERM_14] || !c0[MDRL_TERM_8] || !c1[MDRL_TERM_8]) {
        set_error("md_rl_table_from_arrays: the r^-14 and r^-8 terms are required");
        return MDRL_EINVAL;
    }
    static std::atomic<uint64_t> next_id{1};
    md_rl_table *t = new md_rl_table;
    t->id = next_id.fetch_add(1);
    t->exp_offset = exp_offset;
    t->nseg = nseg;
    t->nent = nseg * 256;
    for (int k = 0; k < MDRL_NTERMS; ++k) {
        t->has[k] = c0[k] && c1[k];
        t->c0[k].assign(t->nent, 0.0f);
        t->c1[k].assign(t->nent, 0.0f);
        if (t->has[k]) {
            memcpy(t->c0[k].data(), c0[k], sizeof(float) * t->nent);
            memcpy(t->c1[k].data(), c1[k], sizeof(float) * t->nent);
        }
    }
    *out = t;
    return MDRL_OK;
}

extern "C" int md_rl_table_load(md_rl_table **out, const char *dir, int exp_offset, int nseg)
{
    if (!out || !dir) { set_error("md_rl_table_load: bad argument"); return MDRL_EINVAL; }
    if (nseg <= 0 || nseg > 32 || exp_offset <= 0 || exp_offset + nseg > 255) {
        set_error("md_rl_table_load: bad geometry (offset %d, nseg %d)", exp_offset, nseg);
        return MDRL_EINVAL;
    }
    const int nent = nseg * 256;
    std::vector<float> buf0[MDRL_NTERMS], buf1[MDRL_NTERMS];
    const float *p0[MDRL_NTERMS] = {}, *p1[MDRL_NTERMS] = {};
    static const char *exts[3] = {"mif", "hex", "txt"};
    for (int k = 0; k < MDRL_NTERMS; ++k) {
        for (int c = 0; c < 2; ++c) {
            std::vector<float> &b = c ? buf1[k] : buf0[k];
            b.assign(nent, 0.0f);
            int got = -1;
            for (const char *e : exts) {
                char path[4096];
                snprintf(path, sizeof path, "%s/c%d_%d.%s", dir, c, kTermPow[k], e);
                FILE *f = fopen(path, "r");
                if (!f) continue;
                fclose(f);
                got = md_rl_table_read_file(path, b.data(), nent);
                if (got < 0) return got;            // present but unreadable: not an absent term
                break;
            }
            if (got < 0) { b.clear(); break; }
            if (got != nent) {
                set_error("table c%d_%d in %s has %d entries, expected %d", c, kTermPow[k], dir, got, nent);
                return MDRL_EIO;
            }
        }
        if (!buf0[k].empty() && !buf1[k].empty()) { p0[k] = buf0[k].data(); p1[k] = buf1[k].data(); }
    }
    if (!p0[MDRL_TERM_14] || !p0[MDRL_TERM_8]) {
        set_error("no c0/c1 r^-14 and r^-8 tables in %s", dir);
        return MDRL_EIO;
    }
    return md_rl_table_from_arrays(out, exp_offset, nseg, p0, p1);
}

extern "C" int md_rl_table_get(const md_rl_table *t, int term, const float **c0, const float **c1,
                               int *nentries, int *exp_offset, int *nseg)
{
    if (!t || term < 0 || term >= MDRL_NTERMS) { set_error("md_rl_table_get: bad argument"); return MDRL_EINVAL; }
    if (c0) *c0 = t->has[term] ? t->c0[term].data() : nullptr;
    if (c1) *c1 = t->has[term] ? t->c1[term].data() : nullptr;
    if (nentries) *nentries = t->nent;
    if (exp_offset) *exp_offset = t->exp_offset;
    if (nseg) *nseg = t->nseg;
    return MDRL_OK;
}

extern "C" void md_rl_table_free(md_rl_table *t) { delete t; }

extern "C" const char *md_rl_default_table_dir(void)
{
    static std::string dir;
    if (!dir.empty()) return dir.c_str();
    const char *env = getenv("MDRL_TABLE_DIR");
    if (env && *env) { dir = env; return dir.c_str(); }
    Dl_info info;
    if (dladdr((void *)&md_rl_default_table_dir, &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        size_t slash = so.rfind('/');
        dir = (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tables/apoa1";
    } else {
        dir = "tables/apoa1";
    }
    return dir.c_str();
}

extern "C" int md_rl_table_load_default(md_rl_table **out)
{
    return md_rl_table_load(out, md_rl_default_table_dir(), 121, 14);
}
