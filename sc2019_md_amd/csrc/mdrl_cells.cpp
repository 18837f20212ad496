// mdrl_cells.cpp -- the reference's cell-memory initialisation files, read and written.
//
// The FPGA design initialises one position memory per cell from cell_ini_file_X_Y_Z
// (cells numbered from 1 along each axis) and its Motion_Update boundary memory from
// cell_ini_boundary_value; both come as Quartus .mif and as Intel HEX converted from it
// (D/CellMemoryModules/{ApoA1,LJArgon}_Mem_Ini_File/, LJArgon_v_File_64_Cells/).
//
// Cell file (written by M/GenInputCellList_LJArgon.m:140-205 and
// M/GenInputCellList_ApoA1.m:138-203): 96-bit words, DEPTH = the cell capacity;
//   word 0          {32'd0, 32'd0, count}   (.mif: "0 : " 16 zeros, 6 zeros, %X of count)
//   words 1..count  {z, y, x} as big-endian IEEE singles
//   the rest        zero (.mif only up to DEPTH - 1 when count + 1 < DEPTH)
// Boundary file (M/GenCellBoundary_LJArgon.m:48-60): 32-bit words, word k = k * edge.
// .hex: one data record per word (":0C" or ":04", 16-bit word address, type 00,
// data, two's-complement checksum), ":00000001FF" at the end.
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "mdrl_internal.h"

using mdrl::set_error;
using mdrl::WordFile;

namespace {

bool ends_with(const char *s, const char *suffix)
{
    const size_t a = strlen(s), b = strlen(suffix);
    return a >= b && strcmp(s + a - b, suffix) == 0;
}

int hex_digit(char c)
{
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    return -1;
}

// Two hex digits at s as a byte, or -1 (no sign, space or "0x" accepted).
int hex_byte(const char *s)
{
    const int hi = hex_digit(s[0]);
    if (hi < 0) return -1;
    const int lo = hex_digit(s[1]);
    return lo < 0 ? -1 : (hi << 4) | lo;
}

bool blank(const char *s)
{
    for (; *s; ++s)
        if (*s != ' ' && *s != '\t' && *s != '\r' && *s != '\n') return false;
    return true;
}

// A hex string of any length, right-aligned into nparts 32-bit parts (most significant
// first).  The cell files' word 0 has 23 to 25 digits, depending on the count.
bool parse_hex_value(const char *h, size_t len, int nparts, uint32_t *out)
{
    for (int k = 0; k < nparts; ++k) out[k] = 0;
    int part = nparts - 1, shift = 0;
    for (size_t i = len; i-- > 0;) {
        const char c = h[i];
        unsigned v;
        if (c >= '0' && c <= '9') v = (unsigned)(c - '0');
        else if (c >= 'A' && c <= 'F') v = (unsigned)(c - 'A' + 10);
        else if (c >= 'a' && c <= 'f') v = (unsigned)(c - 'a' + 10);
        else return false;
        if (part < 0) {
            if (v != 0) return false;      // value wider than the word
            continue;
        }
        out[part] |= (uint32_t)v << shift;
        shift += 4;
        if (shift == 32) { shift = 0; --part; }
    }
    return true;
}

// Only white space up to the end of the line, or a "--" comment after it.
bool is_blank_or_comment(const char *s)
{
    while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n') ++s;
    return *s == 0 || (s[0] == '-' && s[1] == '-');
}

const char *skip_ws(const char *s)
{
    while (*s == ' ' || *s == '\t') ++s;
    return s;
}

// Decimal digits at *s into v (no sign; at most 9 digits), advancing *s.
bool read_dec(const char **s, long *v)
{
    const char *p = *s;
    long x = 0;
    int nd = 0;
    for (; *p >= '0' && *p <= '9'; ++p, ++nd) {
        if (nd == 9) return false;
        x = 10 * x + (*p - '0');
    }
    if (nd == 0) return false;
    *s = p;
    *v = x;
    return true;
}

// The identifier at *s (letters and '_'), advancing *s.
std::string read_word(const char **s)
{
    const char *p = *s;
    while ((*p >= 'A' && *p <= 'Z') || (*p >= 'a' && *p <= 'z') || *p == '_') ++p;
    std::string w(*s, (size_t)(p - *s));
    for (char &c : w) c = (char)toupper((unsigned char)c);
    *s = p;
    return w;
}

// Quartus memory initialisation file, the form the generators write: header lines
// "KEY = value;" (DEPTH, WIDTH, ADDRESS_RADIX = DEC, DATA_RADIX = HEX), "CONTENT", "BEGIN",
// then "addr : value;" or "[lo..hi] : value;" lines and "END;".  "--" starts a comment.
// Anything else is an error, so a damaged file is not read as a different memory.
int read_mif_words(FILE *f, const char *path, WordFile &w)
{
    char line[1024];
    bool content = false, begun = false, addr_dec = false, data_hex = false, ended = false;
    while (fgets(line, sizeof line, f)) {
        if (is_blank_or_comment(line)) continue;
        if (ended) { set_error("%s: text after END", path); return MDRL_EIO; }
        const char *p = skip_ws(line);
        if (!begun) {
            const std::string key = read_word(&p);
            if (key == "CONTENT" || key == "BEGIN") {
                if (key == "CONTENT") {
                    content = true;
                    p = skip_ws(p);
                    const char *q = p;
                    if (read_word(&q) == "BEGIN") { begun = true; p = q; }
                } else {
                    begun = content;
                }
                if (!begun && key == "BEGIN") { set_error("%s: BEGIN without CONTENT", path); return MDRL_EIO; }
                if (!is_blank_or_comment(p)) { set_error("%s: malformed line '%.40s'", path, line); return MDRL_EIO; }
                if (begun) {
                    if (w.width != 32 && w.width != 96) { set_error("%s: unsupported WIDTH %d", path, w.width); return MDRL_EIO; }
                    if (w.depth <= 0 || w.depth > (1 << 20)) { set_error("%s: unsupported DEPTH %d", path, w.depth); return MDRL_EIO; }
                    if (!addr_dec || !data_hex) {
                        set_error("%s: ADDRESS_RADIX = DEC and DATA_RADIX = HEX expected", path);
                        return MDRL_EIO;
                    }
                    w.parts.assign((size_t)w.depth * (w.width / 32), 0u);
                    w.given.assign((size_t)w.depth, 0);
                }
                continue;
            }
            p = skip_ws(p);
            if (content || *p != '=') { set_error("%s: malformed header line '%.40s'", path, line); return MDRL_EIO; }
            p = skip_ws(p + 1);
            long v = 0;
            std::string val;
            const bool numeric = key == "DEPTH" || key == "WIDTH";
            if (numeric ? !read_dec(&p, &v) : (val = read_word(&p)).empty()) {
                set_error("%s: malformed header line '%.40s'", path, line);
                return MDRL_EIO;
            }
            p = skip_ws(p);
            if (*p != ';' || !is_blank_or_comment(p + 1)) { set_error("%s: malformed header line '%.40s'", path, line); return MDRL_EIO; }
            if (key == "DEPTH") w.depth = (int)v;
            else if (key == "WIDTH") w.width = (int)v;
            else if (key == "ADDRESS_RADIX" && val == "DEC") addr_dec = true;
            else if (key == "DATA_RADIX" && val == "HEX") data_hex = true;
            else { set_error("%s: unsupported header line '%.40s'", path, line); return MDRL_EIO; }
            continue;
        }
        {
            const char *q = p;
            if (read_word(&q) == "END") {
                q = skip_ws(q);
                if (*q != ';' || !is_blank_or_comment(q + 1)) { set_error("%s: malformed END line", path); return MDRL_EIO; }
                ended = true;
                continue;
            }
        }
        long lo, hi;
        if (*p == '[') {
            p = skip_ws(p + 1);
            bool ok = read_dec(&p, &lo);
            p = skip_ws(p);
            ok = ok && p[0] == '.' && p[1] == '.';
            if (ok) p = skip_ws(p + 2);
            ok = ok && read_dec(&p, &hi);
            if (ok) p = skip_ws(p);
            if (!ok || *p != ']') { set_error("%s: malformed range '%.40s'", path, line); return MDRL_EIO; }
            ++p;
        } else if (!read_dec(&p, &lo)) {
            set_error("%s: malformed address '%.40s'", path, line);
            return MDRL_EIO;
        } else {
            hi = lo;
        }
        p = skip_ws(p);
        if (*p != ':') { set_error("%s: malformed line '%.40s'", path, line); return MDRL_EIO; }
        const char *h = skip_ws(p + 1);
        size_t len = 0;
        while (hex_digit(h[len]) >= 0) ++len;
        p = skip_ws(h + len);
        if (*p != ';' || !is_blank_or_comment(p + 1)) { set_error("%s: malformed line '%.40s'", path, line); return MDRL_EIO; }
        const int np = w.width / 32;
        uint32_t v[3];
        if (len == 0 || !parse_hex_value(h, len, np, v)) { set_error("%s: malformed value '%.40s'", path, line); return MDRL_EIO; }
        if (hi < lo || hi >= w.depth) { set_error("%s: address %ld beyond DEPTH %d", path, hi, w.depth); return MDRL_EIO; }
        for (long a = lo; a <= hi; ++a) {
            for (int k = 0; k < np; ++k) w.parts[(size_t)a * np + k] = v[k];
            w.given[(size_t)a] = 1;
        }
    }
    if (!ended) { set_error("%s: no END", path); return MDRL_EIO; }
    return MDRL_OK;
}

int read_hex_words(FILE *f, const char *path, WordFile &w)
{
    char line[1024];
    std::vector<uint32_t> parts;
    std::vector<uint8_t> given;
    int width = 0, depth = 0;
    bool eof = false;
    while (fgets(line, sizeof line, f)) {
        if (line[0] != ':') {
            if (blank(line)) continue;
            set_error("%s: line without record mark '%.40s'", path, line);
            return MDRL_EIO;
        }
        // ':' cc aaaa tt data[cc] ss, every field two hex digits a byte, then the line end
        const size_t len = strlen(line);
        const int cnt = len >= 3 ? hex_byte(line + 1) : -1;
        if (cnt < 0 || len < 11 + 2 * (size_t)cnt) { set_error("%s: malformed record '%.40s'", path, line); return MDRL_EIO; }
        int b[5 + 255];
        unsigned sum = 0;
        for (int k = 0; k < 5 + cnt; ++k) {
            b[k] = hex_byte(line + 1 + 2 * k);
            if (b[k] < 0) { set_error("%s: malformed record '%.40s'", path, line); return MDRL_EIO; }
            sum += (unsigned)b[k];
        }
        if (!blank(line + 11 + 2 * cnt)) { set_error("%s: trailing characters in record '%.40s'", path, line); return MDRL_EIO; }
        const unsigned addr = ((unsigned)b[1] << 8) | (unsigned)b[2], type = (unsigned)b[3];
        const int *data = b + 4;
        if ((sum & 0xffu) != 0) {
            set_error("%s: checksum mismatch at word %u", path, addr);
            return MDRL_EIO;
        }
        if (type == 1) { eof = true; break; }
        if (type != 0) continue;
        if (cnt != 4 && cnt != 12) { set_error("%s: %d-byte record (4 or 12 expected)", path, cnt); return MDRL_EIO; }
        if (width == 0) width = cnt * 8;
        if (cnt * 8 != width) { set_error("%s: mixed record sizes", path); return MDRL_EIO; }
        const int np = width / 32;
        if ((int)addr + 1 > depth) {
            depth = (int)addr + 1;
            parts.resize((size_t)depth * np, 0u);
            given.resize((size_t)depth, 0);
        }
        given[addr] = 1;
        for (int k = 0; k < np; ++k)
            parts[(size_t)addr * np + k] = ((uint32_t)data[4 * k] << 24) | ((uint32_t)data[4 * k + 1] << 16) |
                                           ((uint32_t)data[4 * k + 2] << 8) | (uint32_t)data[4 * k + 3];
    }
    if (!eof) { set_error("%s: no end-of-file record", path); return MDRL_EIO; }
    w.width = width;
    w.depth = depth;
    w.parts.swap(parts);
    w.given.swap(given);
    return MDRL_OK;
}

}  // namespace

int mdrl::read_words(const char *path, WordFile &w)
{
    FILE *f = fopen(path, "rb");
    if (!f) { set_error("cannot open %s", path); return MDRL_EIO; }
    int rc;
    if (ends_with(path, ".mif")) rc = read_mif_words(f, path, w);
    else if (ends_with(path, ".hex")) rc = read_hex_words(f, path, w);
    else { set_error("%s: not a .mif or .hex file", path); rc = MDRL_EINVAL; }
    fclose(f);
    return rc;
}

namespace {

using mdrl::read_words;

float as_float(uint32_t u)
{
    float x;
    memcpy(&x, &u, 4);
    return x;
}

uint32_t as_word(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    return u;
}

// Writes depth words of np parts; word 0 of a cell file gets the generator's "%X" count.
int write_words(const char *path, int width, int depth, const std::vector<uint32_t> &parts, bool cell_count_word,
                int last_mif_word, int flags)
{
    const bool mif = ends_with(path, ".mif"), hex = ends_with(path, ".hex");
    if (!mif && !hex) { set_error("%s: not a .mif or .hex file", path); return MDRL_EINVAL; }
    FILE *f = fopen(path, "wb");
    if (!f) { set_error("cannot create %s", path); return MDRL_EIO; }
    const char *nl = (flags & MDRL_FILE_CRLF) ? "\r\n" : "\n";
    const int np = width / 32;
    if (mif) {
        fprintf(f, "DEPTH = %d;%sWIDTH = %d;%sADDRESS_RADIX = DEC;%sDATA_RADIX = HEX;%sCONTENT%sBEGIN%s",
                depth, nl, width, nl, nl, nl, nl, nl);
        for (int a = 0; a <= last_mif_word; ++a) {
            const uint32_t *v = &parts[(size_t)a * np];
            if (a == 0 && cell_count_word) {
                fprintf(f, "0 : 0000000000000000000000%X;%s", v[2], nl);
                continue;
            }
            fprintf(f, "%d : ", a);
            for (int k = 0; k < np; ++k) fprintf(f, "%08X", v[k]);
            fprintf(f, ";%s", nl);
        }
        fprintf(f, "END;%s", nl);
    } else {
        for (int a = 0; a < depth; ++a) {
            const unsigned nb = (unsigned)np * 4;
            unsigned sum = nb + ((unsigned)a >> 8) + ((unsigned)a & 0xffu);
            fprintf(f, ":%02X%04X00", nb, (unsigned)a);
            for (int k = 0; k < np; ++k) {
                const uint32_t v = parts[(size_t)a * np + k];
                for (int s = 24; s >= 0; s -= 8) {
                    const unsigned byte = (v >> s) & 0xffu;
                    sum += byte;
                    fprintf(f, "%02X", byte);
                }
            }
            fprintf(f, "%02X%s", (0x100u - (sum & 0xffu)) & 0xffu, nl);
        }
        fprintf(f, ":00000001FF%s", nl);
    }
    const bool ok = ferror(f) == 0;
    if (fclose(f) != 0 || !ok) { set_error("write error on %s", path); return MDRL_EIO; }
    return MDRL_OK;
}

// Particle count of a cell file's word 0 (negative status if the words are no cell file).
int cell_count(const WordFile &w, const char *path)
{
    if (w.width != 96 || w.depth < 1) { set_error("%s: not a cell file (WIDTH %d)", path, w.width); return MDRL_EIO; }
    const uint32_t *p = w.parts.data();
    if (p[0] != 0 || p[1] != 0) { set_error("%s: word 0 is not a particle count", path); return MDRL_EIO; }
    if (p[2] > (uint32_t)(w.depth - 1)) { set_error("%s: count %u beyond DEPTH %d", path, p[2], w.depth); return MDRL_EIO; }
    return (int)p[2];
}

// Words 1..count as x, y, z rows (the words hold {z, y, x}).
void cell_xyz(const WordFile &w, int count, float *xyz, size_t stride)
{
    for (int k = 0; k < count; ++k) {
        const uint32_t *v = w.parts.data() + 3 * (size_t)(k + 1);
        xyz[stride * k] = as_float(v[2]);
        xyz[stride * k + 1] = as_float(v[1]);
        xyz[stride * k + 2] = as_float(v[0]);
    }
}

}  // namespace

extern "C" int md_rl_cell_file_read(const char *path, float *xyz, int cap)
{
    if (!path || cap < 0 || (cap > 0 && !xyz)) { set_error("md_rl_cell_file_read: bad argument"); return MDRL_EINVAL; }
    WordFile w;
    int rc = read_words(path, w);
    if (rc) return rc;
    const int count = cell_count(w, path);
    if (count < 0) return count;
    if (count > cap) { set_error("%s: %d particles, room for %d", path, count, cap); return MDRL_EINVAL; }
    cell_xyz(w, count, xyz, 3);
    return count;
}

extern "C" int md_rl_cell_file_write(const char *path, const float *xyz, int count, int depth, int flags)
{
    if (!path || count < 0 || (count > 0 && !xyz) || depth < count + 1 || depth > 65536) {
        set_error("md_rl_cell_file_write: bad argument (count %d, depth %d)", count, depth);
        return MDRL_EINVAL;
    }
    std::vector<uint32_t> parts(3 * (size_t)depth, 0u);
    parts[2] = (uint32_t)count;
    for (int k = 0; k < count; ++k) {
        uint32_t *v = &parts[3 * (size_t)(k + 1)];
        v[0] = as_word(xyz[3 * (size_t)k + 2]);
        v[1] = as_word(xyz[3 * (size_t)k + 1]);
        v[2] = as_word(xyz[3 * (size_t)k]);
    }
    // the generator pads with zero words only while count + 1 < DEPTH (up to DEPTH - 1)
    const int last = count + 1 < depth ? depth - 1 : count;
    return write_words(path, 96, depth, parts, true, last, flags);
}

extern "C" int md_rl_boundary_file_read(const char *path, float *bounds, int cap)
{
    if (!path || cap < 0 || (cap > 0 && !bounds)) { set_error("md_rl_boundary_file_read: bad argument"); return MDRL_EINVAL; }
    WordFile w;
    int rc = read_words(path, w);
    if (rc) return rc;
    if (w.width != 32) { set_error("%s: not a boundary file (WIDTH %d)", path, w.width); return MDRL_EIO; }
    if (w.depth > cap) { set_error("%s: %d words, room for %d", path, w.depth, cap); return MDRL_EINVAL; }
    for (int k = 0; k < w.depth; ++k) bounds[k] = as_float(w.parts[k]);
    return w.depth;
}

extern "C" int md_rl_boundary_file_write(const char *path, float edge, int depth, int flags)
{
    if (!path || depth < 1 || depth > 65536 || !(edge > 0.0f)) { set_error("md_rl_boundary_file_write: bad argument"); return MDRL_EINVAL; }
    std::vector<uint32_t> parts(depth);
    for (int k = 0; k < depth; ++k) parts[k] = as_word((float)k * edge);
    return write_words(path, 32, depth, parts, false, depth - 1, flags);
}

extern "C" int md_rl_cell_dir_read(const char *dir, const char *prefix, const int lo[3], const int hi[3],
                                   float *particle, int32_t *cell, int cap)
{
    if (!dir || !lo || !hi || cap < 0 || (cap > 0 && !particle)) { set_error("md_rl_cell_dir_read: bad argument"); return MDRL_EINVAL; }
    for (int d = 0; d < 3; ++d)
        if (lo[d] < 1 || hi[d] < lo[d] || hi[d] > 999) { set_error("md_rl_cell_dir_read: bad cell range"); return MDRL_EINVAL; }
    const char *pre = prefix ? prefix : "cell_ini_file_";
    int n = 0;
    for (int x = lo[0]; x <= hi[0]; ++x)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int z = lo[2]; z <= hi[2]; ++z) {
                std::string base = std::string(dir) + "/" + pre + std::to_string(x) + "_" + std::to_string(y) + "_" +
                                   std::to_string(z);
                std::string path = base + ".hex";
                FILE *probe = fopen(path.c_str(), "rb");
                if (probe) fclose(probe);
                else path = base + ".mif";
                WordFile w;
                int rc = read_words(path.c_str(), w);
                if (rc) return rc;
                const int m = cell_count(w, path.c_str());
                if (m < 0) return m;
                if (n + m > cap) { set_error("md_rl_cell_dir_read: more than %d particles", cap); return MDRL_EINVAL; }
                for (int k = 0; k < m; ++k)
                    for (int c = 0; c < 11; ++c) particle[11 * (size_t)(n + k) + c] = 0.0f;
                cell_xyz(w, m, particle + 11 * (size_t)n, 11);
                for (int k = 0; k < m; ++k) {
                    if (cell) {
                        cell[3 * (size_t)(n + k)] = x;
                        cell[3 * (size_t)(n + k) + 1] = y;
                        cell[3 * (size_t)(n + k) + 2] = z;
                    }
                }
                n += m;
            }
    return n;
}
