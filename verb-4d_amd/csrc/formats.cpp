// The reference's on-disk formats, host side (no device work):
//   .plt   text: a VARIABLES line, a ZONE line with the sizes (I fastest:
//          I=K, J=V, K=R, L=P for [P][R][V][K]), then one point per line,
//          either the value alone (src/Matrix.cpp:4556-4567, 3011-3024) or
//          the point's grid coordinates and the value, tab-separated
//          (:4833-4853, 3031-3050); values in "%.6e";
//   .pltb  binary: the sizes as int32, then the doubles (:5704-5726,
//          5792-5818, 3781-3838);
//   grid.pltb / grid.plt  the P, R, V, K grids (src/ReadInitialData.cpp:
//          253-346);
//   .tab   one line: data file or .lst list, then scale, then from/to limit
//          pairs per axis (src/UpdatableMatrix.cpp:308-551);
//   .lst   queues of "time value" lines (:345-370);
//   PSD.lst  "info<TAB>file" lines appended per output (src/Matrix.cpp:
//          5764-5785).
// Reading follows the reference's stream semantics: tokens are whitespace
// runs, a number is the longest prefix an istream would take for a double
// ([+-]digits[.digits][e[+-]digits]) and must convert entirely, a value
// line is consumed up to its newline (at most 9999 characters) after the
// value. Where the reference's stream would fail (a token that is not a
// number, "nan"/"inf", a short file) these readers return VERB_ERR_FORMAT
// or VERB_ERR_IO instead of reading zeros or exiting.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <strings.h>
#include <vector>

#include "verb4d_hip.h"

namespace {

// Relative grid tolerance of the checked readers, in decades (src/Matrix.cpp:42)
constexpr double kGridErr = 1e-6;

struct Text {
    std::string s;
    size_t p = 0;
    bool load(const char* path) {
        std::ifstream f(path, std::ios::binary);
        if (!f.is_open()) return false;
        s.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        return true;
    }
    bool eof() const { return p >= s.size(); }
    void skip_ws() {
        while (p < s.size() && isspace((unsigned char)s[p])) p++;
    }
    // istream >> std::string
    bool token(std::string& out) {
        skip_ws();
        if (p >= s.size()) return false;
        size_t b = p;
        while (p < s.size() && !isspace((unsigned char)s[p])) p++;
        out.assign(s, b, p - b);
        return true;
    }
    // getline(istream, str, delim)
    bool getline(std::string& out, char delim = '\n') {
        if (p >= s.size()) return false;
        size_t e = s.find(delim, p);
        if (e == std::string::npos) {
            out.assign(s, p, std::string::npos);
            p = s.size();
        } else {
            out.assign(s, p, e - p);
            p = e + 1;
        }
        return true;
    }
    // istream::ignore(9999, '\n')
    void ignore_line() {
        size_t n = 0;
        while (p < s.size() && n < 9999) {
            n++;
            if (s[p++] == '\n') break;
        }
    }
    // istream >> double
    bool number(double& v) {
        skip_ws();
        size_t b = p;
        if (p < s.size() && (s[p] == '+' || s[p] == '-')) p++;
        while (p < s.size() && isdigit((unsigned char)s[p])) p++;
        if (p < s.size() && s[p] == '.') {
            p++;
            while (p < s.size() && isdigit((unsigned char)s[p])) p++;
        }
        if (p < s.size() && (s[p] == 'e' || s[p] == 'E')) {
            p++;
            if (p < s.size() && (s[p] == '+' || s[p] == '-')) p++;
            while (p < s.size() && isdigit((unsigned char)s[p])) p++;
        }
        return convert(s.substr(b, p - b), v);
    }
    static bool convert(const std::string& t, double& v) {
        if (t.empty()) return false;
        char* end = nullptr;
        errno = 0;
        v = strtod(t.c_str(), &end);
        // an overflow fails the stream too (its value would be +-max)
        return end == t.c_str() + t.size() && !std::isinf(v);
    }
};

// Skip the header of a column reader (src/Matrix.cpp:4919-4942 for 4-D;
// the 3-D reader, :3070-3074, skips tokens to ZONE without the VARIABLES
// handling), then the rest of the ZONE line.
void skip_header(Text& t, bool variables_lines, bool case_insensitive) {
    std::string tok, line;
    t.token(tok);
    auto is_zone = [&](const std::string& x) {
        return case_insensitive ? strcasecmp(x.c_str(), "ZONE") == 0 : x == "ZONE";
    };
    while (!is_zone(tok) && !t.eof()) {
        if (variables_lines && strcasecmp(tok.c_str(), "VARIABLES") != 0) t.getline(line);
        if (!t.token(tok)) break;
    }
    t.ignore_line();
}

size_t count(int ndim, const int* s) {
    size_t n = 1;
    for (int i = 0; i < ndim; i++) n *= (size_t)s[i];
    return n;
}

bool sizes_ok(int ndim, const int* s) {
    if ((ndim != 3 && ndim != 4) || !s) return false;
    for (int i = 0; i < ndim; i++)
        if (s[i] < 1) return false;
    return true;
}

void put_e(std::string& o, double v) {
    char b[64];
    snprintf(b, sizeof(b), "%e", v);
    o += b;
}

bool write_all(const char* path, const std::string& o) {
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    const bool ok = fwrite(o.data(), 1, o.size(), f) == o.size();
    return fclose(f) == 0 && ok;
}

// ZONE line of a .plt file: sizes listed fastest first
void zone(std::string& o, const char* title, int ndim, const int* s) {
    o += "ZONE T=\"";
    o += title;
    o += "\"";
    const char* lab[4] = {"I", "J", "K", "L"};
    for (int i = 0; i < ndim; i++) {
        o += ", ";
        o += lab[i];
        o += "=";
        o += std::to_string(s[ndim - 1 - i]);
    }
    o += "\n";
}

}  // namespace

extern "C" {

int verb_plt_write(const char* path, int ndim, const int* sizes, const double* data, const char* name,
                   const char* info) {
    if (!path || !sizes_ok(ndim, sizes) || !data) return VERB_ERR_ARG;
    const size_t n = count(ndim, sizes);
    std::string o;
    o.reserve(n * 14 + 256);
    o += "VARIABLES = \"";
    o += (name && *name) ? name : "f";
    o += "\" \n";
    zone(o, (info && *info) ? info : path, ndim, sizes);
    for (size_t i = 0; i < n; i++) {
        put_e(o, data[i]);
        o += '\n';
    }
    return write_all(path, o) ? VERB_OK : VERB_ERR_IO;
}

int verb_plt_write_grid(const char* path, int ndim, const int* sizes, const double* const* grids,
                        const char* const* grid_names, const double* data, const char* name) {
    if (!path || !sizes_ok(ndim, sizes) || !data || !grids) return VERB_ERR_ARG;
    for (int a = 0; a < ndim; a++)
        if (!grids[a]) return VERB_ERR_ARG;
    const char* dflt4[4] = {"w", "x", "y", "z"};
    const char* dflt3[3] = {"x", "y", "z"};
    const size_t n = count(ndim, sizes);
    std::string o;
    o.reserve(n * 14 * (ndim + 1) + 256);
    o += "VARIABLES = ";
    for (int a = 0; a < ndim; a++) {
        const char* g = (grid_names && grid_names[a] && *grid_names[a]) ? grid_names[a] : (ndim == 4 ? dflt4 : dflt3)[a];
        o += "\"";
        o += g;
        o += "\", ";
    }
    o += "\"";
    o += (name && *name) ? name : "f";
    o += "\" \n";
    zone(o, path, ndim, sizes);
    for (size_t i = 0; i < n; i++) {
        for (int a = 0; a < ndim; a++) {
            o += '\t';
            put_e(o, grids[a][i]);
        }
        o += '\t';
        put_e(o, data[i]);
        o += '\n';
    }
    return write_all(path, o) ? VERB_OK : VERB_ERR_IO;
}

int verb_plt_sizes(const char* path, int* sizes) {
    if (!path || !sizes) return VERB_ERR_ARG;
    Text t;
    if (!t.load(path)) return VERB_ERR_IO;
    // the grid.plt header parse of src/ReadInitialData.cpp:297-310 (pieces
    // up to each '=' until one ends in the size's letter, then the size),
    // kept within the ZONE line so that files without L stop at K
    std::string tok;
    t.token(tok);
    while (strcasecmp(tok.c_str(), "ZONE") != 0 && !t.eof()) t.token(tok);
    if (strcasecmp(tok.c_str(), "ZONE") != 0) return VERB_ERR_FORMAT;
    size_t eol = t.s.find('\n', t.p);
    if (eol == std::string::npos) eol = t.s.size();
    const std::string zl = t.s.substr(t.p, eol - t.p);
    const char lab[4] = {'I', 'J', 'K', 'L'};
    int v[4] = {0, 0, 0, 0}, nd = 0;
    size_t q = 0;
    for (int i = 0; i < 4; i++) {
        bool found = false;
        for (size_t e; (e = zl.find('=', q)) != std::string::npos;) {
            const std::string chunk = zl.substr(q, e - q);
            q = e + 1;
            if (!chunk.empty() && chunk.back() == lab[i]) {
                found = true;
                break;
            }
        }
        if (!found) break;
        const char* b = zl.c_str() + q;
        char* end = nullptr;
        const long x = strtol(b, &end, 10);
        if (end == b || x < 1 || x > 0x7fffffffL) return VERB_ERR_FORMAT;
        q += (size_t)(end - b);
        v[i] = (int)x;
        nd++;
    }
    if (nd < 3) return VERB_ERR_FORMAT;
    for (int i = 0; i < 4; i++) sizes[i] = 0;
    for (int i = 0; i < nd; i++) sizes[nd - 1 - i] = v[i];
    return nd;
}

int verb_plt_read(const char* path, int column, int ndim, const int* sizes, double* out) {
    if (!path || !sizes_ok(ndim, sizes) || !out || column < 1) return VERB_ERR_ARG;
    Text t;
    if (!t.load(path) || t.eof()) return VERB_ERR_IO;
    skip_header(t, ndim == 4, true);
    const size_t n = count(ndim, sizes);
    std::string tok;
    for (size_t i = 0; i < n; i++) {
        for (int c = 1; c < column; c++)
            if (!t.token(tok)) return VERB_ERR_FORMAT;
        if (!t.number(out[i])) return VERB_ERR_FORMAT;
        t.ignore_line();
    }
    // more data than the sizes: the reference's "Check grid size" (:4957-4961)
    if (t.token(tok)) return VERB_ERR_FORMAT;
    return VERB_OK;
}

int verb_plt_read_checked(const char* path, int ndim, const int* sizes, const double* const* grids, double* out,
                          long long* bad_index) {
    if (bad_index) *bad_index = -1;
    if (!path || !sizes_ok(ndim, sizes) || !out || !grids) return VERB_ERR_ARG;
    for (int a = 0; a < ndim; a++)
        if (!grids[a]) return VERB_ERR_ARG;
    Text t;
    if (!t.load(path) || t.eof()) return VERB_ERR_IO;
    skip_header(t, ndim == 4, false);
    const size_t n = count(ndim, sizes);
    for (size_t i = 0; i < n; i++) {
        double g[4];
        for (int a = 0; a < ndim; a++)
            if (!t.number(g[a])) return VERB_ERR_FORMAT;
        for (int a = 0; a < ndim; a++) {
            // (a NaN difference passes, as the reference's comparison does)
            if (fabs(log10(g[a]) - log10(grids[a][i])) > kGridErr) {
                if (bad_index) *bad_index = (long long)i;
                return VERB_ERR_FORMAT;
            }
        }
        if (!t.number(out[i])) return VERB_ERR_FORMAT;
        t.ignore_line();
    }
    return VERB_OK;
}

int verb_pltb_write(const char* path, int ndim, const int* sizes, const double* data) {
    if (!path || !sizes_ok(ndim, sizes) || !data) return VERB_ERR_ARG;
    FILE* f = fopen(path, "wb");
    if (!f) return VERB_ERR_IO;
    int32_t s[4];
    for (int i = 0; i < ndim; i++) s[i] = sizes[i];
    bool ok = fwrite(s, sizeof(int32_t), ndim, f) == (size_t)ndim;
    const size_t n = count(ndim, sizes);
    ok = ok && fwrite(data, sizeof(double), n, f) == n;
    return (fclose(f) == 0 && ok) ? VERB_OK : VERB_ERR_IO;
}

int verb_pltb_read(const char* path, int ndim, const int* sizes, double* out, int* file_sizes) {
    if (!path || !sizes_ok(ndim, sizes) || !out) return VERB_ERR_ARG;
    FILE* f = fopen(path, "rb");
    if (!f) return VERB_ERR_IO;
    int32_t s[4] = {0, 0, 0, 0};
    bool ok = fread(s, sizeof(int32_t), ndim, f) == (size_t)ndim;
    const size_t n = count(ndim, sizes);
    ok = ok && fread(out, sizeof(double), n, f) == n;
    fclose(f);
    if (file_sizes)
        for (int i = 0; i < ndim; i++) file_sizes[i] = s[i];
    return ok ? VERB_OK : VERB_ERR_IO;
}

int verb_grid_read(const char* folder, int* sizes, double* P, double* R, double* V, double* K, size_t capacity) {
    if (!folder || !sizes) return VERB_ERR_ARG;
    const std::string dir(folder);
    double* g[4] = {P, R, V, K};
    FILE* f = fopen((dir + "grid.pltb").c_str(), "rb");
    if (f) {
        int32_t s[4];
        if (fread(s, sizeof(int32_t), 4, f) != 4) {
            fclose(f);
            return VERB_ERR_IO;
        }
        for (int i = 0; i < 4; i++) sizes[i] = s[i];
        if (s[0] < 1 || s[1] < 1 || s[2] < 1 || s[3] < 1) {
            fclose(f);
            return VERB_ERR_FORMAT;
        }
        const size_t n = (size_t)s[0] * s[1] * s[2] * s[3];
        if (!P) {  // sizes only
            fclose(f);
            return VERB_OK;
        }
        if (n > capacity || !R || !V || !K) {
            fclose(f);
            return VERB_ERR_ARG;
        }
        for (int j = 0; j < 4; j++)
            if (fread(g[j], sizeof(double), n, f) != n) {
                fclose(f);
                return VERB_ERR_IO;
            }
        fclose(f);
        return VERB_OK;
    }
    const std::string plt = dir + "grid.plt";
    int s[4];
    const int nd = verb_plt_sizes(plt.c_str(), s);
    if (nd < 0) return nd;
    if (nd != 4) return VERB_ERR_FORMAT;
    for (int i = 0; i < 4; i++) sizes[i] = s[i];
    if (!P) return VERB_OK;
    const size_t n = (size_t)s[0] * s[1] * s[2] * s[3];
    if (n > capacity || !R || !V || !K) return VERB_ERR_ARG;
    for (int j = 0; j < 4; j++) {
        const int rc = verb_plt_read(plt.c_str(), j + 1, 4, s, g[j]);
        if (rc) return rc;
    }
    return VERB_OK;
}

int verb_psd_lst_append(const char* output_folder, const char* filename, int binary, const char* info) {
    if (!output_folder || !filename) return VERB_ERR_ARG;
    std::string full = std::string(filename) + (binary ? ".pltb" : ".plt");
    const std::string folder(output_folder);
    const size_t pos = full.find(folder);
    if (pos != std::string::npos) full.erase(pos, folder.size());
    FILE* f = fopen((folder + "PSD.lst").c_str(), "ab");
    if (!f) return VERB_ERR_IO;
    std::string line = std::string(info ? info : "") + "\t" + full + "\n";
    const bool ok = fwrite(line.data(), 1, line.size(), f) == line.size();
    return (fclose(f) == 0 && ok) ? VERB_OK : VERB_ERR_IO;
}

int verb_tab_read(const char* tab_path, verb_tab_line* out) {
    if (!tab_path || !out) return VERB_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    Text t;
    if (!t.load(tab_path)) return VERB_ERR_IO;
    std::string line;
    if (!t.getline(line)) return VERB_OK;  // empty file: nothing to read
    // a comment or empty first line leaves the array as it was (:326-337)
    if (line.empty() || line.find('#') != std::string::npos) return VERB_OK;
    out->has_line = 1;
    Text l;
    l.s = line;
    auto copy = [](char* dst, const std::string& src) {
        std::strncpy(dst, src.c_str(), VERB_TAB_STRLEN - 1);
        dst[VERB_TAB_STRLEN - 1] = 0;
        return src.size() < VERB_TAB_STRLEN;
    };
    std::string data;
    l.token(data);
    if (!copy(out->data, data)) return VERB_ERR_FORMAT;
    if (!data.empty() && data != "-") {
        // the reference's suffix test, size_t arithmetic included (:407)
        const size_t ext = std::strlen(".lst");
        out->data_kind = (data.rfind(".lst") == data.length() - ext) ? 2 : 1;
    }
    // each later field is read when the line's stream has not reached its
    // end (:427, 447, ...): a field read after trailing blanks is empty
    std::string tok;
    if (!l.eof()) {
        tok.clear();
        l.token(tok);
        out->scale_present = 1;
        if (!copy(out->scale, tok)) return VERB_ERR_FORMAT;
    }
    for (int a = 0; a < 4; a++) {
        if (l.eof()) break;
        std::string f, to;
        l.token(f);
        l.token(to);
        out->nlimits = a + 1;
        if (!copy(out->from[a], f) || !copy(out->to[a], to)) return VERB_ERR_FORMAT;
    }
    return VERB_OK;
}

int verb_is_number(const char* s) {
    // is_number (src/UpdatableMatrix.cpp:1081-1093)
    if (!s || !*s) return 0;
    for (const char* c = s; *c; c++)
        if (!(isdigit((unsigned char)*c) || *c == '.' || *c == 'e' || *c == '+' || *c == '-')) return 0;
    return 1;
}

double verb_parse_number(const char* s) {
    // std::stringstream(s) >> double: the istream prefix, 0 on failure
    if (!s) return 0.0;
    Text t;
    t.s = s;
    double v = 0.0;
    return t.number(v) ? v : 0.0;
}

int verb_queue_read(const char* path, double* times, double* values, char* names, size_t name_len, size_t capacity,
                    size_t* count) {
    // fill_queue_from_file (src/UpdatableMatrix.cpp:349-370): one "time
    // value" pair per line; values are numbers (values != NULL) or file
    // names (names != NULL, name_len bytes each). Blank lines are skipped.
    if (!path || !count) return VERB_ERR_ARG;
    *count = 0;
    Text t;
    if (!t.load(path)) return VERB_ERR_IO;
    std::string line;
    size_t k = 0;
    while (t.getline(line)) {
        Text l;
        l.s = line;
        l.skip_ws();
        if (l.eof()) continue;
        double tm = 0.0;
        if (!l.number(tm)) tm = 0.0;
        double v = 0.0;
        std::string name;
        if (names) {
            l.token(name);
        } else if (!l.number(v)) {
            v = 0.0;
        }
        if (k < capacity) {
            if (times) times[k] = tm;
            if (values) values[k] = v;
            if (names) {
                if (name.size() >= name_len) return VERB_ERR_FORMAT;
                std::memcpy(names + k * name_len, name.c_str(), name.size() + 1);
            }
        }
        k++;
    }
    *count = k;
    return k > capacity && (times || values || names) ? VERB_ERR_ARG : VERB_OK;
}

}  // extern "C"
