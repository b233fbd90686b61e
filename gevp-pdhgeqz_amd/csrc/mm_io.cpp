// mm_io.cpp -- Matrix Market input for the driver path (the roles of tools/mmio.f mminfo2 / mmread2 and
// tools/PDRDMM.f:6-90 of NLAFET/GEVP-PDHGEQZ on a 1x1 grid): the header of a .mtx file, and the whole
// matrix read into a dense column-major host array.  Host code only; no device work.
//
// Supported: "%%MatrixMarket matrix coordinate|array real|integer|pattern general|symmetric|
// skew-symmetric".  As PDRDMM does, the array is cleared first, each stored entry (i, j, v) is set
// and, for a symmetric file, (j, i) too (skew-symmetric: -v).  Pattern entries are 1.  Complex and
// Hermitian files are rejected.
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gqz.h"

namespace {

struct MmHeader {
    std::string rep, field, symm;
    int m = 0, n = 0;
    long nnz = 0;
};

// whole file in memory; pos points after the size line on success
int mm_parse_header(const std::string &text, MmHeader &h, size_t &pos) {
    size_t e = text.find('\n');
    std::string banner = text.substr(0, e == std::string::npos ? text.size() : e);
    char w0[64] = {}, w1[64] = {}, w2[64] = {}, w3[64] = {}, w4[64] = {};
    if (std::sscanf(banner.c_str(), "%63s %63s %63s %63s %63s", w0, w1, w2, w3, w4) != 5) return -2;
    auto lower = [](char *s) { for (; *s; ++s) *s = (char)std::tolower((unsigned char)*s); };
    lower(w1); lower(w2); lower(w3); lower(w4);
    if (std::strcmp(w0, "%%MatrixMarket") != 0 || std::strcmp(w1, "matrix") != 0) return -2;
    h.rep = w2; h.field = w3; h.symm = w4;
    if (h.rep != "coordinate" && h.rep != "array") return -3;
    if (h.field != "real" && h.field != "integer" && h.field != "pattern" && h.field != "double") return -3;
    if (h.symm != "general" && h.symm != "symmetric" && h.symm != "skew-symmetric") return -3;
    if (h.rep == "array" && h.field == "pattern") return -3;
    pos = (e == std::string::npos) ? text.size() : e + 1;
    // comment lines, then the size line
    while (pos < text.size()) {
        size_t le = text.find('\n', pos);
        if (le == std::string::npos) le = text.size();
        size_t p = pos;
        while (p < le && std::isspace((unsigned char)text[p])) ++p;
        if (p == le || text[p] == '%') { pos = le + 1; continue; }
        std::string line = text.substr(pos, le - pos);
        long a = 0, b = 0, c = 0;
        const int got = std::sscanf(line.c_str(), "%ld %ld %ld", &a, &b, &c);
        pos = le + 1;
        if (h.rep == "coordinate") {
            if (got != 3 || a < 0 || b < 0 || c < 0) return -4;
            h.nnz = c;
        } else {
            if (got < 2 || a < 0 || b < 0) return -4;
            h.nnz = (h.symm == "general") ? a * b : a * (a + 1) / 2 - (h.symm == "skew-symmetric" ? a : 0);
        }
        if (a > 0x7fffffffL || b > 0x7fffffffL) return -4;
        h.m = (int)a; h.n = (int)b;
        return 0;
    }
    return -4;
}

int mm_load(const char *path, std::string &text) {
    if (!path) return -1;
    FILE *f = std::fopen(path, "rb");
    if (!f) return -1;
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) { std::fclose(f); return -1; }
    text.resize((size_t)sz);
    const size_t rd = sz ? std::fread(&text[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    return rd == (size_t)sz ? 0 : -1;
}

void copy_str(char *dst, const std::string &s, size_t cap) {
    if (!dst) return;
    std::strncpy(dst, s.c_str(), cap - 1);
    dst[cap - 1] = '\0';
}

}  // namespace

extern "C" {

int gqz_mm_info(const char *path, char *rep, char *field, char *symm, int *nrows, int *ncols, long *nnz) {
    std::string text;
    if (mm_load(path, text)) return -1;
    MmHeader h;
    size_t pos = 0;
    const int r = mm_parse_header(text, h, pos);
    if (r) return r;
    copy_str(rep, h.rep, 11); copy_str(field, h.field, 8); copy_str(symm, h.symm, 20);
    if (nrows) *nrows = h.m;
    if (ncols) *ncols = h.n;
    if (nnz) *nnz = h.nnz;
    return 0;
}

int gqz_mm_read_dense(const char *path, int nrows, int ncols, double *A, int lda) {
    std::string text;
    if (mm_load(path, text)) return -1;
    MmHeader h;
    size_t pos = 0;
    int r = mm_parse_header(text, h, pos);
    if (r) return r;
    if (h.m != nrows || h.n != ncols) return -7;
    if (!A || lda < (nrows > 1 ? nrows : 1)) return -8;
    if (h.symm != "general" && h.m != h.n) return -4;
    for (int j = 0; j < ncols; ++j)
        std::memset(A + (size_t)j * lda, 0, sizeof(double) * (size_t)nrows);
    const char *p = text.c_str() + pos, *end = text.c_str() + text.size();
    auto skip = [&]() {       // whitespace and comment lines
        for (;;) {
            while (p < end && std::isspace((unsigned char)*p)) ++p;
            if (p < end && *p == '%') { while (p < end && *p != '\n') ++p; continue; }
            return;
        }
    };
    auto num_l = [&](long &v) -> bool {
        skip();
        if (p >= end) return false;
        char *q = nullptr;
        errno = 0;
        v = std::strtol(p, &q, 10);
        if (q == p || errno) return false;
        p = q;
        return true;
    };
    auto num_d = [&](double &v) -> bool {
        skip();
        if (p >= end) return false;
        char *q = nullptr;
        v = std::strtod(p, &q);
        if (q == p) return false;
        p = q;
        return true;
    };
    const bool sym = h.symm == "symmetric", skew = h.symm == "skew-symmetric";
    auto put = [&](long i, long j, double v) -> bool {
        if (i < 1 || j < 1 || i > nrows || j > ncols) return false;
        A[(size_t)(i - 1) + (size_t)(j - 1) * lda] = v;
        if ((sym || skew) && i != j) A[(size_t)(j - 1) + (size_t)(i - 1) * lda] = skew ? -v : v;
        return true;
    };
    if (h.rep == "coordinate") {
        for (long k = 0; k < h.nnz; ++k) {
            long i, j;
            double v = 1.0;
            if (!num_l(i) || !num_l(j)) return -5;
            if (h.field != "pattern" && !num_d(v)) return -5;
            if (!put(i, j, v)) return -6;
        }
    } else {
        // column-major values; symmetric arrays hold the lower triangle (skew: strictly lower)
        for (long j = 1; j <= ncols; ++j) {
            const long i0 = sym ? j : (skew ? j + 1 : 1);
            for (long i = i0; i <= nrows; ++i) {
                double v;
                if (!num_d(v)) return -5;
                if (!put(i, j, v)) return -6;
            }
        }
    }
    return 0;
}

}  // extern "C"
