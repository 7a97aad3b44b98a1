// Matrix input from CSV and the Debug rendering of a matrix: the two data
// formats on either side of the dense path (src/util.rs:9-35,
// src/matrix/mod.rs:810-840).  Host code; la_read_csv_dev stages the parsed
// values into device memory for the solvers.
#include "la_common.hpp"
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

// The grammar of Rust's `str::parse::<f64>` (the parser the reference's
// read_csv callers pass, applied to the trimmed item): an optional sign,
// then "inf", "infinity" or "nan" in any letter case, or a decimal number
// with at least one digit and an optional exponent with at least one digit.
// Anything else (empty items, hex floats, underscores, embedded blanks) is
// a parse error, which the reference turns into a panic through unwrap().
bool rust_f64_syntax(const char *s, size_t n)
{
    size_t i = 0;
    if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
    auto ieq = [&](const char *word) {
        const size_t L = strlen(word);
        if (n - i != L) return false;
        for (size_t k = 0; k < L; ++k)
            if ((s[i + k] | 0x20) != word[k]) return false;
        return true;
    };
    if (ieq("inf") || ieq("infinity") || ieq("nan")) return true;
    size_t digits = 0;
    while (i < n && s[i] >= '0' && s[i] <= '9') { ++i; ++digits; }
    if (i < n && s[i] == '.') {
        ++i;
        while (i < n && s[i] >= '0' && s[i] <= '9') { ++i; ++digits; }
    }
    if (digits == 0) return false;
    if (i < n && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
        size_t ed = 0;
        while (i < n && s[i] >= '0' && s[i] <= '9') { ++i; ++ed; }
        if (ed == 0) return false;
    }
    return i == n;
}

// Rust's str::trim: strips Unicode White_Space from both ends; for the
// ASCII range that is ' ', '\t', '\n', '\v', '\f', '\r' (and the items never
// contain '\n').  Non-ASCII blanks make the item unparseable either way.
void trim(const char *&b, const char *&e)
{
    auto ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' ||
                                  c == '\r'; };
    while (b < e && ws(*b)) ++b;
    while (e > b && ws(e[-1])) --e;
}

// Parses the whole file.  want_values = false only counts.  Returns LA_OK,
// LA_ERR_IO (missing file, unparseable item: the reference's unwrap()
// panics) or LA_ERR_SHAPE (a row with a different item count, or no rows:
// util.rs:25-27 and :33 assert).
int parse_csv(const char *path, int64_t *rows, int64_t *cols, std::vector<double> *values,
              std::string *err)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        *err = std::string("read_csv: cannot open ") + path;
        return LA_ERR_IO;
    }
    std::string line;
    int64_t nrows = 0, ncols = -1;
    int status = LA_OK;
    std::vector<char> buf(1 << 16);
    std::string pending;
    bool eof = false;
    while (!eof && status == LA_OK) {
        // read one line (BufRead::lines: split at '\n', a trailing "\r" is
        // removed with it; a final line without '\n' still counts)
        line.clear();
        bool got = false;
        for (;;) {
            int c = fgetc(f);
            if (c == EOF) { eof = true; break; }
            got = true;
            if (c == '\n') break;
            line.push_back((char)c);
        }
        if (!got) break;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        int64_t count = 0;
        const char *p = line.data(), *end = p + line.size();
        for (;;) {                         // split(","): always at least one item
            const char *q = p;
            while (q < end && *q != ',') ++q;
            const char *b = p, *e = q;
            trim(b, e);
            if (!rust_f64_syntax(b, (size_t)(e - b))) {
                *err = "read_csv: item '" + std::string(b, e) + "' on line " +
                       std::to_string(nrows + 1) + " is not an f64";
                status = LA_ERR_IO;
                break;
            }
            if (values) {
                std::string tok(b, e);
                // strtod rounds correctly and reads inf/infinity/nan in any case
                values->push_back(strtod(tok.c_str(), nullptr));
            }
            ++count;
            if (q == end) break;
            p = q + 1;
        }
        if (status != LA_OK) break;
        if (ncols < 0) {
            ncols = count;
        } else if (count != ncols) {
            *err = "read_csv: line " + std::to_string(nrows + 1) + " has " + std::to_string(count) +
                   " items, expected " + std::to_string(ncols);
            status = LA_ERR_SHAPE;
            break;
        }
        ++nrows;
    }
    fclose(f);
    if (status != LA_OK) return status;
    if (ncols < 0) {
        *err = "read_csv: empty file";
        return LA_ERR_SHAPE;
    }
    *rows = nrows;
    *cols = ncols;
    return LA_OK;
}

// Rust's `{:?}` of an f64: the shortest decimal that reads back to the same
// value; plain notation with at least one fractional digit when
// 1e-4 <= |v| < 1e16 (or v == 0), otherwise d[.ddd]e<exp> with no '+' and no
// leading exponent zeros; "NaN", "inf", "-inf"; the sign of -0.0 kept.
std::string rust_debug_f64(double v)
{
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    char sci[64];
    auto r = std::to_chars(sci, sci + sizeof sci, v, std::chars_format::scientific);
    *r.ptr = 0;
    // sci = [-]d[.ddd]e(+|-)xx
    std::string s(sci);
    std::string sign;
    if (s[0] == '-') { sign = "-"; s.erase(0, 1); }
    const size_t epos = s.find('e');
    std::string mant = s.substr(0, epos);
    int exp10 = atoi(s.c_str() + epos + 1);
    std::string digits;
    for (char c : mant)
        if (c != '.') digits.push_back(c);
    const double a = std::fabs(v);
    if (v == 0.0 || (a >= 1e-4 && a < 1e16)) {
        // plain notation: the decimal point sits after digit exp10 + 1
        std::string out;
        const int point = exp10 + 1;
        if (point <= 0) {
            out = "0." + std::string((size_t)(-point), '0') + digits;
        } else if ((size_t)point >= digits.size()) {
            out = digits + std::string((size_t)point - digits.size(), '0') + ".0";
        } else {
            out = digits.substr(0, (size_t)point) + "." + digits.substr((size_t)point);
        }
        return sign + out;
    }
    std::string out = digits.substr(0, 1);
    if (digits.size() > 1) out += "." + digits.substr(1);
    return sign + out + "e" + std::to_string(exp10);
}

}  // namespace

using namespace lahip;

extern "C" int la_csv_shape(const char *path, int64_t *rows, int64_t *cols)
{
    if (!path || !rows || !cols) return LA_ERR_ARG;
    std::string err;
    return parse_csv(path, rows, cols, nullptr, &err);
}

extern "C" int la_read_csv_rm(const char *path, int64_t rows, int64_t cols, double *out)
{
    if (!path || !out || rows <= 0 || cols <= 0) return LA_ERR_ARG;
    int64_t r = 0, c = 0;
    std::vector<double> v;
    std::string err;
    int st = parse_csv(path, &r, &c, &v, &err);
    if (st) return st;
    if (r != rows || c != cols) return LA_ERR_ARG;
    memcpy(out, v.data(), v.size() * sizeof(double));
    return LA_OK;
}

extern "C" int la_read_csv_dev(la_ctx *ctx, const char *path, int64_t rows, int64_t cols,
                               double *d_out, int64_t ld)
{
    if (!ctx || !path || !d_out || rows <= 0 || cols <= 0 || ld < cols) return LA_ERR_ARG;
    int64_t r = 0, c = 0;
    std::vector<double> v;
    std::string err;
    int st = parse_csv(path, &r, &c, &v, &err);
    if (st) return set_error(ctx, st, err.c_str());
    if (r != rows || c != cols) return set_error(ctx, LA_ERR_ARG, "read_csv_dev: shape differs");
    LA_HIP_CHECK(ctx, hipMemcpy2DAsync(d_out, (size_t)ld * 8, v.data(), (size_t)cols * 8,
                                       (size_t)cols * 8, (size_t)rows, hipMemcpyHostToDevice,
                                       ctx->stream));
    LA_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return LA_OK;
}

extern "C" int64_t la_format_debug(int64_t rows, int64_t cols, const double *data, char *buf,
                                   int64_t cap)
{
    if (rows <= 0 || cols <= 0 || !data) return -1;
    std::vector<std::string> items((size_t)(rows * cols));
    size_t width = 0;
    for (int64_t e = 0; e < rows * cols; ++e) {
        items[(size_t)e] = rust_debug_f64(data[e]);
        width = std::max(width, items[(size_t)e].size());
    }
    std::string out = "\n";
    for (int64_t i = 0; i < rows; ++i) {
        out += "|";
        for (int64_t j = 0; j < cols; ++j) {
            const std::string &it = items[(size_t)(i * cols + j)];
            out.append(1 + width - it.size(), ' ');
            out += it;
        }
        out += " |\n";
    }
    const int64_t need = (int64_t)out.size();
    if (buf && cap > 0) {
        const int64_t k = std::min(need, cap - 1);
        memcpy(buf, out.data(), (size_t)k);
        buf[k] = 0;
    }
    return need;
}
