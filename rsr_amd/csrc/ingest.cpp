// ingest.cpp -- datasets read from disk (Dataset::from_csv / from_json,
// src/engine/data/dataset.rs:265-357), host code.
//
// The readers produce the row-major rows and targets that rsr_dataset_create takes, so a
// dataset read from disk goes through the same packing, padding and z-score as one built
// from memory.
//
// CSV, as the reference reads it with the csv crate's default reader: a header row; comma
// separated fields, double-quoted fields with "" for a quote; every record with the
// header's field count; the target in column target_col (the last column when negative);
// features in the remaining columns, left to right, as many as the schema's flat size;
// fields parsed as f32 with no surrounding whitespace (Rust's f32::from_str: decimal or
// inf / infinity / nan, correctly rounded; strtof on a field that contains no hex prefix).
//
// JSON, as serde reads [{"x": [f32, ...], "y": f32}, ...]: other keys are ignored; every
// x must hold the schema's flat size.
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/rsr_amd.h"
#include "isa.h"
#include "runtime_internal.h"

using namespace rsr;

namespace {

bool read_file(const char* path, std::string& out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

int flat_size(const uint8_t* types, int n) {
    int s = 0;
    for (int i = 0; i < n; ++i) s += vt_size(types[i]) ? vt_size(types[i]) : 1;
    return s;
}

// f32::from_str: the whole field, no whitespace, no hex
bool parse_f32(const std::string& s, float& v) {
    if (s.empty() || isspace((unsigned char)s[0]) || isspace((unsigned char)s.back())) return false;
    if (s.find_first_of("xX") != std::string::npos) return false;  // strtof would take hex floats
    const char* b = s.c_str();
    char* e = nullptr;
    errno = 0;
    v = strtof(b, &e);
    return e == b + s.size();
}

// One CSV record starting at pos; false at the end of the input.
bool csv_record(const std::string& t, size_t& pos, std::vector<std::string>& fields) {
    fields.clear();
    while (pos < t.size() && (t[pos] == '\n' || t[pos] == '\r')) ++pos;  // blank lines
    if (pos >= t.size()) return false;
    std::string cur;
    bool quoted = false;
    for (; pos < t.size(); ++pos) {
        const char c = t[pos];
        if (quoted) {
            if (c == '"') {
                if (pos + 1 < t.size() && t[pos + 1] == '"') {
                    cur += '"';
                    ++pos;
                } else {
                    quoted = false;
                }
            } else {
                cur += c;
            }
        } else if (c == '"') {
            quoted = true;
        } else if (c == ',') {
            fields.push_back(cur);
            cur.clear();
        } else if (c == '\n' || c == '\r') {
            break;
        } else {
            cur += c;
        }
    }
    fields.push_back(cur);
    if (pos < t.size() && t[pos] == '\r') ++pos;
    if (pos < t.size() && t[pos] == '\n') ++pos;
    return true;
}

int read_csv(const char* path, int F, int target_col, std::vector<float>& rows, std::vector<float>& y) {
    std::string t;
    if (!path || !read_file(path, t)) return set_error(RSR_E_INVALID_ARG, "cannot read %s", path ? path : "(null)");
    size_t pos = 0;
    std::vector<std::string> f;
    if (!csv_record(t, pos, f)) return RSR_OK;  // no header: no records
    const size_t width = f.size();
    for (int64_t line = 1; csv_record(t, pos, f); ++line) {
        if (f.size() != width)
            return set_error(RSR_E_INVALID_ARG, "record %lld has %zu fields, the header has %zu", (long long)line,
                             f.size(), width);
        const size_t ti = target_col >= 0 ? (size_t)target_col : f.size() - 1;
        if (f.size() < (size_t)F + 1)
            return set_error(RSR_E_INVALID_ARG, "line %lld: not enough columns (expected %d, got %zu)",
                             (long long)line, F + 1, f.size());
        if (ti >= f.size()) return set_error(RSR_E_INVALID_ARG, "target column %zu out of range", ti);
        float v;
        if (!parse_f32(f[ti], v)) return set_error(RSR_E_INVALID_ARG, "line %lld: bad number '%s'", (long long)line, f[ti].c_str());
        y.push_back(v);
        size_t col = 0;
        for (int k = 0; k < F; ++k) {
            if (col == ti) ++col;
            if (col >= f.size()) return set_error(RSR_E_INVALID_ARG, "line %lld: not enough columns", (long long)line);
            if (!parse_f32(f[col], v))
                return set_error(RSR_E_INVALID_ARG, "line %lld: bad number '%s'", (long long)line, f[col].c_str());
            rows.push_back(v);
            ++col;
        }
    }
    return RSR_OK;
}

// A minimal JSON reader for the record array.
struct Json {
    const std::string& t;
    size_t p = 0;
    explicit Json(const std::string& s) : t(s) {}
    void ws() {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\r' || t[p] == '\t')) ++p;
    }
    bool eat(char c) {
        ws();
        if (p < t.size() && t[p] == c) {
            ++p;
            return true;
        }
        return false;
    }
    bool string(std::string& out) {
        ws();
        if (p >= t.size() || t[p] != '"') return false;
        ++p;
        out.clear();
        while (p < t.size() && t[p] != '"') {
            if (t[p] == '\\') {
                if (++p >= t.size()) return false;
                const char e = t[p];
                out += e == 'n' ? '\n' : e == 't' ? '\t' : e;
                if (e == 'u') p += 4;  // \uXXXX in a key: kept as a placeholder character
            } else {
                out += t[p];
            }
            ++p;
        }
        if (p >= t.size()) return false;
        ++p;
        return true;
    }
    // a JSON number (-?int(.frac)?([eE][+-]?digits)?), read as f64 and then narrowed
    // (serde's f32 path)
    bool number(float& v) {
        ws();
        const size_t b = p;
        auto digits = [&]() {
            const size_t s = p;
            while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
            return p > s;
        };
        if (p < t.size() && t[p] == '-') ++p;
        if (!digits()) return false;
        if (p < t.size() && t[p] == '.') {
            ++p;
            if (!digits()) return false;
        }
        if (p < t.size() && (t[p] == 'e' || t[p] == 'E')) {
            ++p;
            if (p < t.size() && (t[p] == '+' || t[p] == '-')) ++p;
            if (!digits()) return false;
        }
        v = (float)strtod(t.substr(b, p - b).c_str(), nullptr);
        return true;
    }
    // skip any value
    bool skip() {
        ws();
        if (p >= t.size()) return false;
        const char c = t[p];
        std::string s;
        float f;
        if (c == '"') return string(s);
        if (c == '{' || c == '[') {
            const char close = c == '{' ? '}' : ']';
            ++p;
            if (eat(close)) return true;
            do {
                if (c == '{' && !(string(s) && eat(':'))) return false;
                if (!skip()) return false;
            } while (eat(','));
            return eat(close);
        }
        if (t.compare(p, 4, "true") == 0 || t.compare(p, 4, "null") == 0) {
            p += 4;
            return true;
        }
        if (t.compare(p, 5, "false") == 0) {
            p += 5;
            return true;
        }
        return number(f);
    }
};

int read_json(const char* path, int F, std::vector<float>& rows, std::vector<float>& y) {
    std::string t;
    if (!path || !read_file(path, t)) return set_error(RSR_E_INVALID_ARG, "cannot read %s", path ? path : "(null)");
    Json j(t);
    if (!j.eat('[')) return set_error(RSR_E_INVALID_ARG, "expected a JSON array of records");
    if (j.eat(']')) return RSR_OK;
    int64_t rec = 0;
    do {
        if (!j.eat('{')) return set_error(RSR_E_INVALID_ARG, "record %lld: expected an object", (long long)rec);
        bool has_x = false, has_y = false;
        std::vector<float> x;
        float yv = 0.0f;
        if (!j.eat('}')) {
            do {
                std::string key;
                if (!(j.string(key) && j.eat(':'))) return set_error(RSR_E_INVALID_ARG, "record %lld: bad key", (long long)rec);
                if (key == "x") {
                    if (!j.eat('[')) return set_error(RSR_E_INVALID_ARG, "record %lld: x is not an array", (long long)rec);
                    x.clear();
                    if (!j.eat(']')) {
                        do {
                            float v;
                            if (!j.number(v)) return set_error(RSR_E_INVALID_ARG, "record %lld: bad number in x", (long long)rec);
                            x.push_back(v);
                        } while (j.eat(','));
                        if (!j.eat(']')) return set_error(RSR_E_INVALID_ARG, "record %lld: unterminated x", (long long)rec);
                    }
                    has_x = true;
                } else if (key == "y") {
                    if (!j.number(yv)) return set_error(RSR_E_INVALID_ARG, "record %lld: bad y", (long long)rec);
                    has_y = true;
                } else if (!j.skip()) {
                    return set_error(RSR_E_INVALID_ARG, "record %lld: bad value of '%s'", (long long)rec, key.c_str());
                }
            } while (j.eat(','));
            if (!j.eat('}')) return set_error(RSR_E_INVALID_ARG, "record %lld: unterminated object", (long long)rec);
        }
        if (!has_x || !has_y) return set_error(RSR_E_INVALID_ARG, "record %lld: missing x or y", (long long)rec);
        if ((int)x.size() != F)
            return set_error(RSR_E_INVALID_ARG, "record %lld: x has %zu values, the schema %d", (long long)rec, x.size(), F);
        rows.insert(rows.end(), x.begin(), x.end());
        y.push_back(yv);
        ++rec;
    } while (j.eat(','));
    if (!j.eat(']')) return set_error(RSR_E_INVALID_ARG, "unterminated record array");
    j.ws();
    if (j.p != t.size()) return set_error(RSR_E_INVALID_ARG, "trailing characters after the record array");
    return RSR_OK;
}

int copy_out(const std::vector<float>& rows, const std::vector<float>& y, float* rows_out, float* y_out,
             int64_t capacity, int64_t* n_out) {
    *n_out = (int64_t)y.size();
    if (!rows_out && !y_out) return RSR_OK;  // size query
    if ((int64_t)y.size() > capacity) return set_error(RSR_E_INVALID_ARG, "%zu records exceed the capacity", y.size());
    if (rows_out && !rows.empty()) memcpy(rows_out, rows.data(), rows.size() * sizeof(float));
    if (y_out && !y.empty()) memcpy(y_out, y.data(), y.size() * sizeof(float));
    return RSR_OK;
}

}  // namespace

extern "C" int rsr_read_csv(const char* path, const uint8_t* feature_types, int n_feature_types, int target_col,
                            float* rows_out, float* y_out, int64_t capacity, int64_t* n_out) {
    if (!n_out || (n_feature_types > 0 && !feature_types)) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::vector<float> rows, y;
    const int rc = read_csv(path, flat_size(feature_types, n_feature_types), target_col, rows, y);
    return rc ? rc : copy_out(rows, y, rows_out, y_out, capacity, n_out);
}

extern "C" int rsr_read_json(const char* path, const uint8_t* feature_types, int n_feature_types, float* rows_out,
                             float* y_out, int64_t capacity, int64_t* n_out) {
    if (!n_out || (n_feature_types > 0 && !feature_types)) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::vector<float> rows, y;
    const int rc = read_json(path, flat_size(feature_types, n_feature_types), rows, y);
    return rc ? rc : copy_out(rows, y, rows_out, y_out, capacity, n_out);
}

extern "C" int rsr_dataset_from_file(rsr_context* ctx, const char* path, int format, const uint8_t* feature_types,
                                     int n_feature_types, int target_col, int normalize, rsr_dataset** out) {
    if (!ctx || !out || (n_feature_types > 0 && !feature_types)) return set_error(RSR_E_INVALID_ARG, "null argument");
    std::vector<float> rows, y;
    const int F = flat_size(feature_types, n_feature_types);
    int rc = format == RSR_FILE_JSON ? read_json(path, F, rows, y)
             : format == RSR_FILE_CSV ? read_csv(path, F, target_col, rows, y)
                                      : set_error(RSR_E_INVALID_ARG, "unknown file format %d", format);
    if (rc) return rc;
    const float zero = 0.0f;
    return rsr_dataset_create(ctx, rows.empty() ? &zero : rows.data(), y.empty() ? &zero : y.data(),
                              (int64_t)y.size(), feature_types, n_feature_types, normalize, out);
}
