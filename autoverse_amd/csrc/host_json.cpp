// host_json.cpp -- SimulationConfig and Seed from the reference's JSON files, in C++.
//
// The reference reads both with serde_json (main.rs:409-436: the config from its path, the seed
// from the same path with the extension "seed.json", Seed::default() when that file is missing).
// This file restates the parts of serde's derived Deserialize that those types use:
//   * SimulationConfig (config.rs:14-35): every field required except depth (default 1) and
//     embedding (default: disabled, temperature 1, softmax); KernelConfig, RingConfig and
//     FlowConfig (config.rs:55-110) have no defaults;
//   * Seed (seed.rs:13-32) holds a Pattern (seed.rs:36-135), internally tagged by "type";
//     tuples are arrays of exactly their length; Noise's channel is an Option (missing or null
//     means every channel);
//   * numbers: f32 fields take any JSON number (parsed to f64, then rounded to f32); usize and
//     u64 fields take only non-negative integers; unknown fields are ignored.
// Errors name the field and its place, and carry HP_ERR_PARSE.
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/flowlenia_hip.h"
#include "host_internal.hpp"

namespace {

thread_local std::string g_json_error;

struct JsonError {
    std::string msg;
};

// ---- a small JSON value tree and its parser (RFC 8259) -------------------------------------
struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    bool integer = false;  // the literal had no fraction and no exponent
    bool negative = false;
    unsigned long long uval = 0;  // exact value of a non-negative integer literal
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    const Value* get(const char* k) const {
        const Value* found = nullptr;
        for (const auto& kv : obj)
            if (kv.first == k) found = &kv.second;  // the last duplicate wins, as serde_json's map
        return found;
    }
};

struct Parser {
    const char* s;
    size_t n, i = 0;
    int depth = 0;

    [[noreturn]] void fail(const std::string& m) {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < i && k < n; k++) {
            if (s[k] == '\n') {
                line++;
                col = 1;
            } else {
                col++;
            }
        }
        throw JsonError{m + " at line " + std::to_string(line) + " column " + std::to_string(col)};
    }
    void ws() {
        while (i < n && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
    }
    bool lit(const char* w) {
        const size_t l = strlen(w);
        if (i + l <= n && memcmp(s + i, w, l) == 0) {
            i += l;
            return true;
        }
        return false;
    }
    static void utf8(std::string& o, unsigned cp) {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18));
            o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (i + 4 > n) fail("EOF while parsing a string");
        unsigned v = 0;
        for (int k = 0; k < 4; k++) {
            const char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("invalid escape");
        }
        return v;
    }
    std::string string() {
        i++;  // opening quote
        std::string o;
        for (;;) {
            if (i >= n) fail("EOF while parsing a string");
            const unsigned char c = (unsigned char)s[i++];
            if (c == '"') return o;
            if (c < 0x20) fail("control character in string");
            if (c != '\\') {
                o += (char)c;
                continue;
            }
            if (i >= n) fail("EOF while parsing a string");
            const char e = s[i++];
            switch (e) {
            case '"': o += '"'; break;
            case '\\': o += '\\'; break;
            case '/': o += '/'; break;
            case 'b': o += '\b'; break;
            case 'f': o += '\f'; break;
            case 'n': o += '\n'; break;
            case 'r': o += '\r'; break;
            case 't': o += '\t'; break;
            case 'u': {
                unsigned cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && i + 1 < n && s[i] == '\\' && s[i + 1] == 'u') {
                    i += 2;
                    const unsigned lo = hex4();
                    if (lo < 0xDC00 || lo > 0xDFFF) fail("invalid surrogate pair");
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(o, cp);
                break;
            }
            default: fail("invalid escape");
            }
        }
    }
    Value number() {
        const size_t st = i;
        Value v;
        v.kind = Value::Number;
        if (s[i] == '-') {
            v.negative = true;
            i++;
        }
        if (i >= n || !(s[i] >= '0' && s[i] <= '9')) fail("invalid number");
        if (s[i] == '0') i++;
        else
            while (i < n && s[i] >= '0' && s[i] <= '9') i++;
        bool integer = true;
        if (i < n && s[i] == '.') {
            integer = false;
            i++;
            if (i >= n || !(s[i] >= '0' && s[i] <= '9')) fail("invalid number");
            while (i < n && s[i] >= '0' && s[i] <= '9') i++;
        }
        if (i < n && (s[i] == 'e' || s[i] == 'E')) {
            integer = false;
            i++;
            if (i < n && (s[i] == '+' || s[i] == '-')) i++;
            if (i >= n || !(s[i] >= '0' && s[i] <= '9')) fail("invalid number");
            while (i < n && s[i] >= '0' && s[i] <= '9') i++;
        }
        const std::string t(s + st, i - st);
        v.num = strtod(t.c_str(), nullptr);
        v.integer = integer;
        if (integer && !v.negative) {
            errno = 0;
            v.uval = strtoull(t.c_str(), nullptr, 10);
            if (errno == ERANGE) v.integer = false;  // beyond u64: only a float fits
        }
        return v;
    }
    Value value() {
        ws();
        if (i >= n) fail("EOF while parsing a value");
        if (++depth > 128) fail("recursion limit exceeded");
        Value v;
        const char c = s[i];
        if (c == '{') {
            v.kind = Value::Object;
            i++;
            ws();
            if (i < n && s[i] == '}') {
                i++;
            } else {
                for (;;) {
                    ws();
                    if (i >= n) fail("EOF while parsing an object");
                    if (s[i] != '"') fail("key must be a string");
                    std::string k = string();
                    ws();
                    if (i >= n || s[i] != ':') fail("expected `:`");
                    i++;
                    v.obj.emplace_back(std::move(k), value());
                    ws();
                    if (i >= n) fail("EOF while parsing an object");
                    if (s[i] == ',') {
                        i++;
                        continue;
                    }
                    if (s[i] == '}') {
                        i++;
                        break;
                    }
                    fail("expected `,` or `}`");
                }
            }
        } else if (c == '[') {
            v.kind = Value::Array;
            i++;
            ws();
            if (i < n && s[i] == ']') {
                i++;
            } else {
                for (;;) {
                    v.arr.push_back(value());
                    ws();
                    if (i >= n) fail("EOF while parsing a list");
                    if (s[i] == ',') {
                        i++;
                        continue;
                    }
                    if (s[i] == ']') {
                        i++;
                        break;
                    }
                    fail("expected `,` or `]`");
                }
            }
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string();
        } else if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Value::Bool;
        } else if (lit("null")) {
            v.kind = Value::Null;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            v = number();
        } else {
            fail("expected value");
        }
        depth--;
        return v;
    }
    Value document() {
        Value v = value();
        ws();
        if (i != n) fail("trailing characters");
        return v;
    }
};

// ---- typed reads with serde's rules --------------------------------------------------------
[[noreturn]] void schema_fail(const std::string& where, const std::string& m) {
    throw JsonError{where.empty() ? m : where + ": " + m};
}

const Value& field(const Value& o, const char* k, const std::string& where) {
    if (o.kind != Value::Object) schema_fail(where, "invalid type: expected a map");
    const Value* v = o.get(k);
    if (!v) schema_fail(where, std::string("missing field `") + k + "`");
    return *v;
}

std::string at(const std::string& where, const char* k) { return where.empty() ? k : where + "." + k; }

float f32_of(const Value& v, const std::string& where) {
    if (v.kind != Value::Number) schema_fail(where, "invalid type: expected f32");
    return (float)v.num;
}

unsigned long long uint_of(const Value& v, const std::string& where, unsigned long long max) {
    if (v.kind != Value::Number) schema_fail(where, "invalid type: expected an unsigned integer");
    if (!v.integer) schema_fail(where, "invalid type: floating point, expected an unsigned integer");
    if (v.negative) schema_fail(where, "invalid value: negative integer, expected an unsigned integer");
    if (v.uval > max) schema_fail(where, "integer out of range");
    return v.uval;
}

float get_f32(const Value& o, const char* k, const std::string& where) {
    return f32_of(field(o, k, where), at(where, k));
}
uint32_t get_u32(const Value& o, const char* k, const std::string& where) {
    return (uint32_t)uint_of(field(o, k, where), at(where, k), 0xFFFFFFFFull);
}

const std::vector<Value>& get_array(const Value& o, const char* k, const std::string& where) {
    const Value& v = field(o, k, where);
    if (v.kind != Value::Array) schema_fail(at(where, k), "invalid type: expected a sequence");
    return v.arr;
}

// a tuple of `len` f32 values (serde's (f32, f32[, f32]): an array of exactly that length)
void get_tuple(const Value& o, const char* k, const std::string& where, float* out, size_t len) {
    const std::vector<Value>& a = get_array(o, k, where);
    if (a.size() != len)
        schema_fail(at(where, k), "invalid length " + std::to_string(a.size()) + ", expected a tuple of size " +
                                      std::to_string(len));
    for (size_t j = 0; j < len; j++) out[j] = f32_of(a[j], at(where, k) + "[" + std::to_string(j) + "]");
}

std::string read_file(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) throw JsonError{std::string("cannot read ") + path};
    std::string t;
    char buf[65536];
    size_t r;
    while ((r = fread(buf, 1, sizeof(buf), f)) > 0) t.append(buf, r);
    fclose(f);
    return t;
}

}  // namespace

// Owned SimulationConfig: the hp_config points into the document's arrays.
struct hp_config_doc {
    hp_config cfg{};
    std::vector<hp_kernel> kernels;
    std::vector<hp_ring> rings;
    hp_embedding embedding{0, 1.0f, 0};
};

// Owned Seed.
struct hp_seed_doc {
    hp_seed seed{};
    std::vector<hp_blob> blobs;
    std::vector<hp_custom_value> values;
};

namespace {

void parse_config(const Value& root, hp_config_doc* d) {  // config.rs:14-120
    const std::string w;
    if (root.kind != Value::Object) schema_fail(w, "invalid type: expected a SimulationConfig map");
    hp_config& c = d->cfg;
    c.width = get_u32(root, "width", w);
    c.height = get_u32(root, "height", w);
    c.depth = root.get("depth") ? get_u32(root, "depth", w) : 1;
    c.channels = get_u32(root, "channels", w);
    c.dt = get_f32(root, "dt", w);
    c.kernel_radius = get_u32(root, "kernel_radius", w);
    const std::vector<Value>& ks = get_array(root, "kernels", w);
    for (size_t k = 0; k < ks.size(); k++) {
        const std::string kw = "kernels[" + std::to_string(k) + "]";
        hp_kernel kc{};
        kc.radius = get_f32(ks[k], "radius", kw);
        const std::vector<Value>& rs = get_array(ks[k], "rings", kw);
        kc.ring_offset = (uint32_t)d->rings.size();
        kc.ring_count = (uint32_t)rs.size();
        for (size_t r = 0; r < rs.size(); r++) {
            const std::string rw = kw + ".rings[" + std::to_string(r) + "]";
            d->rings.push_back(hp_ring{get_f32(rs[r], "amplitude", rw), get_f32(rs[r], "distance", rw),
                                       get_f32(rs[r], "width", rw)});
        }
        kc.weight = get_f32(ks[k], "weight", kw);
        kc.mu = get_f32(ks[k], "mu", kw);
        kc.sigma = get_f32(ks[k], "sigma", kw);
        kc.source_channel = get_u32(ks[k], "source_channel", kw);
        kc.target_channel = get_u32(ks[k], "target_channel", kw);
        d->kernels.push_back(kc);
    }
    const Value& fl = field(root, "flow", w);
    c.flow.beta_a = get_f32(fl, "beta_a", "flow");
    c.flow.n = get_f32(fl, "n", "flow");
    c.flow.distribution_size = get_f32(fl, "distribution_size", "flow");
    if (const Value* e = root.get("embedding")) {  // EmbeddingConfig (embedding.rs:326-345)
        if (e->kind != Value::Object) schema_fail("embedding", "invalid type: expected a map");
        const Value* en = e->get("enabled");
        if (!en) schema_fail("embedding", "missing field `enabled`");
        if (en->kind != Value::Bool) schema_fail("embedding.enabled", "invalid type: expected a boolean");
        d->embedding.enabled = en->b ? 1 : 0;
        d->embedding.mixing_temperature = get_f32(*e, "mixing_temperature", "embedding");
        const Value* lm = e->get("linear_mixing");
        if (!lm) schema_fail("embedding", "missing field `linear_mixing`");
        if (lm->kind != Value::Bool) schema_fail("embedding.linear_mixing", "invalid type: expected a boolean");
        d->embedding.linear_mixing = lm->b ? 1 : 0;
    }
    c.num_kernels = (uint32_t)d->kernels.size();
    c.kernels = d->kernels.data();
    c.num_rings = (uint32_t)d->rings.size();
    c.rings = d->rings.data();
}

void parse_seed(const Value& root, hp_seed_doc* d) {  // seed.rs:13-152
    if (root.kind != Value::Object) schema_fail("", "invalid type: expected a Seed map");
    const Value& p = field(root, "pattern", "");
    if (p.kind != Value::Object) schema_fail("pattern", "invalid type: expected a map");
    const Value* tv = p.get("type");
    if (!tv) schema_fail("pattern", "missing field `type`");
    if (tv->kind != Value::String) schema_fail("pattern.type", "invalid type: expected a string");
    const std::string& t = tv->str;
    const std::string w = "pattern";
    hp_seed& s = d->seed;
    s.channel = 0;
    auto blob = [&](const Value& b, const std::string& bw, int dim) {
        hp_blob o{};
        get_tuple(b, "center", bw, o.center, (size_t)dim);
        o.radius = get_f32(b, "radius", bw);
        o.amplitude = get_f32(b, "amplitude", bw);
        o.channel = get_u32(b, "channel", bw);
        return o;
    };
    if (t == "GaussianBlob" || t == "GaussianSphere") {
        const int dim = t == "GaussianBlob" ? 2 : 3;
        s.kind = dim == 2 ? HP_PATTERN_GAUSSIAN_BLOB : HP_PATTERN_GAUSSIAN_SPHERE;
        get_tuple(p, "center", w, s.center, (size_t)dim);
        s.radius = get_f32(p, "radius", w);
        s.amplitude = get_f32(p, "amplitude", w);
        s.channel = (int32_t)get_u32(p, "channel", w);
    } else if (t == "MultiBlob" || t == "MultiSphere") {
        const bool sph = t == "MultiSphere";
        s.kind = sph ? HP_PATTERN_MULTI_SPHERE : HP_PATTERN_MULTI_BLOB;
        const char* key = sph ? "spheres" : "blobs";
        const std::vector<Value>& a = get_array(p, key, w);
        for (size_t k = 0; k < a.size(); k++)
            d->blobs.push_back(blob(a[k], w + "." + key + "[" + std::to_string(k) + "]", sph ? 3 : 2));
    } else if (t == "Ring" || t == "Shell") {
        const int dim = t == "Ring" ? 2 : 3;
        s.kind = dim == 2 ? HP_PATTERN_RING : HP_PATTERN_SHELL;
        get_tuple(p, "center", w, s.center, (size_t)dim);
        s.inner_radius = get_f32(p, "inner_radius", w);
        s.outer_radius = get_f32(p, "outer_radius", w);
        s.amplitude = get_f32(p, "amplitude", w);
        s.channel = (int32_t)get_u32(p, "channel", w);
    } else if (t == "Torus3D") {
        s.kind = HP_PATTERN_TORUS3D;
        get_tuple(p, "center", w, s.center, 3);
        s.major_radius = get_f32(p, "major_radius", w);
        s.minor_radius = get_f32(p, "minor_radius", w);
        s.amplitude = get_f32(p, "amplitude", w);
        s.channel = (int32_t)get_u32(p, "channel", w);
    } else if (t == "Noise") {
        s.kind = HP_PATTERN_NOISE;
        s.amplitude = get_f32(p, "amplitude", w);
        const Value* ch = p.get("channel");  // Option<usize>: missing or null means every channel
        s.channel = (!ch || ch->kind == Value::Null) ? -1 : (int32_t)uint_of(*ch, "pattern.channel", 0x7FFFFFFFull);
        s.noise_seed = (uint64_t)uint_of(field(p, "seed", w), "pattern.seed", ~0ull);
    } else if (t == "Custom" || t == "Custom3D") {
        const bool d3 = t == "Custom3D";
        s.kind = d3 ? HP_PATTERN_CUSTOM3D : HP_PATTERN_CUSTOM;
        const std::vector<Value>& a = get_array(p, "values", w);
        const size_t len = d3 ? 5 : 4;
        for (size_t k = 0; k < a.size(); k++) {
            const std::string vw = w + ".values[" + std::to_string(k) + "]";
            if (a[k].kind != Value::Array || a[k].arr.size() != len)
                schema_fail(vw, "invalid length, expected a tuple of size " + std::to_string(len));
            const std::vector<Value>& e = a[k].arr;
            hp_custom_value o{};
            o.x = (uint32_t)uint_of(e[0], vw + "[0]", 0xFFFFFFFFull);
            o.y = (uint32_t)uint_of(e[1], vw + "[1]", 0xFFFFFFFFull);
            o.z = d3 ? (uint32_t)uint_of(e[2], vw + "[2]", 0xFFFFFFFFull) : 0;
            o.channel = (uint32_t)uint_of(e[d3 ? 3 : 2], vw + "[" + std::to_string(d3 ? 3 : 2) + "]", 0xFFFFFFFFull);
            o.value = f32_of(e[d3 ? 4 : 3], vw + "[" + std::to_string(d3 ? 4 : 3) + "]");
            d->values.push_back(o);
        }
    } else {
        schema_fail("pattern.type", "unknown variant `" + t + "`");
    }
    s.num_blobs = (uint32_t)d->blobs.size();
    s.blobs = d->blobs.data();
    s.num_values = (uint32_t)d->values.size();
    s.values = d->values.data();
}

hp_seed_doc* default_seed() {  // Seed::default (seed.rs:18-29)
    hp_seed_doc* d = new hp_seed_doc();
    d->seed.kind = HP_PATTERN_GAUSSIAN_BLOB;
    d->seed.center[0] = 0.5f;
    d->seed.center[1] = 0.5f;
    d->seed.radius = 0.1f;
    d->seed.amplitude = 1.0f;
    d->seed.channel = 0;
    return d;
}

// Path::with_extension("seed.json"): the file name's last extension replaced (a name whose only
// dot leads it has no extension).
std::string seed_path_of(const std::string& p) {
    const size_t slash = p.find_last_of('/');
    const size_t name = slash == std::string::npos ? 0 : slash + 1;
    const size_t dot = p.find_last_of('.');
    if (dot != std::string::npos && dot > name) return p.substr(0, dot) + ".seed.json";
    return p + ".seed.json";
}

template <class Doc, class F>
hp_status parse_doc(const std::string& text, Doc** out, F&& fill) {
    Doc* d = new Doc();
    try {
        Parser ps{text.data(), text.size()};
        const Value v = ps.document();
        fill(v, d);
    } catch (const JsonError& e) {
        delete d;
        g_json_error = e.msg;
        return HP_ERR_PARSE;
    } catch (const std::exception& e) {
        delete d;
        g_json_error = e.what();
        return HP_ERR_PARSE;
    }
    *out = d;
    return HP_OK;
}

}  // namespace

extern "C" {

const char* hp_json_last_error(void) { return g_json_error.c_str(); }

hp_status hp_config_from_json(const char* text, hp_config_doc** out) {
    if (!out) return HP_ERR_NULL;
    *out = nullptr;
    if (!text) return HP_ERR_NULL;
    return parse_doc(std::string(text), out, parse_config);
}

hp_status hp_seed_from_json(const char* text, hp_seed_doc** out) {
    if (!out) return HP_ERR_NULL;
    *out = nullptr;
    if (!text) return HP_ERR_NULL;
    return parse_doc(std::string(text), out, parse_seed);
}

const hp_config* hp_config_doc_config(const hp_config_doc* d) { return d ? &d->cfg : nullptr; }
hp_status hp_config_doc_embedding(const hp_config_doc* d, hp_embedding* out) {
    if (!d || !out) return HP_ERR_NULL;
    *out = d->embedding;
    return HP_OK;
}
void hp_config_doc_free(hp_config_doc* d) { delete d; }
const hp_seed* hp_seed_doc_seed(const hp_seed_doc* d) { return d ? &d->seed : nullptr; }
void hp_seed_doc_free(hp_seed_doc* d) { delete d; }

hp_status hp_load_config_and_seed(const char* config_path, hp_config_doc** cfg, hp_seed_doc** seed) {
    if (!config_path || !cfg || !seed) return HP_ERR_NULL;
    *cfg = nullptr;
    *seed = nullptr;
    std::string text;
    try {
        text = read_file(config_path);
    } catch (const JsonError& e) {
        g_json_error = "Error reading config file: " + e.msg;
        return HP_ERR_IO;
    }
    hp_status st = hp_config_from_json(text.c_str(), cfg);
    if (st != HP_OK) {
        g_json_error = "Error parsing config: " + g_json_error;
        return st;
    }
    const std::string sp = seed_path_of(config_path);
    FILE* f = fopen(sp.c_str(), "rb");
    if (!f) {  // no seed file: Seed::default()
        *seed = default_seed();
        return HP_OK;
    }
    fclose(f);
    try {
        text = read_file(sp.c_str());
    } catch (const JsonError& e) {
        hp_config_doc_free(*cfg);
        *cfg = nullptr;
        g_json_error = "Error reading seed file: " + e.msg;
        return HP_ERR_IO;
    }
    st = hp_seed_from_json(text.c_str(), seed);
    if (st != HP_OK) {
        hp_config_doc_free(*cfg);
        *cfg = nullptr;
        g_json_error = "Error parsing seed: " + g_json_error;
    }
    return st;
}

}  // extern "C"
