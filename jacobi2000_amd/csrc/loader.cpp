// Checkpoint loading for libj2k (host only).
#include "loader.h"

#include <cmath>
#include <cstring>
#include <limits>

namespace j2k {

namespace {

struct Parser {
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void bad(const char* what) {
        fail(J2K_ERR_BAD_CONFIG, std::string("invalid JSON: ") + what + " at byte " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) i++;
    }
    bool lit(const char* w) {
        size_t n = std::strlen(w);
        if (s.compare(i, n, w) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    std::string string() {
        if (s[i] != '"') bad("expected string");
        i++;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c != '\\') {
                out.push_back(c);
                continue;
            }
            if (i >= s.size()) bad("truncated escape");
            char e = s[i++];
            switch (e) {
                case '"': out.push_back('"'); break;
                case '\\': out.push_back('\\'); break;
                case '/': out.push_back('/'); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'n': out.push_back('\n'); break;
                case 'r': out.push_back('\r'); break;
                case 't': out.push_back('\t'); break;
                case 'u': {
                    if (i + 4 > s.size()) bad("truncated \\u escape");
                    unsigned cp = std::stoul(s.substr(i, 4), nullptr, 16);
                    i += 4;
                    if (cp < 0x80) {
                        out.push_back((char)cp);
                    } else if (cp < 0x800) {
                        out.push_back((char)(0xc0 | (cp >> 6)));
                        out.push_back((char)(0x80 | (cp & 0x3f)));
                    } else {
                        out.push_back((char)(0xe0 | (cp >> 12)));
                        out.push_back((char)(0x80 | ((cp >> 6) & 0x3f)));
                        out.push_back((char)(0x80 | (cp & 0x3f)));
                    }
                    break;
                }
                default: bad("unknown escape");
            }
        }
        if (i >= s.size()) bad("unterminated string");
        i++;
        return out;
    }
    Json value() {
        ws();
        if (i >= s.size()) bad("unexpected end");
        Json v;
        char c = s[i];
        if (c == '{') {
            v.kind = Json::Obj;
            i++;
            ws();
            if (i < s.size() && s[i] == '}') {
                i++;
                return v;
            }
            for (;;) {
                ws();
                std::string k = string();
                ws();
                if (i >= s.size() || s[i] != ':') bad("expected ':'");
                i++;
                v.obj[k] = value();
                ws();
                if (i < s.size() && s[i] == ',') {
                    i++;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    i++;
                    return v;
                }
                bad("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Json::Arr;
            i++;
            ws();
            if (i < s.size() && s[i] == ']') {
                i++;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') {
                    i++;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    i++;
                    return v;
                }
                bad("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Json::Str;
            v.str = string();
            return v;
        }
        if (lit("true")) {
            v.kind = Json::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = Json::Bool;
            return v;
        }
        if (lit("null")) return v;
        size_t st = i;
        while (i < s.size() && (std::strchr("+-0123456789.eE", s[i]) != nullptr)) i++;
        if (st == i) bad("unexpected character");
        v.kind = Json::Num;
        try {
            v.num = std::stod(s.substr(st, i - st));
        } catch (...) {
            bad("bad number");
        }
        return v;
    }
};

const Json& need(const Json& j, const std::string& key) {
    const Json* v = j.find(key);
    if (!v) fail(J2K_ERR_BAD_CONFIG, "config.json: missing field `" + key + "`");
    return *v;
}

double num(const Json& j, const std::string& key) {
    const Json& v = need(j, key);
    if (v.kind != Json::Num) fail(J2K_ERR_BAD_CONFIG, "config.json: field `" + key + "` is not a number");
    return v.num;
}

int uint_field(const Json& j, const std::string& key) {
    double d = num(j, key);
    if (d < 0 || d != std::floor(d) || d > 1e9) fail(J2K_ERR_BAD_CONFIG, "config.json: field `" + key + "` is not a size");
    return (int)d;
}

}  // namespace

Json Json::parse(const std::string& text) {
    Parser p{text};
    Json v = p.value();
    p.ws();
    if (p.i != text.size()) p.bad("trailing characters");
    return v;
}

const Json* Json::find(const std::string& key) const {
    if (kind != Obj) return nullptr;
    auto it = obj.find(key);
    return it == obj.end() ? nullptr : &it->second;
}

float e4m3_decode(uint8_t code) {
    const int sgn = code >> 7, e = (code >> 3) & 15, mant = code & 7;
    if (e == 15 && mant == 7) return std::numeric_limits<float>::quiet_NaN();
    const float mag = e == 0 ? std::ldexp((float)mant, -9) : std::ldexp(1.0f + mant / 8.0f, e - 7);
    return sgn ? -mag : mag;
}

uint8_t e4m3_encode(float v) {
    if (std::isnan(v)) return 0x7f;
    const uint8_t sgn = std::signbit(v) ? 0x80 : 0;
    double a = std::fabs((double)v);
    if (a >= 448.0) return sgn | 0x7e;  // saturate
    // the nearest of the 127 non-negative finite codes, ties to the even code
    int lo = 0, hi = 0x7e;
    while (lo < hi) {  // largest code with value <= a
        const int mid = (lo + hi + 1) / 2;
        if ((double)e4m3_decode((uint8_t)mid) <= a) lo = mid;
        else hi = mid - 1;
    }
    if (lo < 0x7e) {
        const double below = e4m3_decode((uint8_t)lo), above = e4m3_decode((uint8_t)(lo + 1));
        if (a - below > above - a || (a - below == above - a && (lo & 1))) lo++;
    }
    return sgn | (uint8_t)lo;
}

Weights read_safetensors(const uint8_t* bytes, size_t n) {
    if (!bytes || n < 8) fail(J2K_ERR_SHAPE, "safetensors: buffer shorter than its 8-byte header length");
    uint64_t hl = 0;
    for (int b = 7; b >= 0; b--) hl = (hl << 8) | bytes[b];
    if (hl > n - 8) fail(J2K_ERR_SHAPE, "safetensors: header length past the end of the buffer");
    Json h = Json::parse(std::string((const char*)bytes + 8, (size_t)hl));
    if (h.kind != Json::Obj) fail(J2K_ERR_SHAPE, "safetensors: header is not a JSON object");
    size_t data_start = 8 + (size_t)hl;
    Weights out;
    for (auto& [name, v] : h.obj) {
        if (name == "__metadata__") continue;
        const Json* dt = v.find("dtype");
        if (!dt || dt->kind != Json::Str) fail(J2K_ERR_SHAPE, "safetensors: tensor `" + name + "` has no dtype");
        const std::string& dtype = dt->str;
        const size_t esize = dtype == "F32" ? 4 : dtype == "BF16" ? 2 : dtype == "F8_E4M3" ? 1 : 0;
        if (!esize)
            fail(J2K_ERR_BAD_DTYPE, "tensor `" + name + "`: dtype " + dtype + " (F32, BF16 or F8_E4M3 with a scale)");
        const Json* shp = v.find("shape");
        const Json* off = v.find("data_offsets");
        if (!shp || shp->kind != Json::Arr || !off || off->kind != Json::Arr || off->arr.size() != 2)
            fail(J2K_ERR_SHAPE, "safetensors: tensor `" + name + "` lacks shape/data_offsets");
        // Offsets and dimensions must be non-negative integers inside the
        // buffer (weights.rs reads them as u64 and bounds-checks the slice);
        // every product and sum below is checked before it can wrap.
        const size_t avail = n - data_start;
        auto count = [&](const Json& j, const char* what) -> uint64_t {
            if (j.kind != Json::Num || !(j.num >= 0.0) || j.num != std::floor(j.num) || j.num > 9007199254740992.0)
                fail(J2K_ERR_SHAPE, "safetensors: tensor `" + name + "` has a " + what + " that is not a non-negative integer");
            return (uint64_t)j.num;
        };
        Tensor t;
        uint64_t elems = 1;
        for (auto& d : shp->arr) {
            const uint64_t dim = count(d, "dimension");
            if (dim != 0 && elems > (avail / esize) / dim)
                fail(J2K_ERR_SHAPE, "safetensors: tensor `" + name + "` is larger than the buffer");
            elems *= dim;
            t.shape.push_back((int64_t)dim);
        }
        const uint64_t a = count(off->arr[0], "data offset"), b = count(off->arr[1], "data offset");
        if (a > b || b > avail || b - a != elems * esize)
            fail(J2K_ERR_SHAPE, "safetensors: tensor `" + name + "` data range does not match its shape");
        const uint8_t* src = bytes + data_start + a;
        t.data.resize((size_t)t.numel());
        if (esize == 4) {
            std::memcpy(t.data.data(), src, b - a);  // little-endian host
        } else if (esize == 2) {
            for (size_t i = 0; i < t.data.size(); i++) {
                const uint32_t u = (uint32_t)src[2 * i] | ((uint32_t)src[2 * i + 1] << 8);
                const uint32_t w = u << 16;
                std::memcpy(&t.data[i], &w, 4);
            }
        } else {
            t.f8.assign(src, src + t.data.size());
        }
        out.emplace(name, std::move(t));
    }
    // F8 tensors take their scale from `<name>.scale`
    for (auto& [name, t] : out) {
        if (t.f8.empty()) continue;
        auto it = out.find(name + ".scale");
        if (it == out.end() || it->second.data.size() != 1 || !it->second.f8.empty())
            fail(J2K_ERR_MISSING_TENSOR, "F8_E4M3 tensor `" + name + "` needs an F32 scale `" + name + ".scale`");
        t.f8_scale = it->second.data[0];
        for (size_t i = 0; i < t.f8.size(); i++) t.data[i] = e4m3_decode(t.f8[i]) * t.f8_scale;
    }
    return out;
}

const Tensor& get(const Weights& w, const std::string& name) {
    auto it = w.find(name);
    if (it == w.end()) fail(J2K_ERR_MISSING_TENSOR, "missing tensor `" + name + "`");
    return it->second;
}

j2k_config parse_config(const std::string& json_text, const Weights& w, int precision) {
    Json j = Json::parse(json_text);
    if (j.kind != Json::Obj) fail(J2K_ERR_BAD_CONFIG, "config.json: not an object");
    j2k_config c{};
    c.precision = precision;
    const Json* form = j.find("form");
    if (!form || (form->kind == Json::Str && form->str == "patch2d")) {
        c.form = J2K_FORM_PATCH2D;
    } else if (form->kind == Json::Str && form->str == "cell_st") {
        c.form = J2K_FORM_CELL_ST;
    } else {
        fail(J2K_ERR_BAD_CONFIG, "config.json: unknown `form` (want patch2d | cell_st)");
    }
    c.channels = uint_field(j, "channels");
    c.dim = uint_field(j, "dim");
    c.layers = uint_field(j, "layers");
    c.heads = uint_field(j, "heads");
    c.meta_dim = uint_field(j, "meta_dim");
    const Json& nrm = need(j, "norm");
    const Json* mean = nrm.find("mean");
    const Json* stdv = nrm.find("std");
    const Json* depth = nrm.find("depth");
    if (!mean || !stdv || !depth || mean->kind != Json::Arr || stdv->kind != Json::Arr || mean->arr.size() != 3 ||
        stdv->arr.size() != 3 || depth->kind != Json::Num)
        fail(J2K_ERR_BAD_CONFIG, "config.json: `norm` needs mean[3], std[3], depth");
    for (int i = 0; i < 3; i++) {
        c.norm_mean[i] = (float)mean->arr[i].num;
        c.norm_std[i] = (float)stdv->arr[i].num;
    }
    c.norm_depth = (float)depth->num;
    c.dx = (float)num(j, "dx");
    c.dt = (float)num(j, "dt");
    c.g = (float)num(j, "g");
    c.depth = (float)num(j, "depth");
    c.nu = (float)num(j, "nu");
    const Json* rm = j.find("energy_r_max");
    c.energy_r_max = rm ? (float)rm->num : 1.0f;  // config.rs:33-35 default
    // export.py:67-68; the reference engine ignores both
    auto opt = [&](const char* key) {
        const Json* e = j.find(key);
        return e && e->kind == Json::Num ? (float)e->num : std::numeric_limits<float>::quiet_NaN();
    };
    c.excess_r = opt("excess_r");
    c.excess_r_min = opt("excess_r_min");
    if (c.channels != 3) fail(J2K_ERR_BAD_CONFIG, "config.json: channels must be 3 (h, u, v)");
    if (c.dim == 0 || c.heads == 0 || c.dim % c.heads != 0)
        fail(J2K_ERR_BAD_CONFIG, "config.json: dim must be a positive multiple of heads");
    c.ff_inner = (int)get(w, "blocks.0.ff1.weight").shape.at(1);
    c.max_iter = 32;
    c.tol = 1e-4f;
    c.stall_stop = 1;

    if (c.form == J2K_FORM_PATCH2D) {
        c.nx = uint_field(j, "nx");
        c.ny = uint_field(j, "ny");
        c.patch = uint_field(j, "patch");
        if (c.patch == 0 || c.nx % c.patch || c.ny % c.patch)
            fail(J2K_ERR_BAD_CONFIG, "config.json: nx and ny must be multiples of patch");
        c.head_dim = (int)get(w, "head.to_cells.weight").shape.at(0);
        c.tokens_x = c.nx / c.patch;
        c.tokens_y = c.ny / c.patch;
        c.tokens = c.tokens_x * c.tokens_y;
        c.patch_dim = c.channels * c.patch * c.patch;
        const Json* m = j.find("mask");  // config.rs:37-39: default Full
        if (!m || (m->kind == Json::Str && m->str == "Full")) {
            c.mask_kind = J2K_MASK_FULL;
            c.radius = 0;
            c.dilation = 1;
        } else if (m->kind == Json::Obj && m->find("Stencil")) {
            const Json& s = *m->find("Stencil");
            c.mask_kind = J2K_MASK_STENCIL;
            c.radius = uint_field(s, "radius");
            c.dilation = uint_field(s, "dilation");
            if (c.dilation == 0) fail(J2K_ERR_BAD_CONFIG, "config.json: stencil dilation must be >= 1");
        } else {
            fail(J2K_ERR_BAD_CONFIG, "config.json: `mask` must be \"Full\" or {\"Stencil\": {radius, dilation}}");
        }
    } else {
        const Json& s = need(j, "stencil");
        c.mask_kind = J2K_MASK_STENCIL;
        c.radius = uint_field(s, "radius");
        c.time_radius = uint_field(s, "time_radius");
        c.dilation = 1;
        c.head_dim = 0;
        c.patch = 1;
        c.patch_dim = c.channels + 1;
        if (const Json* mi = j.find("max_iter")) c.max_iter = (int)mi->num;
        if (const Json* t = j.find("tol")) c.tol = (float)t->num;
        if (const Json* st = j.find("stall_stop")) c.stall_stop = st->kind == Json::Bool ? (st->b ? 1 : 0) : (int)st->num;
    }
    return c;
}

}  // namespace j2k
