// hgx_wxf.hpp -- Wolfram Exchange Format (WXF) values: a small tree, a bounds-checked reader and a
// writer. This is the envelope the reference's FFI and its hg_evolve worker speak
// (paclet_source/hypergraph_ffi.cpp:136-376 parse, hg_gpu_backend.cpp:362-560 result;
// hg_evolve_main.cpp:1-120 frames), restated from the published WXF layout:
//
//   header      "8:" (version 8, uncompressed; "8C:" compressed is refused)
//   'f' n head args...   function with n arguments      'A' n ('-'|':' key value)...  association
//   'C' i8  'j' i16  'i' i32  'L' i64 (little endian)    'r' f64
//   'S' n utf8   's' n name   'B' n bytes   'I' n digits   'R' n digits
//   0xC1 / 0xC2 type rank dims... data                   packed / numeric array
//   n = unsigned LEB128 varint (at most 10 bytes)
//
// Packed and numeric integer arrays are read as nested lists of integers, which is how the FFI
// reads rules and states. The writer emits the smallest integer token that holds a value, as
// the Wolfram kernel does.
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace hgx {
namespace wxf {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Value {
    enum Kind : uint8_t { Int, Real, String, Symbol, Binary, BigInt, BigReal, Function, Assoc };
    Kind kind = Int;
    int64_t i = 0;
    double r = 0;
    std::string s;                                        // String/Symbol/Binary/BigInt/BigReal
    std::vector<Value> args;                              // Function: args[0] is the head
    std::vector<std::pair<Value, Value>> rules;           // Assoc
    std::vector<uint8_t> delayed;                         // Assoc: 1 for ':' (RuleDelayed)

    static Value integer(int64_t v) { Value x; x.kind = Int; x.i = v; return x; }
    static Value real(double v) { Value x; x.kind = Real; x.r = v; return x; }
    static Value string(std::string v) { Value x; x.kind = String; x.s = std::move(v); return x; }
    static Value symbol(std::string v) { Value x; x.kind = Symbol; x.s = std::move(v); return x; }
    static Value boolean(bool b) { return symbol(b ? "True" : "False"); }
    static Value function(const std::string& head, std::vector<Value> a = {}) {
        Value x; x.kind = Function;
        x.args.reserve(a.size() + 1);
        x.args.push_back(symbol(head));
        for (auto& v : a) x.args.push_back(std::move(v));
        return x;
    }
    static Value list(std::vector<Value> a = {}) { return function("List", std::move(a)); }
    static Value assoc() { Value x; x.kind = Assoc; return x; }

    bool is_symbol(const char* name) const { return kind == Symbol && s == name; }
    bool is_head(const char* name) const {
        return kind == Function && !args.empty() && args[0].is_symbol(name);
    }
    bool is_list() const { return is_head("List"); }
    size_t nargs() const { return kind == Function ? args.size() - 1 : 0; }
    const Value& arg(size_t k) const { return args.at(k + 1); }
    void push(Value v) { args.push_back(std::move(v)); }                       // list append
    void put(const std::string& key, Value v) {                                 // assoc append
        rules.emplace_back(string(key), std::move(v));
        delayed.push_back(0);
    }
    void put(int64_t key, Value v) {
        rules.emplace_back(integer(key), std::move(v));
        delayed.push_back(0);
    }
    const Value* find(const char* key) const {
        for (auto& kv : rules)
            if ((kv.first.kind == String || kv.first.kind == Symbol) && kv.first.s == key) return &kv.second;
        return nullptr;
    }
};

// ---- reader ---------------------------------------------------------------------------------

class Reader {
public:
    Reader(const uint8_t* p, size_t n) : p_(p), n_(n) {}

    // The whole input: header, one expression, nothing after it.
    Value document() {
        if (n_ >= 3 && p_[0] == '8' && p_[1] == 'C' && p_[2] == ':')
            throw Error("compressed WXF (8C:) is not supported; serialize without compression");
        if (n_ < 2 || p_[0] != '8' || p_[1] != ':') throw Error("not WXF: missing the 8: header");
        k_ = 2;
        Value v = expr(0);
        if (k_ != n_) throw Error("trailing bytes after the WXF expression");
        return v;
    }

private:
    const uint8_t* p_;
    size_t n_, k_ = 0;
    static constexpr int kMaxDepth = 512;

    uint8_t byte() {
        if (k_ >= n_) throw Error("unexpected end of WXF data");
        return p_[k_++];
    }
    const uint8_t* take(size_t m) {
        if (m > n_ - k_) throw Error("unexpected end of WXF data");
        const uint8_t* q = p_ + k_;
        k_ += m;
        return q;
    }
    uint64_t varint() {
        uint64_t v = 0;
        for (int sh = 0, i = 0;; sh += 7, ++i) {
            if (i == 10) throw Error("varint longer than 10 bytes");
            const uint8_t b = byte();
            if (i == 9 && (b & 0x7E)) throw Error("varint overflows 64 bits");
            v |= (uint64_t)(b & 0x7F) << sh;
            if (!(b & 0x80)) return v;
        }
    }
    // a count of items that each take at least `min_bytes`: refuse counts the input cannot hold
    size_t count(size_t min_bytes) {
        const uint64_t c = varint();
        if (c > (n_ - k_) / (min_bytes ? min_bytes : 1)) throw Error("WXF length exceeds the input");
        return (size_t)c;
    }
    template <class T>
    T le() {
        const uint8_t* q = take(sizeof(T));
        T v;
        std::memcpy(&v, q, sizeof(T));
        return v;
    }
    std::string text() {
        const size_t m = count(1);
        const uint8_t* q = take(m);
        return std::string((const char*)q, m);
    }

    Value expr(int depth) {
        if (depth > kMaxDepth) throw Error("WXF nesting too deep");
        const uint8_t t = byte();
        Value v;
        switch (t) {
            case 'C': return Value::integer(le<int8_t>());
            case 'j': return Value::integer(le<int16_t>());
            case 'i': return Value::integer(le<int32_t>());
            case 'L': return Value::integer(le<int64_t>());
            case 'r': return Value::real(le<double>());
            case 'S': v.kind = Value::String; v.s = text(); return v;
            case 's': v.kind = Value::Symbol; v.s = text(); return v;
            case 'B': v.kind = Value::Binary; v.s = text(); return v;
            case 'I': v.kind = Value::BigInt; v.s = text(); return v;
            case 'R': v.kind = Value::BigReal; v.s = text(); return v;
            case 'f': {
                const size_t m = count(1);
                v.kind = Value::Function;
                v.args.reserve(m + 1);
                v.args.push_back(expr(depth + 1));
                for (size_t a = 0; a < m; ++a) v.args.push_back(expr(depth + 1));
                return v;
            }
            case 'A': {
                const size_t m = count(3);
                v.kind = Value::Assoc;
                v.rules.reserve(m);
                for (size_t a = 0; a < m; ++a) {
                    const uint8_t rt = byte();
                    if (rt != '-' && rt != ':') throw Error("association entry is not a rule");
                    Value key = expr(depth + 1);
                    Value val = expr(depth + 1);
                    v.rules.emplace_back(std::move(key), std::move(val));
                    v.delayed.push_back(rt == ':' ? 1 : 0);
                }
                return v;
            }
            case 0xC1:
            case 0xC2:
                return array(t);
            default:
                throw Error("unknown WXF token " + std::to_string((int)t));
        }
    }

    // Integer packed/numeric arrays become nested List[...] of integers; real arrays of reals.
    Value array(uint8_t t) {
        const uint8_t ty = byte();
        const uint64_t rank = varint();
        if (rank == 0 || rank > 16) throw Error("array rank out of range");
        std::vector<uint64_t> dims(rank);
        // Every prefix product of the dims counts nodes the nested lists will hold (list nodes
        // at the inner levels, elements at the last). Elements are bounded by the data bytes
        // below; list nodes by the input left plus a fixed allowance of empty lists, so dims
        // like {2^28, 0} (no data, 2^28 empty lists) are refused before anything is reserved.
        const uint64_t room = (uint64_t)(n_ - k_) + 4096;
        uint64_t total = 1;
        for (auto& d : dims) {
            d = varint();
            if (d && total > room / d) throw Error("array larger than the input");
            total *= d;
        }
        size_t w;
        switch (ty) {
            case 0x00: case 0x10: w = 1; break;
            case 0x01: case 0x11: w = 2; break;
            case 0x02: case 0x12: case 0x22: w = 4; break;
            case 0x03: case 0x13: case 0x23: w = 8; break;
            default: throw Error("unsupported array element type");
        }
        if (t == 0xC1 && ty >= 0x10 && ty < 0x20) throw Error("unsigned element type in a packed array");
        if (total > (n_ - k_) / w) throw Error("array data exceeds the input");
        const uint8_t* q = take((size_t)(total * w));
        size_t at = 0;
        auto elem = [&]() -> Value {
            const uint8_t* e = q + (at++) * w;
            switch (ty) {
                case 0x00: { int8_t x; std::memcpy(&x, e, 1); return Value::integer(x); }
                case 0x01: { int16_t x; std::memcpy(&x, e, 2); return Value::integer(x); }
                case 0x02: { int32_t x; std::memcpy(&x, e, 4); return Value::integer(x); }
                case 0x03: { int64_t x; std::memcpy(&x, e, 8); return Value::integer(x); }
                case 0x10: return Value::integer(e[0]);
                case 0x11: { uint16_t x; std::memcpy(&x, e, 2); return Value::integer(x); }
                case 0x12: { uint32_t x; std::memcpy(&x, e, 4); return Value::integer(x); }
                case 0x13: {
                    uint64_t x; std::memcpy(&x, e, 8);
                    if (x > (uint64_t)INT64_MAX) throw Error("UnsignedInteger64 element above 2^63-1");
                    return Value::integer((int64_t)x);
                }
                case 0x22: { float x; std::memcpy(&x, e, 4); return Value::real(x); }
                default: { double x; std::memcpy(&x, e, 8); return Value::real(x); }
            }
        };
        return nest(dims, 0, elem);
    }
    template <class F>
    Value nest(const std::vector<uint64_t>& dims, size_t lvl, F& elem) {
        Value v = Value::list();
        v.args.reserve(dims[lvl] + 1);
        for (uint64_t a = 0; a < dims[lvl]; ++a)
            v.push(lvl + 1 == dims.size() ? elem() : nest(dims, lvl + 1, elem));
        return v;
    }
};

// ---- writer ---------------------------------------------------------------------------------

class Writer {
public:
    std::vector<uint8_t> out;
    Writer() { out.push_back('8'); out.push_back(':'); }

    void value(const Value& v) {
        switch (v.kind) {
            case Value::Int: integer(v.i); break;
            case Value::Real: out.push_back('r'); raw(&v.r, 8); break;
            case Value::String: text('S', v.s); break;
            case Value::Symbol: text('s', v.s); break;
            case Value::Binary: text('B', v.s); break;
            case Value::BigInt: text('I', v.s); break;
            case Value::BigReal: text('R', v.s); break;
            case Value::Function:
                out.push_back('f');
                varint(v.args.size() - 1);
                for (auto& a : v.args) value(a);
                break;
            case Value::Assoc:
                out.push_back('A');
                varint(v.rules.size());
                for (size_t k = 0; k < v.rules.size(); ++k) {
                    out.push_back(k < v.delayed.size() && v.delayed[k] ? ':' : '-');
                    value(v.rules[k].first);
                    value(v.rules[k].second);
                }
                break;
        }
    }
    void integer(int64_t x) {
        if (x >= INT8_MIN && x <= INT8_MAX) { out.push_back('C'); const int8_t y = (int8_t)x; raw(&y, 1); }
        else if (x >= INT16_MIN && x <= INT16_MAX) { out.push_back('j'); const int16_t y = (int16_t)x; raw(&y, 2); }
        else if (x >= INT32_MIN && x <= INT32_MAX) { out.push_back('i'); const int32_t y = (int32_t)x; raw(&y, 4); }
        else { out.push_back('L'); raw(&x, 8); }
    }
    void varint(uint64_t v) {
        while (v >= 0x80) { out.push_back((uint8_t)((v & 0x7F) | 0x80)); v >>= 7; }
        out.push_back((uint8_t)v);
    }

private:
    void raw(const void* p, size_t m) {
        const uint8_t* b = (const uint8_t*)p;
        out.insert(out.end(), b, b + m);
    }
    void text(uint8_t tag, const std::string& s) {
        out.push_back(tag);
        varint(s.size());
        raw(s.data(), s.size());
    }
};

inline Value parse(const uint8_t* p, size_t n) { return Reader(p, n).document(); }
inline std::vector<uint8_t> serialize(const Value& v) {
    Writer w;
    w.value(v);
    return std::move(w.out);
}

}  // namespace wxf
}  // namespace hgx
