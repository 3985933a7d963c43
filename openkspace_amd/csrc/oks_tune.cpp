// oks_tune.cpp -- reading the autotuner's launch-shape table (JSON).
#include "oks_tune.hpp"

#include <dlfcn.h>

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace oks {

namespace {

// A JSON value: just enough of the grammar for the table (objects, arrays,
// strings without unicode escapes beyond \uXXXX passthrough, numbers,
// true/false/null).
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const std::string& s;
    size_t i = 0;
    explicit JParser(const std::string& t) : s(t) {}
    [[noreturn]] void bad(const char* what) {
        throw std::runtime_error(std::string("tuning table: ") + what + " at byte " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    std::string string() {
        if (!eat('"')) bad("expected a string");
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) bad("bad escape");
                const char e = s[i++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'u':
                        if (i + 4 > s.size()) bad("bad \\u escape");
                        i += 4;
                        out += '?';
                        break;
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (!eat('"')) bad("unterminated string");
        return out;
    }
    JVal value(int depth = 0) {
        if (depth > 32) bad("nesting too deep");
        ws();
        if (i >= s.size()) bad("unexpected end");
        JVal v;
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.kind = JVal::Obj;
            if (eat('}')) return v;
            do {
                std::string k = string();
                if (!eat(':')) bad("expected ':'");
                v.obj.emplace_back(std::move(k), value(depth + 1));
            } while (eat(','));
            if (!eat('}')) bad("expected '}'");
        } else if (c == '[') {
            ++i;
            v.kind = JVal::Arr;
            if (eat(']')) return v;
            do {
                v.arr.push_back(value(depth + 1));
            } while (eat(','));
            if (!eat(']')) bad("expected ']'");
        } else if (c == '"') {
            v.kind = JVal::Str;
            v.str = string();
        } else if (s.compare(i, 4, "true") == 0) {
            i += 4;
            v.kind = JVal::Bool;
            v.b = true;
        } else if (s.compare(i, 5, "false") == 0) {
            i += 5;
            v.kind = JVal::Bool;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            const char* p = s.c_str() + i;
            char* end = nullptr;
            v.num = std::strtod(p, &end);
            if (end == p) bad("expected a value");
            i += (size_t)(end - p);
            v.kind = JVal::Num;
        }
        return v;
    }
};

int64_t as_int(const JVal& v, const char* what) {
    if (v.kind != JVal::Num || v.num != (double)(int64_t)v.num)
        throw std::runtime_error(std::string("tuning table: ") + what + " must be an integer");
    return (int64_t)v.num;
}

std::string library_dir() {
    Dl_info info{};
    if (dladdr((void*)&library_dir, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t slash = p.rfind('/');
        if (slash != std::string::npos) return p.substr(0, slash);
    }
    return ".";
}

}  // namespace

const TuneEntry* TuneTable::find(const std::string& op, const std::vector<int64_t>& key) const {
    for (auto& e : entries)
        if (e.op == op && e.key == key) return &e;
    return nullptr;
}

TuneTable parse_tune_table(const std::string& text) {
    JParser p(text);
    const JVal root = p.value();
    p.ws();
    if (p.i != text.size()) p.bad("trailing characters");
    if (root.kind != JVal::Obj) throw std::runtime_error("tuning table: the root must be an object");
    TuneTable t;
    const JVal* arch = root.get("arch");
    if (!arch || arch->kind != JVal::Str) throw std::runtime_error("tuning table: missing \"arch\"");
    t.arch = arch->str;
    const JVal* ents = root.get("entries");
    if (!ents || ents->kind != JVal::Arr) throw std::runtime_error("tuning table: missing \"entries\"");
    for (const JVal& e : ents->arr) {
        const JVal* op = e.get("op");
        const JVal* key = e.get("key");
        const JVal* opts = e.get("opts");
        if (e.kind != JVal::Obj || !op || op->kind != JVal::Str || !key || key->kind != JVal::Arr || !opts ||
            opts->kind != JVal::Obj)
            throw std::runtime_error("tuning table: an entry needs \"op\", \"key\" and \"opts\"");
        TuneEntry te;
        te.op = op->str;
        for (const JVal& k : key->arr) te.key.push_back(as_int(k, "a key element"));
        for (auto& kv : opts->obj) te.opts[kv.first] = as_int(kv.second, "an option value");
        t.entries.push_back(std::move(te));
    }
    return t;
}

const TuneTable* tune_table_for(const std::string& arch) {
    static std::mutex mu;
    static std::map<std::string, std::unique_ptr<TuneTable>> tables;  // arch -> table (null: none)
    std::lock_guard<std::mutex> lk(mu);
    auto it = tables.find(arch);
    if (it != tables.end()) return it->second.get();
    std::string path;
    if (const char* env = std::getenv("OKS_TUNE_FILE")) {
        path = env;
    } else {
        path = library_dir() + "/tuning/" + arch + ".json";
    }
    std::unique_ptr<TuneTable> t;
    if (!path.empty()) {
        std::ifstream f(path);
        if (f) {
            std::stringstream ss;
            ss << f.rdbuf();
            t.reset(new TuneTable(parse_tune_table(ss.str())));
            t->source = path;
            if (t->arch != arch) t.reset();
        }
    }
    const TuneTable* r = t.get();
    tables[arch] = std::move(t);
    return r;
}

}  // namespace oks
