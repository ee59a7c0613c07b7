// Simulation state files: XML and JSON readers and writers with the layouts of
// src/flamegpu/io/{XMLStateReader,XMLStateWriter,JSONStateReader,JSONStateWriter}.cu.
//
//   XML:  <states> <config><simulation>..</simulation><cuda>..</cuda></config>
//         <itno>N</itno> <stats><step_count>N</step_count></stats>
//         <environment><prop type="..">v[,v..]</prop>..</environment>
//         <xagent><name>A</name><state>S</state><var>v[,v..]</var>..</xagent>.. </states>
//   JSON: {"config":{"simulation":{..},"cuda":{..}}, "stats":{"step_count":N},
//          "environment":{"prop":v|[v..]}, "agents":{"A":{"S":[{"var":v|[v..]},..]}}}
//
// Values are written with enough digits to read back bit-exactly (9 significant
// digits for float, 17 for double), so an exported state resumes exactly; the
// reference writes floats through a default-precision stream.  The file is read on the
// host and handed to setPopulationData / the environment, as the reference does.
#pragma once

#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "flamegpu/model.h"

namespace flamegpu {
namespace io {

// Format-independent content of a state file: values as text tokens.
struct StateData {
    std::map<std::string, std::string> simulation;   // config/simulation key -> text
    std::map<std::string, std::string> cuda;         // config/cuda key -> text
    bool has_step = false;
    unsigned int step_count = 0;
    std::map<std::string, uint32_t> next_id;         // stats/next_id: agent -> next agent ID
    std::map<std::string, std::vector<std::string>> environment;
    struct AgentRecord {
        std::vector<std::pair<std::string, std::vector<std::string>>> vars;
    };
    // (agent, state) -> records in file order; state "" = the agent's initial state
    std::map<std::pair<std::string, std::string>, std::vector<AgentRecord>> agents;
};

namespace detail {

inline std::string lower_ext(const std::string& path) {
    const size_t d = path.find_last_of('.');
    std::string e = d == std::string::npos ? "" : path.substr(d + 1);
    for (auto& c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

inline std::vector<std::string> split_commas(const std::string& s) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : s) {
        if (c == ',') {
            out.push_back(cur);
            cur.clear();
        } else if (!std::isspace((unsigned char)c)) {
            cur += c;
        }
    }
    if (!cur.empty() || !out.empty()) out.push_back(cur);
    return out;
}

inline std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw exception::InvalidArgument("Unable to open state file '" + path + "'");
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// ---------------------------------------------------------------- minimal XML
struct XNode {
    std::string name, text;
    std::vector<XNode> kids;
    const XNode* child(const std::string& n) const {
        for (auto& k : kids)
            if (k.name == n) return &k;
        return nullptr;
    }
};

class XmlCursor {
 public:
    explicit XmlCursor(const std::string& s) : s_(s) {}
    // Skips the prolog, comments and whitespace; positions at the root element.
    void skip_misc() {
        for (;;) {
            while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
            if (s_.compare(p_, 4, "<!--") == 0) {
                const size_t e = s_.find("-->", p_);
                if (e == std::string::npos) fail("unterminated comment");
                p_ = e + 3;
            } else if (s_.compare(p_, 2, "<?") == 0 || s_.compare(p_, 2, "<!") == 0) {
                const size_t e = s_.find('>', p_);
                if (e == std::string::npos) fail("unterminated declaration");
                p_ = e + 1;
            } else {
                return;
            }
        }
    }
    // Reads "<name ...>" or "<name .../>"; returns false at "</...".
    bool open_tag(std::string& name, bool& empty) {
        skip_misc();
        if (p_ >= s_.size() || s_[p_] != '<') fail("expected an element");
        if (s_.compare(p_, 2, "</") == 0) return false;
        ++p_;
        const size_t b = p_;
        while (p_ < s_.size() && !std::isspace((unsigned char)s_[p_]) && s_[p_] != '>' && s_[p_] != '/') ++p_;
        name = s_.substr(b, p_ - b);
        // attributes are not used by the reader (types come from the model)
        bool quote = false;
        char q = 0;
        while (p_ < s_.size()) {
            const char c = s_[p_];
            if (quote) {
                if (c == q) quote = false;
            } else if (c == '"' || c == '\'') {
                quote = true;
                q = c;
            } else if (c == '>') {
                break;
            }
            ++p_;
        }
        if (p_ >= s_.size()) fail("unterminated tag <" + name);
        empty = s_[p_ - 1] == '/';
        ++p_;
        return true;
    }
    void close_tag(const std::string& name) {
        skip_misc();
        if (s_.compare(p_, 2, "</") != 0) fail("expected </" + name + ">");
        const size_t e = s_.find('>', p_);
        if (e == std::string::npos) fail("unterminated closing tag");
        std::string n = s_.substr(p_ + 2, e - p_ - 2);
        while (!n.empty() && std::isspace((unsigned char)n.back())) n.pop_back();
        if (n != name) fail("mismatched </" + n + ">, expected </" + name + ">");
        p_ = e + 1;
    }
    // Parses one element subtree (the opening tag already read).
    void element_body(XNode& n, bool empty) {
        if (empty) return;
        for (;;) {
            // text up to the next '<'
            const size_t lt = s_.find('<', p_);
            if (lt == std::string::npos) fail("unterminated element <" + n.name + ">");
            n.text += unescape(s_.substr(p_, lt - p_));
            p_ = lt;
            if (s_.compare(p_, 4, "<!--") == 0) {
                const size_t e = s_.find("-->", p_);
                if (e == std::string::npos) fail("unterminated comment");
                p_ = e + 3;
                continue;
            }
            if (s_.compare(p_, 2, "</") == 0) break;
            XNode k;
            bool ke;
            open_tag(k.name, ke);
            element_body(k, ke);
            n.kids.push_back(std::move(k));
        }
        close_tag(n.name);
        trim(n.text);
    }
    [[noreturn]] void fail(const std::string& m) const {
        throw exception::InvalidArgument("XML state file: " + m + " (at byte " + std::to_string(p_) + ")");
    }

 private:
    static void trim(std::string& t) {
        size_t a = 0, b = t.size();
        while (a < b && std::isspace((unsigned char)t[a])) ++a;
        while (b > a && std::isspace((unsigned char)t[b - 1])) --b;
        t = t.substr(a, b - a);
    }
    static std::string unescape(const std::string& t) {
        if (t.find('&') == std::string::npos) return t;
        std::string o;
        for (size_t i = 0; i < t.size(); ++i) {
            if (t[i] != '&') {
                o += t[i];
                continue;
            }
            const size_t e = t.find(';', i);
            const std::string ent = e == std::string::npos ? "" : t.substr(i + 1, e - i - 1);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else {
                o += '&';
                continue;
            }
            i = e;
        }
        return o;
    }
    const std::string& s_;
    size_t p_ = 0;
};

// src/flamegpu/io/XMLStateReader.cu: parse()
inline StateData read_xml(const std::string& path) {
    const std::string text = read_file(path);
    XmlCursor c(text);
    StateData d;
    std::string root;
    bool empty;
    if (!c.open_tag(root, empty)) c.fail("missing root element");
    if (empty) return d;
    for (;;) {
        XNode n;
        bool e;
        if (!c.open_tag(n.name, e)) break;
        c.element_body(n, e);
        if (n.name == "config") {
            if (auto* s = n.child("simulation"))
                for (auto& k : s->kids) d.simulation[k.name] = k.text;
            if (auto* s = n.child("cuda"))
                for (auto& k : s->kids) d.cuda[k.name] = k.text;
        } else if (n.name == "itno") {
            d.has_step = true;
            d.step_count = (unsigned int)std::stoul(n.text);
        } else if (n.name == "stats") {
            if (auto* s = n.child("step_count")) {
                d.has_step = true;
                d.step_count = (unsigned int)std::stoul(s->text);
            }
            if (auto* s = n.child("next_id"))
                for (auto& k : s->kids) d.next_id[k.name] = (uint32_t)std::stoul(k.text);
        } else if (n.name == "environment") {
            for (auto& k : n.kids) {
                if (d.environment.count(k.name))
                    throw exception::InvalidArgument("Input file contains environment property '" + k.name + "' multiple times");
                d.environment[k.name] = split_commas(k.text);
            }
        } else if (n.name == "xagent") {
            const XNode* nm = n.child("name");
            if (!nm) c.fail("xagent without <name>");
            const XNode* st = n.child("state");
            StateData::AgentRecord r;
            for (auto& k : n.kids)
                if (k.name != "name" && k.name != "state") r.vars.emplace_back(k.name, split_commas(k.text));
            d.agents[{nm->text, st ? st->text : std::string()}].push_back(std::move(r));
        }
        // macro_environment and unknown blocks are skipped
    }
    c.close_tag(root);
    return d;
}

// ---------------------------------------------------------------- minimal JSON
struct JValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    std::string text;                                   // number / string / bool text
    std::vector<JValue> items;                          // array
    std::vector<std::pair<std::string, JValue>> fields; // object
    const JValue* get(const std::string& k) const {
        for (auto& f : fields)
            if (f.first == k) return &f.second;
        return nullptr;
    }
};

class JsonParser {
 public:
    explicit JsonParser(const std::string& s) : s_(s) {}
    JValue parse() {
        JValue v = value();
        ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

 private:
    void ws() {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
    }
    [[noreturn]] void fail(const std::string& m) const {
        throw exception::InvalidArgument("JSON state file: " + m + " (at byte " + std::to_string(p_) + ")");
    }
    std::string str() {
        if (s_[p_] != '"') fail("expected a string");
        ++p_;
        std::string o;
        while (p_ < s_.size() && s_[p_] != '"') {
            if (s_[p_] == '\\') {
                ++p_;
                if (p_ >= s_.size()) break;
                const char e = s_[p_];
                o += e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'b' ? '\b' : e == 'f' ? '\f' : e;
                if (e == 'u') {  // keep \uXXXX as is; names in state files are ASCII
                    o.pop_back();
                    o += "\\u";
                }
            } else {
                o += s_[p_];
            }
            ++p_;
        }
        if (p_ >= s_.size()) fail("unterminated string");
        ++p_;
        return o;
    }
    JValue value() {
        ws();
        if (p_ >= s_.size()) fail("unexpected end");
        JValue v;
        const char c = s_[p_];
        if (c == '{') {
            v.kind = JValue::Object;
            ++p_;
            ws();
            if (s_[p_] == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                std::string k = str();
                ws();
                if (s_[p_] != ':') fail("expected ':'");
                ++p_;
                v.fields.emplace_back(std::move(k), value());
                ws();
                if (s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (s_[p_] == '}') {
                    ++p_;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JValue::Array;
            ++p_;
            ws();
            if (s_[p_] == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.items.push_back(value());
                ws();
                if (s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (s_[p_] == ']') {
                    ++p_;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JValue::String;
            v.text = str();
            return v;
        }
        const size_t b = p_;
        while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '-' || s_[p_] == '+' || s_[p_] == '.'))
            ++p_;
        v.text = s_.substr(b, p_ - b);
        if (v.text.empty()) fail("unexpected character");
        if (v.text == "true" || v.text == "false") v.kind = JValue::Bool;
        else if (v.text == "null") v.kind = JValue::Null;
        else v.kind = JValue::Number;
        return v;
    }
    const std::string& s_;
    size_t p_ = 0;
};

inline std::vector<std::string> json_tokens(const JValue& v) {
    if (v.kind == JValue::Array) {
        std::vector<std::string> o;
        for (auto& i : v.items) o.push_back(i.text);
        return o;
    }
    return {v.text};
}

// src/flamegpu/io/JSONStateReader.cu
inline StateData read_json(const std::string& path) {
    const std::string text = read_file(path);
    const JValue root = JsonParser(text).parse();
    if (root.kind != JValue::Object) throw exception::InvalidArgument("JSON state file: root is not an object");
    StateData d;
    if (auto* cfg = root.get("config")) {
        if (auto* s = cfg->get("simulation"))
            for (auto& f : s->fields) d.simulation[f.first] = f.second.text;
        if (auto* s = cfg->get("cuda"))
            for (auto& f : s->fields) d.cuda[f.first] = f.second.text;
    }
    if (auto* st = root.get("stats")) {
        if (auto* sc = st->get("step_count")) {
            d.has_step = true;
            d.step_count = (unsigned int)std::stoul(sc->text);
        }
        if (auto* ni = st->get("next_id"))
            for (auto& f : ni->fields) d.next_id[f.first] = (uint32_t)std::stoul(f.second.text);
    }
    if (auto* env = root.get("environment"))
        for (auto& f : env->fields) d.environment[f.first] = json_tokens(f.second);
    if (auto* ag = root.get("agents")) {
        for (auto& a : ag->fields)
            for (auto& s : a.second.fields) {
                auto& list = d.agents[{a.first, s.first}];
                for (auto& rec : s.second.items) {
                    StateData::AgentRecord r;
                    for (auto& f : rec.fields) r.vars.emplace_back(f.first, json_tokens(f.second));
                    list.push_back(std::move(r));
                }
            }
    }
    return d;
}

// ---------------------------------------------------------------- typed values
// Parses one token into element `e` of a value of `type` at dst.
inline void parse_token(const std::type_index& type, const std::string& tok, char* dst, size_t e) {
    try {
        if (type == std::type_index(typeid(float))) {
            const float v = std::stof(tok);
            std::memcpy(dst + e * 4, &v, 4);
        } else if (type == std::type_index(typeid(double))) {
            const double v = std::stod(tok);
            std::memcpy(dst + e * 8, &v, 8);
        } else if (type == std::type_index(typeid(int64_t)) || type == std::type_index(typeid(long long))) {
            const int64_t v = std::stoll(tok);
            std::memcpy(dst + e * 8, &v, 8);
        } else if (type == std::type_index(typeid(uint64_t)) || type == std::type_index(typeid(unsigned long long))) {
            const uint64_t v = std::stoull(tok);
            std::memcpy(dst + e * 8, &v, 8);
        } else if (type == std::type_index(typeid(int32_t))) {
            const int32_t v = (int32_t)std::stoll(tok);
            std::memcpy(dst + e * 4, &v, 4);
        } else if (type == std::type_index(typeid(uint32_t))) {
            const uint32_t v = (uint32_t)std::stoull(tok);
            std::memcpy(dst + e * 4, &v, 4);
        } else if (type == std::type_index(typeid(int16_t))) {
            const int16_t v = (int16_t)std::stoll(tok);
            std::memcpy(dst + e * 2, &v, 2);
        } else if (type == std::type_index(typeid(uint16_t))) {
            const uint16_t v = (uint16_t)std::stoull(tok);
            std::memcpy(dst + e * 2, &v, 2);
        } else if (type == std::type_index(typeid(int8_t)) || type == std::type_index(typeid(char))) {
            const int8_t v = (int8_t)std::stoll(tok);
            std::memcpy(dst + e, &v, 1);
        } else if (type == std::type_index(typeid(uint8_t))) {
            const uint8_t v = (uint8_t)std::stoull(tok);
            std::memcpy(dst + e, &v, 1);
        } else {
            throw exception::UnsupportedVarType(std::string("state files do not support type ") + type.name());
        }
    } catch (const std::invalid_argument&) {
        throw exception::InvalidArgument("state file value '" + tok + "' is not a number");
    } catch (const std::out_of_range&) {
        throw exception::InvalidArgument("state file value '" + tok + "' is out of range");
    }
}

inline std::string format_value(const std::type_index& type, const char* src, size_t e) {
    char b[64];
    if (type == std::type_index(typeid(float))) {
        float v;
        std::memcpy(&v, src + e * 4, 4);
        snprintf(b, sizeof b, "%.9g", (double)v);
    } else if (type == std::type_index(typeid(double))) {
        double v;
        std::memcpy(&v, src + e * 8, 8);
        snprintf(b, sizeof b, "%.17g", v);
    } else if (type == std::type_index(typeid(int64_t)) || type == std::type_index(typeid(long long))) {
        int64_t v;
        std::memcpy(&v, src + e * 8, 8);
        snprintf(b, sizeof b, "%lld", (long long)v);
    } else if (type == std::type_index(typeid(uint64_t)) || type == std::type_index(typeid(unsigned long long))) {
        uint64_t v;
        std::memcpy(&v, src + e * 8, 8);
        snprintf(b, sizeof b, "%llu", (unsigned long long)v);
    } else if (type == std::type_index(typeid(int32_t))) {
        int32_t v;
        std::memcpy(&v, src + e * 4, 4);
        snprintf(b, sizeof b, "%d", v);
    } else if (type == std::type_index(typeid(uint32_t))) {
        uint32_t v;
        std::memcpy(&v, src + e * 4, 4);
        snprintf(b, sizeof b, "%u", v);
    } else if (type == std::type_index(typeid(int16_t))) {
        int16_t v;
        std::memcpy(&v, src + e * 2, 2);
        snprintf(b, sizeof b, "%d", (int)v);
    } else if (type == std::type_index(typeid(uint16_t))) {
        uint16_t v;
        std::memcpy(&v, src + e * 2, 2);
        snprintf(b, sizeof b, "%u", (unsigned)v);
    } else if (type == std::type_index(typeid(int8_t)) || type == std::type_index(typeid(char))) {
        int8_t v;
        std::memcpy(&v, src + e, 1);
        snprintf(b, sizeof b, "%d", (int)v);
    } else if (type == std::type_index(typeid(uint8_t))) {
        uint8_t v;
        std::memcpy(&v, src + e, 1);
        snprintf(b, sizeof b, "%u", (unsigned)v);
    } else {
        throw exception::UnsupportedVarType(std::string("state files do not support type ") + type.name());
    }
    return b;
}

// Element e of a value of `type` at src, as double (exact for every type up to 53 bits).
inline double as_double(const std::type_index& type, const char* src, size_t e) {
    if (type == std::type_index(typeid(float))) { float v; std::memcpy(&v, src + e * 4, 4); return v; }
    if (type == std::type_index(typeid(double))) { double v; std::memcpy(&v, src + e * 8, 8); return v; }
    if (type == std::type_index(typeid(int64_t)) || type == std::type_index(typeid(long long))) { int64_t v; std::memcpy(&v, src + e * 8, 8); return (double)v; }
    if (type == std::type_index(typeid(uint64_t)) || type == std::type_index(typeid(unsigned long long))) { uint64_t v; std::memcpy(&v, src + e * 8, 8); return (double)v; }
    if (type == std::type_index(typeid(int32_t))) { int32_t v; std::memcpy(&v, src + e * 4, 4); return v; }
    if (type == std::type_index(typeid(uint32_t))) { uint32_t v; std::memcpy(&v, src + e * 4, 4); return v; }
    if (type == std::type_index(typeid(int16_t))) { int16_t v; std::memcpy(&v, src + e * 2, 2); return v; }
    if (type == std::type_index(typeid(uint16_t))) { uint16_t v; std::memcpy(&v, src + e * 2, 2); return v; }
    if (type == std::type_index(typeid(int8_t)) || type == std::type_index(typeid(char))) { int8_t v; std::memcpy(&v, src + e, 1); return v; }
    if (type == std::type_index(typeid(uint8_t))) { uint8_t v; std::memcpy(&v, src + e, 1); return v; }
    throw exception::UnsupportedVarType(std::string("state files do not support type ") + type.name());
}

// A readable type tag for the XML "type" attribute (the reference writes typeid names).
inline const char* type_tag(const std::type_index& t) {
    if (t == std::type_index(typeid(float))) return "float";
    if (t == std::type_index(typeid(double))) return "double";
    if (t == std::type_index(typeid(int64_t)) || t == std::type_index(typeid(long long))) return "int64_t";
    if (t == std::type_index(typeid(uint64_t)) || t == std::type_index(typeid(unsigned long long))) return "uint64_t";
    if (t == std::type_index(typeid(int32_t))) return "int32_t";
    if (t == std::type_index(typeid(uint32_t))) return "uint32_t";
    if (t == std::type_index(typeid(int16_t))) return "int16_t";
    if (t == std::type_index(typeid(uint16_t))) return "uint16_t";
    if (t == std::type_index(typeid(int8_t))) return "int8_t";
    if (t == std::type_index(typeid(uint8_t))) return "uint8_t";
    return "unknown";
}

}  // namespace detail

// What an exported state holds: configuration, step count, environment and every
// agent population (one (agent, state) entry each, variables as raw bytes per agent).
struct StateOut {
    struct Cfg {
        std::string key, text;
        bool is_string;   // quoted in JSON
    };
    std::vector<Cfg> simulation, cuda;
    unsigned int step_count = 0;
    std::map<std::string, uint32_t> next_id;   // agent -> next agent ID (exact resume of births)
    struct Env {
        std::string name;
        std::type_index type = std::type_index(typeid(void));
        unsigned int elements = 1;
        std::vector<char> bytes;
    };
    std::vector<Env> environment;
    struct Pop {
        std::string agent, state;
        uint32_t n = 0;
        struct Var {
            std::string name;
            std::type_index type = std::type_index(typeid(void));
            unsigned int elements = 1;
            const std::vector<char>* bytes = nullptr;
        };
        std::vector<Var> vars;
    };
    std::vector<Pop> pops;
};

// src/flamegpu/io/XMLStateWriter.cu
inline void write_xml(const std::string& path, const StateOut& s, bool pretty) {
    const char* nl = pretty ? "\n" : "";
    auto ind = [&](int k) { return pretty ? std::string(4 * k, ' ') : std::string(); };
    std::string o = "<states>";
    o += nl;
    o += ind(1) + "<config>" + nl + ind(2) + "<simulation>" + nl;
    for (auto& kv : s.simulation) o += ind(3) + "<" + kv.key + ">" + kv.text + "</" + kv.key + ">" + nl;
    o += ind(2) + "</simulation>" + nl + ind(2) + "<cuda>" + nl;
    for (auto& kv : s.cuda) o += ind(3) + "<" + kv.key + ">" + kv.text + "</" + kv.key + ">" + nl;
    o += ind(2) + "</cuda>" + nl + ind(1) + "</config>" + nl;
    o += ind(1) + "<itno>" + std::to_string(s.step_count) + "</itno>" + nl;
    o += ind(1) + "<stats>" + nl + ind(2) + "<step_count>" + std::to_string(s.step_count) + "</step_count>" + nl;
    o += ind(2) + "<next_id>" + nl;
    for (auto& kv : s.next_id) o += ind(3) + "<" + kv.first + ">" + std::to_string(kv.second) + "</" + kv.first + ">" + nl;
    o += ind(2) + "</next_id>" + nl + ind(1) + "</stats>" + nl;
    o += ind(1) + "<environment>" + nl;
    for (auto& e : s.environment) {
        o += ind(2) + "<" + e.name + " type=\"" + detail::type_tag(e.type) + "\">";
        for (unsigned int k = 0; k < e.elements; ++k) {
            if (k) o += ",";
            o += detail::format_value(e.type, e.bytes.data(), k);
        }
        o += "</" + e.name + ">" + nl;
    }
    o += ind(1) + "</environment>" + nl;
    for (auto& p : s.pops) {
        for (uint32_t i = 0; i < p.n; ++i) {
            o += ind(1) + "<xagent>" + nl + ind(2) + "<name>" + p.agent + "</name>" + nl + ind(2) + "<state>" + p.state +
                 "</state>" + nl;
            for (auto& v : p.vars) {
                o += ind(2) + "<" + v.name;
                if (i == 0) o += std::string(" type=\"") + detail::type_tag(v.type) + "\"";
                o += ">";
                const size_t bpe = v.bytes->size() / std::max<uint32_t>(p.n, 1);
                for (unsigned int k = 0; k < v.elements; ++k) {
                    if (k) o += ",";
                    o += detail::format_value(v.type, v.bytes->data() + (size_t)i * bpe, k);
                }
                o += "</" + v.name + ">" + nl;
            }
            o += ind(1) + "</xagent>" + nl;
        }
    }
    o += "</states>";
    o += nl;
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw exception::InvalidArgument("Unable to open '" + path + "' for writing");
    f << o;
}

// src/flamegpu/io/JSONStateWriter.cu
inline void write_json(const std::string& path, const StateOut& s, bool pretty) {
    const char* nl = pretty ? "\n" : "";
    auto ind = [&](int k) { return pretty ? std::string(4 * k, ' ') : std::string(); };
    const char* sp = pretty ? " " : "";
    std::string o = std::string("{") + nl;
    auto kvs = [&](const std::vector<StateOut::Cfg>& v, int lvl) {
        for (size_t i = 0; i < v.size(); ++i) {
            const std::string val = v[i].is_string ? "\"" + v[i].text + "\"" : v[i].text;
            o += ind(lvl) + "\"" + v[i].key + "\":" + sp + val + (i + 1 < v.size() ? "," : "") + nl;
        }
    };
    o += ind(1) + "\"config\":" + sp + "{" + nl + ind(2) + "\"simulation\":" + sp + "{" + nl;
    kvs(s.simulation, 3);
    o += ind(2) + "}," + nl + ind(2) + "\"cuda\":" + sp + "{" + nl;
    kvs(s.cuda, 3);
    o += ind(2) + "}" + nl + ind(1) + "}," + nl;
    o += ind(1) + "\"stats\":" + sp + "{" + nl + ind(2) + "\"step_count\":" + sp + std::to_string(s.step_count) + "," + nl;
    o += ind(2) + "\"next_id\":" + sp + "{";
    {
        size_t k = 0;
        for (auto& kv : s.next_id)
            o += "\"" + kv.first + "\":" + sp + std::to_string(kv.second) + (++k < s.next_id.size() ? "," : "");
    }
    o += std::string("}") + nl + ind(1) + "}," + nl;
    o += ind(1) + "\"environment\":" + sp + "{" + nl;
    for (size_t i = 0; i < s.environment.size(); ++i) {
        const auto& e = s.environment[i];
        o += ind(2) + "\"" + e.name + "\":" + sp;
        if (e.elements != 1) o += "[";
        for (unsigned int k = 0; k < e.elements; ++k) {
            if (k) o += ",";
            o += detail::format_value(e.type, e.bytes.data(), k);
        }
        if (e.elements != 1) o += "]";
        o += (i + 1 < s.environment.size() ? "," : "");
        o += nl;
    }
    o += ind(1) + "}," + nl + ind(1) + "\"agents\":" + sp + "{" + nl;
    // group populations by agent
    std::map<std::string, std::vector<const StateOut::Pop*>> by_agent;
    for (auto& p : s.pops) by_agent[p.agent].push_back(&p);
    size_t ai = 0;
    for (auto& a : by_agent) {
        o += ind(2) + "\"" + a.first + "\":" + sp + "{" + nl;
        for (size_t si = 0; si < a.second.size(); ++si) {
            const auto& p = *a.second[si];
            o += ind(3) + "\"" + p.state + "\":" + sp + "[" + nl;
            for (uint32_t i = 0; i < p.n; ++i) {
                o += ind(4) + "{";
                for (size_t vi = 0; vi < p.vars.size(); ++vi) {
                    const auto& v = p.vars[vi];
                    o += "\"" + v.name + "\":" + sp;
                    const size_t bpe = v.bytes->size() / std::max<uint32_t>(p.n, 1);
                    if (v.elements != 1) o += "[";
                    for (unsigned int k = 0; k < v.elements; ++k) {
                        if (k) o += ",";
                        o += detail::format_value(v.type, v.bytes->data() + (size_t)i * bpe, k);
                    }
                    if (v.elements != 1) o += "]";
                    if (vi + 1 < p.vars.size()) o += std::string(",") + sp;
                }
                o += std::string("}") + (i + 1 < p.n ? "," : "") + nl;
            }
            o += ind(3) + "]" + (si + 1 < a.second.size() ? "," : "") + nl;
        }
        o += ind(2) + "}" + (++ai < by_agent.size() ? "," : "") + nl;
    }
    o += ind(1) + "}" + nl + "}" + nl;
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw exception::InvalidArgument("Unable to open '" + path + "' for writing");
    f << o;
}

// Reads an .xml or .json state file (src/flamegpu/io/StateReaderFactory.h).
inline StateData read_state(const std::string& path) {
    const std::string e = detail::lower_ext(path);
    if (e == "xml") return detail::read_xml(path);
    if (e == "json") return detail::read_json(path);
    throw exception::InvalidArgument("State file '" + path + "' is neither .xml nor .json");
}

inline void write_state(const std::string& path, const StateOut& s, bool pretty) {
    const std::string e = detail::lower_ext(path);
    if (e == "xml") return write_xml(path, s, pretty);
    if (e == "json") return write_json(path, s, pretty);
    throw exception::InvalidArgument("State file '" + path + "' is neither .xml nor .json");
}

}  // namespace io
}  // namespace flamegpu
