// Template-bank persistence in OpenCV FileStorage's YAML layout, byte-compatible with what the
// reference writes (line2Dup.cpp:58-101 Feature/Template::read/write, 1400-1509 Detector::read/write,
// readClass/writeClass, readClasses/writeClasses).  Files ending in ".gz" are gzip-compressed, as
// FileStorage does; reading accepts plain or gzip files transparently.
#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "sbm_detector.h"

namespace sbm {
namespace yaml {

const Node* Node::get(const std::string& key) const {
    if (kind != kMap) return nullptr;
    for (const auto& kv : map)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

int Node::as_int() const {
    if (kind != kScalar) throw Error(kEIO, "YAML: expected an integer");
    char* end = nullptr;
    const double v = strtod(scalar.c_str(), &end);
    if (end == scalar.c_str()) throw Error(kEIO, "YAML: bad integer '" + scalar + "'");
    return (int)std::lrint(v);
}

double Node::as_real() const {
    if (kind != kScalar) throw Error(kEIO, "YAML: expected a number");
    const std::string& s = scalar;
    if (s == ".Nan" || s == ".nan") return NAN;
    if (s == ".Inf" || s == ".inf") return INFINITY;
    if (s == "-.Inf" || s == "-.inf") return -INFINITY;
    char* end = nullptr;
    const double v = strtod(s.c_str(), &end);
    if (end == s.c_str()) throw Error(kEIO, "YAML: bad number '" + s + "'");
    return v;
}

namespace {
struct Line {
    int indent;
    std::string text;
};

std::string rstrip(const std::string& s) {
    size_t e = s.size();
    while (e > 0 && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r')) --e;
    return s.substr(0, e);
}

std::string unquote(const std::string& s) {
    if (s.size() >= 2 && s.front() == '"' && s.back() == '"') {
        std::string o;
        for (size_t i = 1; i + 1 < s.size(); ++i) {
            char c = s[i];
            if (c == '\\' && i + 2 < s.size()) {
                char n = s[++i];
                if (n == 'n') o += '\n';
                else if (n == 'r') o += '\r';
                else if (n == 't') o += '\t';
                else if (n == 'x' && i + 2 < s.size()) {
                    o += (char)strtol(s.substr(i + 1, 2).c_str(), nullptr, 16);
                    i += 2;
                } else o += n;
            } else {
                o += c;
            }
        }
        return o;
    }
    if (s.size() >= 2 && s.front() == '\'' && s.back() == '\'') return s.substr(1, s.size() - 2);
    return s;
}

// Flow collections: [ a, b ] and { k: v, ... }
struct Flow {
    const std::string& s;
    size_t i = 0;
    explicit Flow(const std::string& str) : s(str) {}
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t' || s[i] == '\r')) ++i;
    }
    Node value() {
        ws();
        if (i >= s.size()) throw Error(kEIO, "YAML: truncated flow collection");
        Node n;
        if (s[i] == '[') {
            n.kind = Node::kSeq;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') { ++i; return n; }
            for (;;) {
                n.seq.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == ']') { ++i; break; }
                throw Error(kEIO, "YAML: bad flow sequence");
            }
            return n;
        }
        if (s[i] == '{') {
            n.kind = Node::kMap;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') { ++i; return n; }
            for (;;) {
                ws();
                size_t k0 = i;
                while (i < s.size() && s[i] != ':') ++i;
                std::string key = unquote(rstrip(s.substr(k0, i - k0)));
                ++i;
                n.map.emplace_back(key, value());
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == '}') { ++i; break; }
                throw Error(kEIO, "YAML: bad flow map");
            }
            return n;
        }
        n.kind = Node::kScalar;
        if (s[i] == '"') {
            size_t j = i + 1;
            while (j < s.size() && s[j] != '"') j += (s[j] == '\\') ? 2 : 1;
            n.scalar = unquote(s.substr(i, j + 1 - i));
            i = j + 1;
            return n;
        }
        size_t j = i;
        while (j < s.size() && s[j] != ',' && s[j] != ']' && s[j] != '}') ++j;
        n.scalar = unquote(rstrip(s.substr(i, j - i)));
        i = j;
        return n;
    }
};

int bracket_balance(const std::string& s) {
    int b = 0;
    bool q = false;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (c == '"' && (i == 0 || s[i - 1] != '\\')) q = !q;
        if (q) continue;
        if (c == '[' || c == '{') ++b;
        if (c == ']' || c == '}') --b;
    }
    return b;
}

struct Parser {
    std::vector<Line> lines;
    size_t pos = 0;

    // value text that may continue on the following lines (multi-line flow collection)
    Node inline_value(std::string text) {
        if (!text.empty() && (text[0] == '[' || text[0] == '{')) {
            while (bracket_balance(text) > 0 && pos < lines.size()) text += " " + lines[pos++].text;
            Flow f(text);
            return f.value();
        }
        Node n;
        n.kind = Node::kScalar;
        n.scalar = unquote(text);
        return n;
    }

    Node block(int indent) {
        Node n;
        if (pos >= lines.size()) return n;
        const Line& first = lines[pos];
        if (first.text[0] == '-' && (first.text.size() == 1 || first.text[1] == ' ')) {
            n.kind = Node::kSeq;
            while (pos < lines.size() && lines[pos].indent == indent && lines[pos].text[0] == '-' &&
                   (lines[pos].text.size() == 1 || lines[pos].text[1] == ' ')) {
                std::string rest = lines[pos].text.size() > 1 ? lines[pos].text.substr(2) : std::string();
                size_t k = 0;
                while (k < rest.size() && rest[k] == ' ') ++k;
                rest = rest.substr(k);
                if (rest.empty()) {
                    ++pos;
                    if (pos < lines.size() && lines[pos].indent > indent) n.seq.push_back(block(lines[pos].indent));
                    else n.seq.push_back(Node());
                } else if (rest[0] == '[' || rest[0] == '{' || rest.find(": ") == std::string::npos) {
                    if (rest[0] != '[' && rest[0] != '{' && rest.back() == ':') {
                        // "- key:" opens an inline map
                        lines[pos].indent = indent + 2 + (int)k;
                        lines[pos].text = rest;
                        n.seq.push_back(block(lines[pos].indent));
                    } else {
                        ++pos;
                        n.seq.push_back(inline_value(rest));
                    }
                } else {
                    lines[pos].indent = indent + 2 + (int)k;  // "- key: value" starts an inline map
                    lines[pos].text = rest;
                    n.seq.push_back(block(lines[pos].indent));
                }
            }
            return n;
        }
        n.kind = Node::kMap;
        while (pos < lines.size() && lines[pos].indent == indent) {
            const std::string& t = lines[pos].text;
            size_t c = std::string::npos;
            bool q = false;
            for (size_t i = 0; i < t.size(); ++i) {
                if (t[i] == '"') q = !q;
                if (!q && t[i] == ':' && (i + 1 == t.size() || t[i + 1] == ' ')) { c = i; break; }
            }
            if (c == std::string::npos) throw Error(kEIO, "YAML: expected 'key: value' in '" + t + "'");
            std::string key = unquote(rstrip(t.substr(0, c)));
            std::string val = c + 1 < t.size() ? t.substr(c + 1) : std::string();
            size_t k = 0;
            while (k < val.size() && val[k] == ' ') ++k;
            val = val.substr(k);
            ++pos;
            if (val.empty()) {
                if (pos < lines.size() &&
                    (lines[pos].indent > indent || (lines[pos].indent == indent && lines[pos].text[0] == '-')))
                    n.map.emplace_back(key, block(lines[pos].indent));
                else
                    n.map.emplace_back(key, Node());
            } else {
                n.map.emplace_back(key, inline_value(val));
            }
        }
        return n;
    }
};
}  // namespace

Node parse(const std::string& text) {
    Parser p;
    size_t i = 0;
    while (i <= text.size()) {
        size_t e = text.find('\n', i);
        if (e == std::string::npos) e = text.size();
        std::string raw = rstrip(text.substr(i, e - i));
        i = e + 1;
        size_t ind = 0;
        while (ind < raw.size() && raw[ind] == ' ') ++ind;
        std::string body = raw.substr(ind);
        if (body.empty() || body[0] == '#' || body[0] == '%' || body == "---" || body == "...") {
            if (e == text.size()) break;
            continue;
        }
        p.lines.push_back({(int)ind, body});
        if (e == text.size()) break;
    }
    if (p.lines.empty()) return Node();
    return p.block(p.lines[0].indent);
}

std::string read_file(const std::string& path) {
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) throw Error(kEIO, "cannot open " + path);
    std::string out;
    char buf[1 << 16];
    for (;;) {
        const int n = gzread(f, buf, sizeof(buf));
        if (n < 0) {
            gzclose(f);
            throw Error(kEIO, "read error in " + path);
        }
        if (n == 0) break;
        out.append(buf, (size_t)n);
    }
    gzclose(f);
    return out;
}

void write_file(const std::string& path, const std::string& data) {
    const bool gz = path.size() >= 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
    if (gz) {
        gzFile f = gzopen(path.c_str(), "wb");
        if (!f) throw Error(kEIO, "cannot open " + path + " for writing");
        const int n = gzwrite(f, data.data(), (unsigned)data.size());
        gzclose(f);
        if (n != (int)data.size()) throw Error(kEIO, "write error in " + path);
    } else {
        FILE* f = fopen(path.c_str(), "wb");
        if (!f) throw Error(kEIO, "cannot open " + path + " for writing");
        const size_t n = fwrite(data.data(), 1, data.size(), f);
        fclose(f);
        if (n != data.size()) throw Error(kEIO, "write error in " + path);
    }
}

// String scalars as FileStorage's YAML writer emits them: quoted when empty, when starting with a
// digit, sign or dot, or when holding characters outside [A-Za-z0-9_ -()/+;].
std::string quote_string(const std::string& s) {
    const size_t len = s.size();
    if (len >= 2 && s[0] == s[len - 1] && (s[0] == '"' || s[0] == '\'')) return s;
    bool need = len == 0 || s[0] == ' ';
    std::string body;
    for (char c : s) {
        const bool alnum = std::isalnum((unsigned char)c) != 0;
        if (!need && !alnum && c != '_' && c != ' ' && c != '-' && c != '(' && c != ')' && c != '/' && c != '+' && c != ';')
            need = true;
        if (!alnum && (!std::isprint((unsigned char)c) || c == '\\' || c == '\'' || c == '"')) {
            body += '\\';
            if (std::isprint((unsigned char)c)) body += c;
            else if (c == '\n') body += 'n';
            else if (c == '\r') body += 'r';
            else if (c == '\t') body += 't';
            else {
                char b[8];
                snprintf(b, sizeof(b), "x%02x", (unsigned char)c);
                body += b;
            }
        } else {
            body += c;
        }
    }
    if (!need && len > 0 && (std::isdigit((unsigned char)s[0]) || s[0] == '+' || s[0] == '-' || s[0] == '.')) need = true;
    return need ? "\"" + body + "\"" : body;
}

// Reals as FileStorage writes doubles: "%d." for integral values, else "%.16e"; .Nan / .Inf.
std::string real_to_string(double v) {
    char buf[64];
    if (std::isnan(v)) return ".Nan";
    if (std::isinf(v)) return v > 0 ? ".Inf" : "-.Inf";
    const double r = std::nearbyint(v);
    if (r == v && std::fabs(v) < 2147483647.0) {
        snprintf(buf, sizeof(buf), "%d.", (int)r);
    } else {
        snprintf(buf, sizeof(buf), "%.16e", v);
    }
    return buf;
}

}  // namespace yaml

// ---------------------------------------------------------------------------------------------
// Detector persistence
// ---------------------------------------------------------------------------------------------
static const char* kHeader = "%YAML:1.0\n---\n";

// writeClass (line2Dup.cpp:1458-1485) with Template::write / Feature::write (line2Dup.cpp:64-101)
std::string Detector::write_class_text(const std::string& class_id) const {
    auto it = class_templates.find(class_id);
    if (it == class_templates.end()) throw Error(kENotFound, "class not found: " + class_id);
    std::string o = kHeader;
    o += "class_id: " + yaml::quote_string(it->first) + "\n";
    o += "pyramid_levels: " + std::to_string(pyramid_levels) + "\n";
    o += "template_pyramids:\n";
    if (it->second.empty()) o += "   []\n";
    char buf[96];
    for (size_t i = 0; i < it->second.size(); ++i) {
        const TemplatePyramid& tp = it->second[i];
        o += "   -\n";
        o += "      template_id: " + std::to_string((int)i) + "\n";
        o += "      templates:\n";
        if (tp.empty()) o += "         []\n";
        for (const Template& t : tp) {
            o += "         -\n";
            o += "            width: " + std::to_string(t.width) + "\n";
            o += "            height: " + std::to_string(t.height) + "\n";
            o += "            tl_x: " + std::to_string(t.tl_x) + "\n";
            o += "            tl_y: " + std::to_string(t.tl_y) + "\n";
            o += "            pyramid_level: " + std::to_string(t.pyramid_level) + "\n";
            o += "            features:\n";
            if (t.features.empty()) o += "               []\n";
            for (const Feature& f : t.features) {
                snprintf(buf, sizeof(buf), "               - [ %d, %d, %d ]\n", f.x, f.y, f.label);
                o += buf;
            }
        }
    }
    return o;
}

// readClass (line2Dup.cpp:1417-1456)
std::string Detector::read_class_text(const std::string& text, const std::string& class_id_override) {
    const yaml::Node root = yaml::parse(text);
    std::string class_id;
    if (class_id_override.empty()) {
        const yaml::Node* c = root.get("class_id");
        class_id = (c && c->kind == yaml::Node::kScalar) ? c->scalar : std::string();
        if (class_templates.count(class_id)) throw Error(kEDupClass, "class already loaded: " + class_id);
    } else {
        class_id = class_id_override;
    }
    std::vector<TemplatePyramid> tps;
    const yaml::Node* tpn = root.get("template_pyramids");
    if (tpn && tpn->kind == yaml::Node::kSeq) {
        tps.resize(tpn->seq.size());
        int expected = 0;
        for (const yaml::Node& p : tpn->seq) {
            const yaml::Node* idn = p.get("template_id");
            const int tid = idn ? idn->as_int() : -1;
            if (tid != expected) throw Error(kEIO, "template_id out of order");
            const yaml::Node* tn = p.get("templates");
            if (tn && tn->kind == yaml::Node::kSeq) {
                for (const yaml::Node& t : tn->seq) {
                    Template te;
                    auto geti = [&](const char* k) {
                        const yaml::Node* v = t.get(k);
                        return v ? v->as_int() : 0;
                    };
                    te.width = geti("width");
                    te.height = geti("height");
                    te.tl_x = geti("tl_x");
                    te.tl_y = geti("tl_y");
                    te.pyramid_level = geti("pyramid_level");
                    const yaml::Node* fn = t.get("features");
                    if (fn && fn->kind == yaml::Node::kSeq) {
                        for (const yaml::Node& f : fn->seq) {
                            if (f.kind != yaml::Node::kSeq || f.seq.size() < 3) throw Error(kEIO, "bad feature record");
                            Feature fe;
                            fe.x = f.seq[0].as_int();
                            fe.y = f.seq[1].as_int();
                            fe.label = f.seq[2].as_int();
                            te.features.push_back(fe);
                        }
                    }
                    tps[tid].push_back(te);
                }
            }
            ++expected;
        }
    }
    class_templates.emplace(class_id, std::move(tps));  // std::map::insert keeps an existing class
    bank_dirty_ = true;
    return class_id;
}

static std::string format_path(const std::string& format, const std::string& class_id) {
    const size_t p = format.find("%s");
    if (p == std::string::npos) return format;
    return format.substr(0, p) + class_id + format.substr(p + 2);
}

void Detector::read_classes(const std::vector<std::string>& class_ids, const std::string& format) {
    for (const auto& c : class_ids) read_class_text(yaml::read_file(format_path(format, c)), "");
}

void Detector::write_classes(const std::string& format) const {
    for (const auto& kv : class_templates) yaml::write_file(format_path(format, kv.first), write_class_text(kv.first));
}

// Detector::write (line2Dup.cpp:1409-1415) with ColorGradient::write (line2Dup.cpp:545-551)
std::string Detector::write_detector_text() const {
    std::string o = kHeader;
    o += "pyramid_levels: " + std::to_string(pyramid_levels) + "\n";
    o += "T: [";
    for (size_t i = 0; i < T_at_level.size(); ++i) o += (i ? ", " : " ") + std::to_string(T_at_level[i]);
    o += T_at_level.empty() ? "]\n" : " ]\n";
    o += "type: ColorGradient\n";
    o += "weak_threshold: " + yaml::real_to_string((double)weak_threshold) + "\n";
    o += "num_features: " + std::to_string(num_features) + "\n";
    o += "strong_threshold: " + yaml::real_to_string((double)strong_threshold) + "\n";
    return o;
}

// Detector::read (line2Dup.cpp:1400-1407): clears the bank, reads levels and T, resets the modality.
void Detector::read_detector_text(const std::string& text) {
    const yaml::Node root = yaml::parse(text);
    class_templates.clear();
    const yaml::Node* pl = root.get("pyramid_levels");
    const yaml::Node* tn = root.get("T");
    if (!pl || !tn || tn->kind != yaml::Node::kSeq) throw Error(kEIO, "detector YAML lacks pyramid_levels / T");
    std::vector<int> T;
    for (const auto& v : tn->seq) T.push_back(v.as_int());
    if (pl->as_int() != (int)T.size() || T.empty() || T.size() > 8)
        throw Error(kEIO, "detector YAML: pyramid_levels must equal the size of T (1 to 8 levels)");
    for (int v : T)
        if (v < 1 || v > 16) throw Error(kEIO, "detector YAML: T must be in [1, 16]");
    pyramid_levels = (int)T.size();
    T_at_level = T;
    weak_threshold = 30.f;
    num_features = 63;
    strong_threshold = 60.f;
    bank_dirty_ = true;
}

}  // namespace sbm
