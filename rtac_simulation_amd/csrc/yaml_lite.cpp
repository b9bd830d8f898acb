#include "yaml_lite.h"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace rtsim {
namespace yaml {

namespace {

struct Line {
    int indent;
    std::string text;   // without indentation and comment
    int number;
};

[[noreturn]] void error(int line, const std::string& msg)
{
    std::ostringstream o;
    o << "yaml line " << line << ": " << msg;
    throw std::runtime_error(o.str());
}

std::string rstrip(const std::string& s)
{
    size_t e = s.find_last_not_of(" \t\r");
    return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}

std::string strip(const std::string& s)
{
    size_t b = s.find_first_not_of(" \t\r");
    if (b == std::string::npos) return std::string();
    return rstrip(s.substr(b));
}

// removes a trailing comment (a '#' at line start or after whitespace,
// outside quotes)
std::string drop_comment(const std::string& s)
{
    char quote = 0;
    for (size_t i = 0; i < s.size(); i++) {
        char c = s[i];
        if (quote) {
            if (c == quote) quote = 0;
        } else if (c == '"' || c == '\'') {
            quote = c;
        } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
            return s.substr(0, i);
        }
    }
    return s;
}

std::vector<Line> split_lines(const std::string& text)
{
    std::vector<Line> out;
    std::istringstream in(text);
    std::string raw;
    int no = 0;
    while (std::getline(in, raw)) {
        no++;
        std::string s = rstrip(drop_comment(raw));
        size_t ind = s.find_first_not_of(' ');
        if (ind == std::string::npos) continue;
        if (s.find('\t') < ind) error(no, "tab indentation is not supported");
        std::string body = s.substr(ind);
        if (body == "---" || body == "...") continue;
        if (body.rfind("--- ", 0) == 0) body = strip(body.substr(4));
        if (body.empty()) continue;
        out.push_back({(int)ind, body, no});
    }
    return out;
}

struct Parser {
    std::vector<Line> lines;
    size_t pos = 0;
    std::map<std::string, NodePtr> anchors;

    NodePtr scalar(const std::string& raw, int line)
    {
        auto n = std::make_shared<Node>();
        n->line = line;
        std::string v = strip(raw);
        if (v.empty() || v == "~" || v == "null") {
            n->kind = Node::Null;
            return n;
        }
        if (v.size() >= 2 && ((v.front() == '"' && v.back() == '"') ||
                              (v.front() == '\'' && v.back() == '\''))) {
            v = v.substr(1, v.size() - 2);
        }
        n->kind = Node::Scalar;
        n->scalar = v;
        return n;
    }

    NodePtr flow_seq(const std::string& raw, int line)
    {
        std::string v = strip(raw);
        if (v.back() != ']') error(line, "unterminated flow sequence");
        auto n = std::make_shared<Node>();
        n->kind = Node::Seq;
        n->line = line;
        std::string inner = v.substr(1, v.size() - 2);
        std::string item;
        int depth = 0;
        for (size_t i = 0; i <= inner.size(); i++) {
            char c = i < inner.size() ? inner[i] : ',';
            if (c == '[') depth++;
            if (c == ']') depth--;
            if (c == ',' && depth == 0) {
                std::string it = strip(item);
                if (!it.empty()) n->seq.push_back(value_inline(it, line));
                item.clear();
            } else {
                item += c;
            }
        }
        return n;
    }

    // an inline value: alias, flow sequence or scalar (optionally anchored)
    NodePtr value_inline(std::string v, int line)
    {
        v = strip(v);
        std::string anchor;
        if (!v.empty() && v[0] == '&') {
            size_t sp = v.find(' ');
            anchor = v.substr(1, sp == std::string::npos ? std::string::npos : sp - 1);
            v = sp == std::string::npos ? std::string() : strip(v.substr(sp));
        }
        NodePtr n;
        if (!v.empty() && v[0] == '*') {
            auto it = anchors.find(v.substr(1));
            if (it == anchors.end()) error(line, "unknown alias '" + v + "'");
            n = it->second;
        } else if (!v.empty() && v[0] == '[') {
            n = flow_seq(v, line);
        } else {
            n = scalar(v, line);
        }
        if (!anchor.empty()) anchors[anchor] = n;
        return n;
    }

    static size_t key_colon(const std::string& s)
    {
        char quote = 0;
        for (size_t i = 0; i < s.size(); i++) {
            char c = s[i];
            if (quote) {
                if (c == quote) quote = 0;
            } else if (c == '"' || c == '\'') {
                quote = c;
            } else if (c == ':' && (i + 1 == s.size() || s[i + 1] == ' ')) {
                return i;
            }
        }
        return std::string::npos;
    }

    static bool is_seq_item(const std::string& s) { return s == "-" || s.rfind("- ", 0) == 0; }

    NodePtr block(int indent)
    {
        if (pos >= lines.size()) return std::make_shared<Node>();
        if (is_seq_item(lines[pos].text)) return sequence(lines[pos].indent);
        return mapping(indent);
    }

    // value after "key:" with nothing inline (or only an anchor)
    NodePtr nested(int parent_indent, const std::string& anchor, int line)
    {
        NodePtr n;
        if (pos < lines.size() &&
            (lines[pos].indent > parent_indent ||
             (lines[pos].indent == parent_indent && is_seq_item(lines[pos].text)))) {
            n = block(lines[pos].indent);
        } else {
            n = std::make_shared<Node>();
            n->line = line;
        }
        if (!anchor.empty()) anchors[anchor] = n;
        return n;
    }

    NodePtr mapping(int indent)
    {
        auto n = std::make_shared<Node>();
        n->kind = Node::Map;
        n->line = pos < lines.size() ? lines[pos].number : 0;
        while (pos < lines.size() && lines[pos].indent == indent &&
               !is_seq_item(lines[pos].text)) {
            const Line L = lines[pos++];
            size_t c = key_colon(L.text);
            if (c == std::string::npos) error(L.number, "expected 'key: value'");
            std::string key = strip(L.text.substr(0, c));
            if (key.size() >= 2 && (key.front() == '"' || key.front() == '\''))
                key = key.substr(1, key.size() - 2);
            std::string rest = strip(L.text.substr(c + 1));
            NodePtr v;
            if (rest.empty()) {
                v = nested(indent, "", L.number);
            } else if (rest[0] == '&' && rest.find(' ') == std::string::npos) {
                v = nested(indent, rest.substr(1), L.number);
            } else {
                v = value_inline(rest, L.number);
            }
            n->map.emplace_back(key, v);
        }
        if (pos < lines.size() && lines[pos].indent > indent)
            error(lines[pos].number, "unexpected indentation");
        return n;
    }

    NodePtr sequence(int indent)
    {
        auto n = std::make_shared<Node>();
        n->kind = Node::Seq;
        n->line = pos < lines.size() ? lines[pos].number : 0;
        while (pos < lines.size() && lines[pos].indent == indent && is_seq_item(lines[pos].text)) {
            Line L = lines[pos];
            std::string rest = L.text == "-" ? std::string() : strip(L.text.substr(2));
            if (rest.empty()) {
                pos++;
                n->seq.push_back(nested(indent, "", L.number));
            } else if (key_colon(rest) != std::string::npos && rest[0] != '[' && rest[0] != '"' &&
                       rest[0] != '\'') {
                // "- key: value": a mapping whose first key sits after the dash
                lines[pos].indent = indent + 2;
                lines[pos].text = rest;
                n->seq.push_back(mapping(indent + 2));
            } else {
                pos++;
                n->seq.push_back(value_inline(rest, L.number));
            }
        }
        return n;
    }
};

}  // namespace

NodePtr Node::get(const std::string& key) const
{
    if (kind != Map) return nullptr;
    for (const auto& kv : map)
        if (kv.first == key) return kv.second;
    return nullptr;
}

double Node::as_double(const std::string& what) const
{
    if (kind != Scalar) {
        std::ostringstream o;
        o << what << " (line " << line << "): expected a number";
        throw std::runtime_error(o.str());
    }
    char* end = nullptr;
    errno = 0;
    double v = std::strtod(scalar.c_str(), &end);
    if (errno || end == scalar.c_str() || *end) {
        std::ostringstream o;
        o << what << " (line " << line << "): '" << scalar << "' is not a number";
        throw std::runtime_error(o.str());
    }
    return v;
}

long Node::as_long(const std::string& what) const
{
    double v = as_double(what);
    // range first: converting a double outside long's range is undefined
    if (!(v >= -9.2e18 && v <= 9.2e18) || v != (double)(long)v) {
        std::ostringstream o;
        o << what << " (line " << line << "): '" << scalar << "' is not an integer";
        throw std::runtime_error(o.str());
    }
    return (long)v;
}

const std::string& Node::as_string(const std::string& what) const
{
    if (kind != Scalar) {
        std::ostringstream o;
        o << what << " (line " << line << "): expected a string";
        throw std::runtime_error(o.str());
    }
    return scalar;
}

NodePtr parse(const std::string& text)
{
    Parser p;
    p.lines = split_lines(text);
    if (p.lines.empty()) return std::make_shared<Node>();
    NodePtr root = p.block(p.lines[0].indent);
    if (p.pos < p.lines.size()) error(p.lines[p.pos].number, "trailing content");
    return root;
}

NodePtr parse_file(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("cannot open '" + path + "'");
    std::ostringstream o;
    o << f.rdbuf();
    return parse(o.str());
}

}  // namespace yaml
}  // namespace rtsim
