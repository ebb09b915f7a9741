// rt_xml.cpp -- a small XML reader for scheme files (see rt_xml.h).
#include "rt_xml.h"

#include <string.h>

#include <fstream>
#include <sstream>

namespace rtxml {

namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

bool is_name_char(char c) {
    return !is_space(c) && c != '>' && c != '/' && c != '=' && c != '<' && c != '\0';
}

void put_utf8(std::string& out, unsigned long cp) {
    if (cp < 0x80) {
        out += (char)cp;
    } else if (cp < 0x800) {
        out += (char)(0xC0 | (cp >> 6));
        out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
        out += (char)(0xE0 | (cp >> 12));
        out += (char)(0x80 | ((cp >> 6) & 0x3F));
        out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x110000) {
        out += (char)(0xF0 | (cp >> 18));
        out += (char)(0x80 | ((cp >> 12) & 0x3F));
        out += (char)(0x80 | ((cp >> 6) & 0x3F));
        out += (char)(0x80 | (cp & 0x3F));
    } else {
        throw ParseError("character reference out of range");
    }
}

class Reader {
public:
    explicit Reader(const std::string& d) : doc(d), pos(0) {}

    std::unique_ptr<Node> run() {
        auto root = std::make_unique<Node>();
        std::vector<Node*> stack{root.get()};
        std::string text;
        while (pos < doc.size()) {
            const char c = doc[pos];
            if (c == '<') {
                if (starts("<![CDATA[")) {
                    const size_t end = doc.find("]]>", pos + 9);
                    if (end == std::string::npos) throw ParseError("unterminated CDATA section");
                    text.append(doc, pos + 9, end - pos - 9);
                    pos = end + 3;
                    continue;
                }
                flush(text, stack.back());
                if (starts("<!--")) {
                    const size_t end = doc.find("-->", pos + 4);
                    if (end == std::string::npos) throw ParseError("unterminated comment");
                    pos = end + 3;
                } else if (starts("<?")) {
                    const size_t end = doc.find("?>", pos + 2);
                    if (end == std::string::npos) throw ParseError("unterminated processing instruction");
                    pos = end + 2;
                } else if (starts("<!")) {
                    skip_declaration();
                } else if (starts("</")) {
                    pos += 2;
                    const std::string name = read_name();
                    while (pos < doc.size() && is_space(doc[pos])) ++pos;
                    if (pos >= doc.size() || doc[pos] != '>') throw ParseError("malformed end tag </" + name);
                    ++pos;
                    if (stack.size() < 2 || stack.back()->name != name)
                        throw ParseError("end tag </" + name + "> does not match the open element");
                    stack.pop_back();
                } else {
                    ++pos;
                    auto el = std::make_unique<Node>();
                    el->name = read_name();
                    if (el->name.empty()) throw ParseError("element without a name");
                    const bool empty = skip_attributes();
                    Node* raw = el.get();
                    stack.back()->children.push_back(std::move(el));
                    if (!empty) stack.push_back(raw);
                }
            } else if (c == '&') {
                text += read_reference();
            } else {
                text += c;
                ++pos;
            }
        }
        flush(text, stack.back());
        if (stack.size() != 1) throw ParseError("element <" + stack.back()->name + "> is not closed");
        return root;
    }

private:
    const std::string& doc;
    size_t pos;

    bool starts(const char* s) const { return doc.compare(pos, strlen(s), s) == 0; }

    static void flush(std::string& text, Node* parent) {
        if (text.empty()) return;
        auto t = std::make_unique<Node>();
        t->is_text = true;
        t->text.swap(text);
        parent->children.push_back(std::move(t));
        text.clear();
    }

    std::string read_name() {
        const size_t b = pos;
        while (pos < doc.size() && is_name_char(doc[pos])) ++pos;
        return doc.substr(b, pos - b);
    }

    // attributes of a start tag (the scheme format carries none that matter);
    // true for an empty-element tag <x/>
    bool skip_attributes() {
        for (;;) {
            while (pos < doc.size() && is_space(doc[pos])) ++pos;
            if (pos >= doc.size()) throw ParseError("unterminated start tag");
            if (doc[pos] == '>') {
                ++pos;
                return false;
            }
            if (doc[pos] == '/') {
                if (pos + 1 < doc.size() && doc[pos + 1] == '>') {
                    pos += 2;
                    return true;
                }
                throw ParseError("stray '/' in a start tag");
            }
            if (read_name().empty()) throw ParseError("malformed attribute");
            while (pos < doc.size() && is_space(doc[pos])) ++pos;
            if (pos < doc.size() && doc[pos] == '=') {
                ++pos;
                while (pos < doc.size() && is_space(doc[pos])) ++pos;
                if (pos >= doc.size() || (doc[pos] != '"' && doc[pos] != '\''))
                    throw ParseError("attribute value must be quoted");
                const char q = doc[pos];
                const size_t end = doc.find(q, pos + 1);
                if (end == std::string::npos) throw ParseError("unterminated attribute value");
                pos = end + 1;
            }
        }
    }

    // <!DOCTYPE ...> with an optional internal subset in brackets
    void skip_declaration() {
        int depth = 0;
        for (pos += 2; pos < doc.size(); ++pos) {
            const char c = doc[pos];
            if (c == '[') ++depth;
            if (c == ']') --depth;
            if (c == '>' && depth <= 0) {
                ++pos;
                return;
            }
        }
        throw ParseError("unterminated declaration");
    }

    std::string read_reference() {
        const size_t end = doc.find(';', pos);
        if (end == std::string::npos || end - pos > 12) throw ParseError("unterminated character reference");
        const std::string ent = doc.substr(pos + 1, end - pos - 1);
        pos = end + 1;
        if (ent == "lt") return "<";
        if (ent == "gt") return ">";
        if (ent == "amp") return "&";
        if (ent == "quot") return "\"";
        if (ent == "apos") return "'";
        if (ent.size() > 1 && ent[0] == '#') {
            const bool hex = ent[1] == 'x' || ent[1] == 'X';
            const char* s = ent.c_str() + (hex ? 2 : 1);
            char* e = nullptr;
            const unsigned long cp = strtoul(s, &e, hex ? 16 : 10);
            if (e == s || *e != '\0') throw ParseError("bad character reference &" + ent + ";");
            std::string out;
            put_utf8(out, cp);
            return out;
        }
        throw ParseError("unknown entity &" + ent + ";");
    }
};

void collect(const Node* n, const std::string& name, std::vector<const Node*>& out) {
    for (const auto& c : n->children) {
        if (c->is_text) continue;
        if (c->name == name) out.push_back(c.get());
        collect(c.get(), name, out);
    }
}

const Node* first_descendant(const Node* n, const std::string& name) {
    for (const auto& c : n->children) {
        if (c->is_text) continue;
        if (c->name == name) return c.get();
        if (const Node* d = first_descendant(c.get(), name)) return d;
    }
    return nullptr;
}

const Node* first_child(const Node* n, const std::string& name) {
    for (const auto& c : n->children)
        if (!c->is_text && c->name == name) return c.get();
    return nullptr;
}

}  // namespace

std::unique_ptr<Node> parse(const std::string& doc) { return Reader(doc).run(); }

std::unique_ptr<Node> parse_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw ParseError("cannot open " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return parse(ss.str());
}

std::vector<const Node*> elements(const Node* root, const std::string& name) {
    std::vector<const Node*> out;
    if (root) collect(root, name, out);
    return out;
}

const Node* find_path(const Node* node, const std::string& path) {
    if (!node || path.empty()) return nullptr;
    size_t b = 0;
    const Node* cur = node;
    bool first = true;
    while (b <= path.size()) {
        size_t e = path.find('/', b);
        if (e == std::string::npos) e = path.size();
        const std::string part = path.substr(b, e - b);
        cur = first ? first_descendant(cur, part) : first_child(cur, part);
        if (!cur) return nullptr;
        first = false;
        b = e + 1;
    }
    return cur;
}

const std::string* value(const Node* element) {
    if (!element || element->children.empty()) return nullptr;
    const Node* c = element->children.front().get();
    return c->is_text ? &c->text : nullptr;
}

}  // namespace rtxml
