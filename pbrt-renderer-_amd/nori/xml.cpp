// Recursive-descent XML reader (the subset Nori scene files use).
#include "xml.h"

#include <cctype>

namespace nori {

namespace {

struct Reader {
    const std::string &s;
    size_t i = 0;
    std::string err;
    size_t err_at = 0;

    explicit Reader(const std::string &t) : s(t) {}
    bool fail(const std::string &m, size_t at) {
        if (err.empty()) {
            err = m;
            err_at = at;
        }
        return false;
    }
    bool eof() const { return i >= s.size(); }
    bool starts(const char *p) const { return s.compare(i, std::char_traits<char>::length(p), p) == 0; }
    void skip_ws() {
        while (!eof() && std::isspace((unsigned char)s[i])) ++i;
    }
    static bool name_char(char c) {
        return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '.' || c == ':';
    }
    bool read_name(std::string &out) {
        size_t b = i;
        while (!eof() && name_char(s[i])) ++i;
        if (i == b) return fail("expected a name", i);
        out = s.substr(b, i - b);
        return true;
    }
    bool decode(const std::string &raw, std::string &out, size_t at) {
        out.clear();
        for (size_t k = 0; k < raw.size(); ++k) {
            if (raw[k] != '&') {
                out += raw[k];
                continue;
            }
            size_t e = raw.find(';', k);
            if (e == std::string::npos) return fail("unterminated entity reference", at + k);
            std::string ent = raw.substr(k + 1, e - k - 1);
            if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "amp") out += '&';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else if (!ent.empty() && ent[0] == '#') {
                long code = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                                                              : std::strtol(ent.c_str() + 1, nullptr, 10);
                if (code <= 0 || code > 127) return fail("unsupported character reference", at + k);
                out += (char)code;
            } else {
                return fail("unknown entity &" + ent + ";", at + k);
            }
            k = e;
        }
        return true;
    }
    bool read_misc(std::vector<std::unique_ptr<XmlNode>> &into, bool &done) {
        // comment, declaration / processing instruction or nothing
        done = false;
        if (starts("<!--")) {
            size_t b = i, e = s.find("-->", i + 4);
            if (e == std::string::npos) return fail("unterminated comment", b);
            auto n = std::make_unique<XmlNode>();
            n->type = XmlNode::Comment;
            n->offset = b;
            n->text = s.substr(b + 4, e - b - 4);
            into.push_back(std::move(n));
            i = e + 3;
            done = true;
        } else if (starts("<?")) {
            size_t b = i, e = s.find("?>", i + 2);
            if (e == std::string::npos) return fail("unterminated declaration", b);
            auto n = std::make_unique<XmlNode>();
            n->type = XmlNode::Declaration;
            n->offset = b;
            n->text = s.substr(b + 2, e - b - 2);
            into.push_back(std::move(n));
            i = e + 2;
            done = true;
        } else if (starts("<!DOCTYPE")) {
            size_t e = s.find('>', i);
            if (e == std::string::npos) return fail("unterminated DOCTYPE", i);
            i = e + 1;
            done = true;
        }
        return true;
    }
    bool read_element(std::unique_ptr<XmlNode> &out) {
        size_t b = i;
        if (eof() || s[i] != '<') return fail("expected '<'", i);
        ++i;
        auto n = std::make_unique<XmlNode>();
        n->offset = b;
        if (!read_name(n->name)) return false;
        while (true) {
            skip_ws();
            if (eof()) return fail("unterminated start tag", b);
            if (starts("/>")) {
                i += 2;
                out = std::move(n);
                return true;
            }
            if (s[i] == '>') {
                ++i;
                break;
            }
            std::string an;
            size_t at = i;
            if (!read_name(an)) return false;
            skip_ws();
            if (eof() || s[i] != '=') return fail("expected '=' after attribute name", i);
            ++i;
            skip_ws();
            if (eof() || (s[i] != '"' && s[i] != '\'')) return fail("expected a quoted attribute value", i);
            char q = s[i++];
            size_t e = s.find(q, i);
            if (e == std::string::npos) return fail("unterminated attribute value", i);
            std::string val;
            if (!decode(s.substr(i, e - i), val, i)) return false;
            i = e + 1;
            if (n->attribute(an)) return fail("duplicate attribute \"" + an + "\"", at);
            n->attributes.emplace_back(an, val);
        }
        // content
        while (true) {
            if (eof()) return fail("unterminated element <" + n->name + ">", b);
            if (starts("</")) {
                size_t cb = i;
                i += 2;
                std::string cn;
                if (!read_name(cn)) return false;
                if (cn != n->name) return fail("end tag </" + cn + "> does not match <" + n->name + ">", cb);
                skip_ws();
                if (eof() || s[i] != '>') return fail("expected '>'", i);
                ++i;
                out = std::move(n);
                return true;
            }
            bool done;
            if (!read_misc(n->children, done)) return false;
            if (done) continue;
            if (s[i] == '<') {
                std::unique_ptr<XmlNode> ch;
                if (!read_element(ch)) return false;
                n->children.push_back(std::move(ch));
                continue;
            }
            size_t tb = i;
            while (!eof() && s[i] != '<') ++i;
            std::string raw = s.substr(tb, i - tb), txt;
            bool blank = true;
            for (char c : raw) blank = blank && std::isspace((unsigned char)c);
            if (!blank) {
                if (!decode(raw, txt, tb)) return false;
                auto t = std::make_unique<XmlNode>();
                t->type = XmlNode::Text;
                t->offset = tb;
                while (tb < s.size() && std::isspace((unsigned char)s[tb])) ++tb;
                t->offset = tb;
                t->text = txt;
                n->children.push_back(std::move(t));
            }
        }
    }
};

} // namespace

bool parseXml(const std::string &text, XmlDocument &doc, std::string &error, size_t &error_offset) {
    Reader r(text);
    // a UTF-8 byte order mark before the document is skipped, as pugixml does
    if (text.size() >= 3 && (unsigned char)text[0] == 0xEF && (unsigned char)text[1] == 0xBB &&
        (unsigned char)text[2] == 0xBF)
        r.i = 3;
    bool have_root = false;
    while (true) {
        r.skip_ws();
        if (r.eof()) break;
        bool done;
        if (!r.read_misc(doc.nodes, done)) break;
        if (done) continue;
        if (text[r.i] != '<') {
            r.fail("unexpected text outside the root element", r.i);
            break;
        }
        if (have_root) {
            r.fail("more than one root element", r.i);
            break;
        }
        std::unique_ptr<XmlNode> root;
        if (!r.read_element(root)) break;
        doc.nodes.push_back(std::move(root));
        have_root = true;
    }
    if (r.err.empty() && !have_root) r.fail("no document element", text.size());
    if (!r.err.empty()) {
        error = r.err;
        error_offset = r.err_at;
        return false;
    }
    return true;
}

std::string xmlPosition(const std::string &text, size_t offset) {
    size_t line = 1, start = 0;
    for (size_t k = 0; k < offset && k < text.size(); ++k)
        if (text[k] == '\n') {
            ++line;
            start = k + 1;
        }
    return "line " + std::to_string(line) + ", col " + std::to_string(offset - start + 1);
}

} // namespace nori
