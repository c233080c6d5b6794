// Minimal XML reader for Nori scene files: elements, attributes, comments,
// processing instructions / declarations, CDATA-free text.  Every node keeps
// its byte offset so that errors can name a line and column.
#pragma once
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace nori {

struct XmlNode {
    enum Type { Element, Comment, Declaration, Text };
    Type type = Element;
    std::string name;                                         // element name
    std::vector<std::pair<std::string, std::string>> attributes;
    std::vector<std::unique_ptr<XmlNode>> children;
    std::string text;                                         // Text / Comment content
    size_t offset = 0;                                        // byte offset of '<' (or of the text)

    const std::string *attribute(const std::string &n) const {
        for (auto &a : attributes)
            if (a.first == n) return &a.second;
        return nullptr;
    }
};

struct XmlDocument {
    std::vector<std::unique_ptr<XmlNode>> nodes; // top level (declarations, comments, the root)
};

/// Parses `text`; on error sets `error` (a description) and `error_offset`.
bool parseXml(const std::string &text, XmlDocument &doc, std::string &error, size_t &error_offset);

/// "line L, col C" of a byte offset in `text` (1-based line, column counted from the line start).
std::string xmlPosition(const std::string &text, size_t offset);

} // namespace nori
