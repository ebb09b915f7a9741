// rt_xml.h -- the XML reader behind the Scheme loader (no third-party
// library).
//
// The reference reads scheme files with mxml, loading every text run as one
// "opaque" string (MXML_OPAQUE_CALLBACK, common-lib/scheme.cpp:310): the
// characters between two tags, whitespace included, become one text node of
// the enclosing element.  A scheme attribute is a child *element* whose text
// is the value (<class>conic</class>), found with mxmlFindPath from the
// surface, beam or general element (scheme.cpp:130-149): the first matching
// descendant in document order, whose first child must be text.  This reader
// builds the same tree shape (elements, text runs with the five predefined
// entities and numeric character references decoded, CDATA sections as text)
// and answers the same queries; declarations, processing instructions and
// comments are skipped.
#ifndef RT_XML_H
#define RT_XML_H

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rtxml {

struct Node {
    bool is_text = false;
    std::string name;                       // element name (elements)
    std::string text;                       // character data (text nodes)
    std::vector<std::unique_ptr<Node>> children;
};

class ParseError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// The document's nodes as children of an unnamed root; throws ParseError on
// malformed markup (unbalanced or mismatched tags, an unterminated tag,
// comment, CDATA section or reference).
std::unique_ptr<Node> parse(const std::string& doc);
std::unique_ptr<Node> parse_file(const std::string& path);   // throws ParseError if unreadable too

// Every element named `name` below `root`, in document order
// (mxmlIndexNew over the tree with no attribute key).
std::vector<const Node*> elements(const Node* root, const std::string& name);

// mxmlFindPath(node, path): the first descendant element matching the
// '/'-separated path ("a/b": b a direct child of the first descendant a).
const Node* find_path(const Node* node, const std::string& path);

// mxmlGetOpaque of what mxmlFindPath returns: the text of the element's
// first child when that child is a text node, else nullptr.
const std::string* value(const Node* element);

}  // namespace rtxml

#endif  // RT_XML_H
