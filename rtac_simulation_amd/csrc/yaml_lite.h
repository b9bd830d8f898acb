// Minimal YAML reader for sensor configuration files (the subset the
// reference's configs use, examples/*_receiver.yaml): block mappings and
// sequences by indentation, flow sequences [a, b], plain / quoted scalars,
// comments, '---' document markers, anchors (&name) and aliases (*name).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace rtsim {
namespace yaml {

struct Node;
using NodePtr = std::shared_ptr<Node>;

struct Node {
    enum Kind { Null, Scalar, Map, Seq } kind = Null;
    std::string scalar;
    std::vector<std::pair<std::string, NodePtr>> map;
    std::vector<NodePtr> seq;
    int line = 0;

    // child lookup (nullptr when absent or not a mapping)
    NodePtr get(const std::string& key) const;
    bool is_scalar() const { return kind == Scalar; }
    double as_double(const std::string& what) const;   // throws on bad input
    long as_long(const std::string& what) const;
    const std::string& as_string(const std::string& what) const;
};

// Parses a whole document; throws std::runtime_error with a line number.
NodePtr parse(const std::string& text);
NodePtr parse_file(const std::string& path);

}  // namespace yaml
}  // namespace rtsim
