// xml.h -- minimal XML reader for scene files: elements, attributes,
// comments, <?xml ...?> declarations, character data (ignored).  Enough for
// the reference's scene format (src/core/Parser.cpp reads it with pugixml);
// every node keeps its line and column for error messages.
#pragma once
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace hikari {

struct XmlNode {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::unique_ptr<XmlNode>> children;
  int line = 0, col = 0;
  const std::string *attr(const std::string &key) const {
    for (auto &a : attrs)
      if (a.first == key) return &a.second;
    return nullptr;
  }
};

// Parses a whole document; throws HikariException with line/column on
// malformed input.  Returns the root element.
std::unique_ptr<XmlNode> ParseXml(const std::string &text, const std::string &filename);

}  // namespace hikari
