// xml.cpp -- see xml.h.
#include "xml.h"

#include <cstring>

#include "hikari.h"

namespace hikari {

namespace {

struct Cursor {
  const std::string &s;
  const std::string &file;
  size_t i = 0;
  int line = 1, col = 1;
  Cursor(const std::string &text, const std::string &f) : s(text), file(f) {}
  bool eof() const { return i >= s.size(); }
  char peek(size_t k = 0) const { return i + k < s.size() ? s[i + k] : '\0'; }
  void adv(size_t n = 1) {
    for (size_t k = 0; k < n && i < s.size(); k++, i++) {
      if (s[i] == '\n') { line++; col = 1; }
      else col++;
    }
  }
  bool starts(const char *p) const { return s.compare(i, std::strlen(p), p) == 0; }
  [[noreturn]] void fail(const std::string &what) const {
    throw HikariException("Error while parsing \"%s\": %s (at line %d, col %d)", file.c_str(), what.c_str(), line,
                          col);
  }
  void skip_ws() {
    while (!eof() && (peek() == ' ' || peek() == '\t' || peek() == '\n' || peek() == '\r')) adv();
  }
  void skip_until(const char *end) {
    while (!eof() && !starts(end)) adv();
    if (eof()) fail(std::string("unterminated construct, expected \"") + end + "\"");
    adv(std::strlen(end));
  }
  // comments, declarations, processing instructions, DOCTYPE
  bool skip_misc() {
    if (starts("<!--")) { skip_until("-->"); return true; }
    if (starts("<?")) { skip_until("?>"); return true; }
    if (starts("<!")) { skip_until(">"); return true; }
    return false;
  }
  std::string name() {
    size_t b = i;
    while (!eof()) {
      char c = peek();
      if (std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == ':' || c == '.') adv();
      else break;
    }
    if (b == i) fail("expected a name");
    return s.substr(b, i - b);
  }
};

std::string unescape(const std::string &v) {
  std::string out;
  for (size_t k = 0; k < v.size(); k++) {
    if (v[k] != '&') { out += v[k]; continue; }
    size_t e = v.find(';', k);
    if (e == std::string::npos) { out += v[k]; continue; }
    std::string ent = v.substr(k + 1, e - k - 1);
    if (ent == "lt") out += '<';
    else if (ent == "gt") out += '>';
    else if (ent == "amp") out += '&';
    else if (ent == "quot") out += '"';
    else if (ent == "apos") out += '\'';
    else { out += v.substr(k, e - k + 1); }
    k = e;
  }
  return out;
}

std::unique_ptr<XmlNode> element(Cursor &c) {
  // at '<'
  std::unique_ptr<XmlNode> n(new XmlNode());
  n->line = c.line;
  n->col = c.col;
  c.adv();
  n->name = c.name();
  for (;;) {
    c.skip_ws();
    if (c.starts("/>")) { c.adv(2); return n; }
    if (c.peek() == '>') { c.adv(); break; }
    std::string key = c.name();
    c.skip_ws();
    if (c.peek() != '=') c.fail("expected '=' after attribute \"" + key + "\"");
    c.adv();
    c.skip_ws();
    char q = c.peek();
    if (q != '"' && q != '\'') c.fail("expected a quoted value for attribute \"" + key + "\"");
    c.adv();
    size_t b = c.i;
    while (!c.eof() && c.peek() != q) c.adv();
    if (c.eof()) c.fail("unterminated attribute value");
    std::string val = c.s.substr(b, c.i - b);
    c.adv();
    if (n->attr(key)) c.fail("duplicate attribute \"" + key + "\"");
    n->attrs.push_back({key, unescape(val)});
  }
  // content
  for (;;) {
    if (c.eof()) c.fail("missing closing tag for <" + n->name + ">");
    if (c.starts("</")) {
      c.adv(2);
      std::string end = c.name();
      if (end != n->name) c.fail("closing tag </" + end + "> does not match <" + n->name + ">");
      c.skip_ws();
      if (c.peek() != '>') c.fail("expected '>'");
      c.adv();
      return n;
    }
    if (c.skip_misc()) continue;
    if (c.starts("<![CDATA[")) { c.skip_until("]]>"); continue; }
    if (c.peek() == '<') {
      n->children.push_back(element(c));
      continue;
    }
    char ch = c.peek();
    if (!(ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r'))
      c.fail("unexpected content inside <" + n->name + ">");
    c.adv();
  }
}

}  // namespace

std::unique_ptr<XmlNode> ParseXml(const std::string &text, const std::string &filename) {
  Cursor c(text, filename);
  std::unique_ptr<XmlNode> root;
  for (;;) {
    c.skip_ws();
    if (c.eof()) break;
    if (c.skip_misc()) continue;
    if (c.peek() != '<') c.fail("unexpected content");
    if (root) c.fail("more than one root element");
    root = element(c);
  }
  if (!root) c.fail("no root element");
  return root;
}

}  // namespace hikari
