// A small XML reader for the model loaders (MJCF, URDF): elements,
// attributes and the five predefined entities; comments, processing
// instructions, DOCTYPE and CDATA are skipped; text content is ignored.
#pragma once
#include <cctype>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace phx {

struct XNode {
  std::string tag;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::unique_ptr<XNode>> kids;
  const char* get(const char* k) const {
    for (auto& a : attrs)
      if (a.first == k) return a.second.c_str();
    return nullptr;
  }
};

struct XmlError {
  std::string msg;
};

class XmlParser {
 public:
  explicit XmlParser(const std::string& s) : s_(s), i_(0) {}
  std::unique_ptr<XNode> parse() {
    skip_misc();
    if (i_ >= s_.size() || s_[i_] != '<') throw XmlError{"expected a root element"};
    auto root = element();
    return root;
  }

 private:
  const std::string& s_;
  size_t i_;
  bool starts(const char* p) const { return s_.compare(i_, strlen(p), p) == 0; }
  void ws() {
    while (i_ < s_.size() && isspace((unsigned char)s_[i_])) i_++;
  }
  void skip_until(const char* end) {
    size_t j = s_.find(end, i_);
    if (j == std::string::npos) throw XmlError{std::string("unterminated construct, expected ") + end};
    i_ = j + strlen(end);
  }
  void skip_misc() {
    for (;;) {
      ws();
      if (starts("<?")) skip_until("?>");
      else if (starts("<!--")) skip_until("-->");
      else if (starts("<!")) skip_until(">");
      else break;
    }
  }
  std::string name() {
    size_t j = i_;
    while (j < s_.size() && (isalnum((unsigned char)s_[j]) || s_[j] == '_' || s_[j] == '-' || s_[j] == ':' || s_[j] == '.'))
      j++;
    if (j == i_) throw XmlError{"expected a name"};
    std::string n = s_.substr(i_, j - i_);
    i_ = j;
    return n;
  }
  static std::string unescape(const std::string& v) {
    std::string o;
    for (size_t k = 0; k < v.size(); k++) {
      if (v[k] != '&') {
        o += v[k];
        continue;
      }
      size_t e = v.find(';', k);
      std::string ent = e == std::string::npos ? "" : v.substr(k + 1, e - k - 1);
      if (ent == "lt") o += '<';
      else if (ent == "gt") o += '>';
      else if (ent == "amp") o += '&';
      else if (ent == "quot") o += '"';
      else if (ent == "apos") o += '\'';
      else {
        o += '&';
        continue;
      }
      k = e;
    }
    return o;
  }
  std::unique_ptr<XNode> element() {
    i_++;  // '<'
    auto n = std::make_unique<XNode>();
    n->tag = name();
    for (;;) {
      ws();
      if (i_ >= s_.size()) throw XmlError{"unexpected end of document"};
      if (starts("/>")) {
        i_ += 2;
        return n;
      }
      if (s_[i_] == '>') {
        i_++;
        break;
      }
      std::string k = name();
      ws();
      if (i_ >= s_.size() || s_[i_] != '=') throw XmlError{"expected '=' after attribute " + k};
      i_++;
      ws();
      char q = s_[i_];
      if (q != '"' && q != '\'') throw XmlError{"expected a quoted attribute value"};
      size_t e = s_.find(q, i_ + 1);
      if (e == std::string::npos) throw XmlError{"unterminated attribute value"};
      n->attrs.push_back({k, unescape(s_.substr(i_ + 1, e - i_ - 1))});
      i_ = e + 1;
    }
    for (;;) {
      size_t lt = s_.find('<', i_);
      if (lt == std::string::npos) throw XmlError{"unterminated element <" + n->tag + ">"};
      i_ = lt;
      if (starts("<!--")) skip_until("-->");
      else if (starts("<![CDATA[")) skip_until("]]>");
      else if (starts("<?")) skip_until("?>");
      else if (starts("</")) {
        i_ += 2;
        std::string c = name();
        if (c != n->tag) throw XmlError{"mismatched closing tag </" + c + "> for <" + n->tag + ">"};
        ws();
        if (i_ >= s_.size() || s_[i_] != '>') throw XmlError{"malformed closing tag"};
        i_++;
        return n;
      } else {
        n->kids.push_back(element());
      }
    }
  }
};

}  // namespace phx
