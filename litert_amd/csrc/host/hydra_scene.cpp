// hydra_scene.cpp - reader of the Hydra XML scene parts LoadSceneHydra uses.
#include "hydra_scene.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>

namespace lrt {
namespace {

struct Node {
  std::string name, text;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::unique_ptr<Node>> kids;
  const std::string* attr(const char* k) const {
    for (const auto& a : attrs)
      if (a.first == k) return &a.second;
    return nullptr;
  }
  const Node* child(const char* n) const {
    for (const auto& c : kids)
      if (c->name == n) return c.get();
    return nullptr;
  }
};

class Reader {
 public:
  explicit Reader(const std::string& s) : s_(s) {}

  std::unique_ptr<Node> document() {
    auto root = std::make_unique<Node>();
    root->name = "#document";
    content(*root, true);
    return root;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;

  [[noreturn]] void fail(const std::string& what) const {
    throw std::runtime_error("malformed scene XML at byte " + std::to_string(i_) + ": " + what);
  }
  bool starts(const char* p) const { return s_.compare(i_, std::char_traits<char>::length(p), p) == 0; }
  void skip_until(const char* end) {
    const size_t k = s_.find(end, i_);
    if (k == std::string::npos) fail(std::string("missing ") + end);
    i_ = k + std::char_traits<char>::length(end);
  }
  void spaces() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
  }
  std::string name() {
    const size_t b = i_;
    while (i_ < s_.size() && !std::isspace((unsigned char)s_[i_]) && s_[i_] != '>' && s_[i_] != '/' && s_[i_] != '=') i_++;
    if (i_ == b) fail("expected a name");
    return s_.substr(b, i_ - b);
  }
  static std::string unescape(const std::string& v) {
    std::string o;
    for (size_t k = 0; k < v.size(); k++) {
      if (v[k] != '&') {
        o += v[k];
        continue;
      }
      const size_t e = v.find(';', k);
      const std::string ent = e == std::string::npos ? "" : v.substr(k + 1, e - k - 1);
      const char* rep = ent == "lt" ? "<" : ent == "gt" ? ">" : ent == "amp" ? "&" : ent == "quot" ? "\"" : ent == "apos" ? "'" : nullptr;
      if (!rep) {
        o += v[k];
        continue;
      }
      o += rep;
      k = e;
    }
    return o;
  }
  // children and text of `n` up to its end tag (or the end of the input for the document)
  void content(Node& n, bool document) {
    for (;;) {
      const size_t b = i_;
      while (i_ < s_.size() && s_[i_] != '<') i_++;
      n.text += unescape(s_.substr(b, i_ - b));
      if (i_ >= s_.size()) {
        if (!document) fail("missing </" + n.name + ">");
        return;
      }
      if (starts("<?")) {
        skip_until("?>");
      } else if (starts("<!--")) {
        skip_until("-->");
      } else if (starts("<!")) {
        skip_until(">");
      } else if (starts("</")) {
        i_ += 2;
        const std::string end = name();
        spaces();
        if (i_ >= s_.size() || s_[i_] != '>') fail("bad end tag");
        i_++;
        if (document || end != n.name) fail("unexpected </" + end + ">");
        return;
      } else {
        i_++;
        auto c = std::make_unique<Node>();
        c->name = name();
        for (;;) {
          spaces();
          if (i_ >= s_.size()) fail("unterminated tag");
          if (s_[i_] == '/') {
            if (i_ + 1 >= s_.size() || s_[i_ + 1] != '>') fail("bad empty tag");
            i_ += 2;
            break;
          }
          if (s_[i_] == '>') {
            i_++;
            content(*c, false);
            break;
          }
          std::string k = name();
          spaces();
          if (i_ >= s_.size() || s_[i_] != '=') fail("expected = after " + k);
          i_++;
          spaces();
          if (i_ >= s_.size() || (s_[i_] != '"' && s_[i_] != '\'')) fail("expected a quoted value");
          const char q = s_[i_++];
          const size_t e = s_.find(q, i_);
          if (e == std::string::npos) fail("unterminated value");
          c->attrs.emplace_back(k, unescape(s_.substr(i_, e - i_)));
          i_ = e + 1;
        }
        n.kids.push_back(std::move(c));
      }
    }
  }
};

void floats(const std::string& s, float* out, int n, const char* what) {
  std::istringstream in(s);
  for (int k = 0; k < n; k++)
    if (!(in >> out[k])) throw std::runtime_error(std::string("scene XML: expected ") + std::to_string(n) + " numbers in " + what);
}

}  // namespace

HydraScene load_hydra_scene(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open scene " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  Reader rd(text);
  std::unique_ptr<Node> doc = rd.document();
  HydraScene S;
  const size_t slash = path.find_last_of('/');
  const std::string dir = slash == std::string::npos ? "." : path.substr(0, slash);
  for (const auto& top : doc->kids) {
    if (top->name == "cam_lib" && !S.has_camera) {
      for (const auto& c : top->kids) {
        if (c->name != "camera") continue;
        HydraCamera& C = S.camera;
        if (const Node* n = c->child("fov")) floats(n->text, &C.fov, 1, "fov");
        if (const Node* n = c->child("nearClipPlane")) floats(n->text, &C.znear, 1, "nearClipPlane");
        if (const Node* n = c->child("farClipPlane")) floats(n->text, &C.zfar, 1, "farClipPlane");
        if (const Node* n = c->child("position")) floats(n->text, C.pos, 3, "position");
        if (const Node* n = c->child("look_at")) floats(n->text, C.look_at, 3, "look_at");
        if (const Node* n = c->child("up")) floats(n->text, C.up, 3, "up");
        S.has_camera = true;
        break;  // the first camera, as LoadSceneHydra takes it
      }
    } else if (top->name == "geometry_lib") {
      for (const auto& g : top->kids) {
        HydraGeom G;
        G.type = g->name;
        if (const std::string* loc = g->attr("loc")) G.path = dir + "/" + *loc;
        if (const std::string* id = g->attr("id")) G.id = std::atoi(id->c_str());
        S.geoms.push_back(G);
      }
    } else if (top->name == "scenes") {
      const Node* sc = top->child("scene");
      if (!sc || !S.instances.empty()) continue;
      for (const auto& in : sc->kids) {
        if (in->name != "instance") continue;
        HydraInstance I;
        if (const std::string* m = in->attr("mesh_id")) I.geom_id = std::atoi(m->c_str());
        if (const std::string* m = in->attr("matrix")) floats(*m, I.matrix, 16, "instance matrix");
        S.instances.push_back(I);
      }
    }
  }
  return S;
}

}  // namespace lrt
