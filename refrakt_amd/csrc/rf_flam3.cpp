// flam3 XML reader and the flame model helpers of the host library.
//
// Follows flame::load_flame (src/flame.cpp:160-226): the attributes read, the
// way numbers are converted (std::stof / std::stod / std::stoi, pugixml
// as_float / as_int), the palette scaling rgb/256 with alpha 1, and the
// failure on an attribute that is neither a known field, a variation nor a
// variation parameter.  pugixml is not available, so a small reader for the
// element/attribute subset flam3 files use is written here.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/refrakt_amd.h"
#include "rf_host.hpp"

namespace rfh {

namespace {

struct XmlElem {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<XmlElem> children;
};

struct XmlReader {
  const char* p;
  const char* end;
  std::string err;

  void skip_ws() { while (p < end && std::isspace((unsigned char)*p)) ++p; }
  bool starts(const char* s) const { size_t n = strlen(s); return (size_t)(end - p) >= n && memcmp(p, s, n) == 0; }
  bool skip_until(const char* s) {
    size_t n = strlen(s);
    while ((size_t)(end - p) >= n) {
      if (memcmp(p, s, n) == 0) { p += n; return true; }
      ++p;
    }
    return false;
  }
  // skip prolog, comments, doctype, processing instructions and text
  bool skip_misc() {
    for (;;) {
      while (p < end && *p != '<') ++p;
      if (p >= end) return false;
      if (starts("<?")) { if (!skip_until("?>")) return false; continue; }
      if (starts("<!--")) { if (!skip_until("-->")) return false; continue; }
      if (starts("<!")) { if (!skip_until(">")) return false; continue; }
      return true;
    }
  }
  static std::string decode(const std::string& s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
      if (s[i] == '&') {
        size_t j = s.find(';', i);
        if (j != std::string::npos) {
          std::string e = s.substr(i + 1, j - i - 1);
          if (e == "amp") { o += '&'; i = j; continue; }
          if (e == "lt") { o += '<'; i = j; continue; }
          if (e == "gt") { o += '>'; i = j; continue; }
          if (e == "quot") { o += '"'; i = j; continue; }
          if (e == "apos") { o += '\''; i = j; continue; }
        }
      }
      o += s[i];
    }
    return o;
  }
  bool parse_elem(XmlElem& e) {
    if (p >= end || *p != '<') { err = "expected '<'"; return false; }
    ++p;
    const char* s = p;
    while (p < end && !std::isspace((unsigned char)*p) && *p != '>' && *p != '/') ++p;
    e.name.assign(s, p);
    for (;;) {
      skip_ws();
      if (p >= end) { err = "unterminated element <" + e.name + ">"; return false; }
      if (*p == '/') {
        if (p + 1 < end && p[1] == '>') { p += 2; return true; }
        err = "bad '/' in <" + e.name + ">";
        return false;
      }
      if (*p == '>') { ++p; break; }
      const char* a = p;
      while (p < end && *p != '=' && !std::isspace((unsigned char)*p)) ++p;
      std::string an(a, p);
      skip_ws();
      if (p >= end || *p != '=') { err = "attribute '" + an + "' without value"; return false; }
      ++p;
      skip_ws();
      if (p >= end || (*p != '"' && *p != '\'')) { err = "attribute '" + an + "' not quoted"; return false; }
      char q = *p++;
      const char* v = p;
      while (p < end && *p != q) ++p;
      if (p >= end) { err = "unterminated attribute '" + an + "'"; return false; }
      e.attrs.emplace_back(an, decode(std::string(v, p)));
      ++p;
    }
    // content
    for (;;) {
      while (p < end && *p != '<') ++p;
      if (p >= end) { err = "missing </" + e.name + ">"; return false; }
      if (starts("</")) {
        p += 2;
        const char* s2 = p;
        while (p < end && *p != '>') ++p;
        std::string cn(s2, p);
        while (!cn.empty() && std::isspace((unsigned char)cn.back())) cn.pop_back();
        if (p < end) ++p;
        if (cn != e.name) { err = "mismatched </" + cn + "> for <" + e.name + ">"; return false; }
        return true;
      }
      if (starts("<!--")) { if (!skip_until("-->")) { err = "unterminated comment"; return false; } continue; }
      if (starts("<?")) { if (!skip_until("?>")) { err = "unterminated PI"; return false; } continue; }
      if (starts("<![CDATA[")) { if (!skip_until("]]>")) { err = "unterminated CDATA"; return false; } continue; }
      e.children.emplace_back();
      if (!parse_elem(e.children.back())) return false;
    }
  }
};

const XmlElem* find_flame(const XmlElem& e) {
  if (e.name == "flame") return &e;
  for (auto& c : e.children) {
    const XmlElem* f = find_flame(c);
    if (f) return f;
  }
  return nullptr;
}

const std::string* attr(const XmlElem& e, const char* n) {
  for (auto& a : e.attrs) if (a.first == n) return &a.second;
  return nullptr;
}

// pugixml as_float: strtod of the text, 0 when absent.
float as_float(const XmlElem& e, const char* n) {
  const std::string* v = attr(e, n);
  return v ? (float)strtod(v->c_str(), nullptr) : 0.0f;
}
// pugixml as_int: strtol base 10, 0 when absent.
int as_int(const XmlElem& e, const char* n) {
  const std::string* v = attr(e, n);
  return v ? (int)strtol(v->c_str(), nullptr, 10) : 0;
}

// parse_strings (src/util.hpp:12-25): whitespace-separated tokens, at most N,
// remaining entries zero; conv is applied per token.
template <class F>
int parse_list(const std::string& s, float* out, int n, F conv) {
  for (int i = 0; i < n; ++i) out[i] = 0.0f;
  const char* c = s.c_str();
  int k = 0;
  while (k < n) {
    while (*c && std::isspace((unsigned char)*c)) ++c;
    if (!*c) break;
    const char* b = c;
    while (*c && !std::isspace((unsigned char)*c)) ++c;
    out[k++] = conv(std::string(b, c));
  }
  return k;
}

float conv_stof(const std::string& t) { return strtof(t.c_str(), nullptr); }
float conv_stod(const std::string& t) { return (float)strtod(t.c_str(), nullptr); }
float conv_stoi_256(const std::string& t) { return (float)(int)strtol(t.c_str(), nullptr, 10) / 256.0f; }

bool parse_xform(const XmlElem& node, bool is_final, rf_xform& xf, std::string& err) {
  memset(&xf, 0, sizeof(xf));
  struct Pend { int var; int k; float v; };
  std::vector<Pend> params;
  for (auto& a : node.attrs) {
    const std::string& n = a.first;
    const char* v = a.second.c_str();
    if (n == "weight") xf.weight = (float)strtod(v, nullptr);
    else if (n == "color") xf.color = (float)strtod(v, nullptr);
    else if (n == "color_speed") xf.color_speed = (float)strtod(v, nullptr);
    else if (n == "animate") {
      // src/flame.cpp:190: the "final_xform" test never matches the element
      // name "finalxform", so a final xform's animate counts too.
      (void)is_final;
      xf.rotation_frequency = ((float)strtod(v, nullptr) > 0.0f) ? 1.0f : 0.0f;
    } else if (n == "opacity") xf.opacity = (float)strtod(v, nullptr);
    else if (n == "coefs") parse_list(a.second, xf.affine, 6, conv_stod);
    else if (n == "post") { parse_list(a.second, xf.post, 6, conv_stod); xf.has_post = 1; }
    else {
      int pv = -1, pk = -1;
      if (param_lookup(n.c_str(), &pv, &pk)) {
        params.push_back({pv, pk, (float)strtod(v, nullptr)});
        continue;
      }
      int id = rf_variation_id(n.c_str());
      if (id >= 0) {
        int slot = -1;
        for (int t = 0; t < xf.n_vars; ++t) if (xf.vars[t].id == id) slot = t;
        if (slot < 0) {
          if (xf.n_vars >= RF_MAX_VARS_PER_XFORM) { err = "too many variations in one xform"; return false; }
          slot = xf.n_vars++;
          memset(&xf.vars[slot], 0, sizeof(rf_var_term));
          xf.vars[slot].id = id;
        }
        xf.vars[slot].weight = (float)strtod(v, nullptr);
        continue;
      }
      err = "Unknown attribute " + n;
      return false;
    }
  }
  for (auto& pp : params)
    for (int t = 0; t < xf.n_vars; ++t)
      if (xf.vars[t].id == pp.var) xf.vars[t].params[pp.k] = pp.v;
  return true;
}

}  // namespace

int parse_flam3(const char* xml, size_t len, rf_flame_desc* out, std::string& err) {
  XmlReader rd{xml, xml + len, {}};
  XmlElem root;
  if (!rd.skip_misc()) { err = "no XML element found"; return RF_ERR_PARSE; }
  if (!rd.parse_elem(root)) { err = rd.err; return RF_ERR_PARSE; }
  const XmlElem* f = find_flame(root);
  if (!f) { err = "no <flame> element"; return RF_ERR_PARSE; }
  rf_flame_desc& d = *out;
  memset(&d, 0, sizeof(d));
  if (const std::string* c = attr(*f, "center")) parse_list(*c, d.center, 2, conv_stof);
  d.scale = as_float(*f, "scale");
  d.rotate = as_float(*f, "rotate");
  d.estimator_curve = as_float(*f, "estimator_curve");
  d.estimator_min = as_int(*f, "estimator_min");  // src/flame.cpp:171 (flam3 spells it estimator_minimum)
  d.estimator_radius = as_int(*f, "estimator_radius");
  d.brightness = as_float(*f, "brightness");
  d.gamma = as_float(*f, "gamma");
  d.vibrancy = as_float(*f, "vibrancy");
  if (const std::string* s = attr(*f, "size")) {
    float sz[2];
    parse_list(*s, sz, 2, [](const std::string& t) { return (float)(int)strtol(t.c_str(), nullptr, 10); });
    d.size[0] = (uint32_t)sz[0];
    d.size[1] = (uint32_t)sz[1];
  }
  for (auto& c : f->children) {
    if (c.name == "xform" || c.name == "finalxform") {
      rf_xform xf;
      if (!parse_xform(c, c.name == "finalxform", xf, err)) return RF_ERR_PARSE;
      if (c.name == "finalxform") { d.final_xform = xf; d.has_final = 1; }
      else {
        if (d.n_xforms >= RF_MAX_XFORMS) { err = "more than RF_MAX_XFORMS xforms"; return RF_ERR_LIMIT; }
        d.xforms[d.n_xforms++] = xf;
      }
    } else if (c.name == "color") {
      const std::string* iv = attr(c, "index");
      const std::string* rgb = attr(c, "rgb");
      unsigned long long idx = iv ? strtoull(iv->c_str(), nullptr, 10) : 0ull;
      if (idx >= RF_PALETTE_SIZE) { err = "palette index out of range"; return RF_ERR_PARSE; }
      float v[4] = {0, 0, 0, 0};
      if (rgb) parse_list(*rgb, v, 3, conv_stoi_256);
      d.palette[idx][0] = v[0];
      d.palette[idx][1] = v[1];
      d.palette[idx][2] = v[2];
      d.palette[idx][3] = 1.0f;
    }
  }
  if (d.n_xforms < 1) { err = "flame has no xforms"; return RF_ERR_PARSE; }
  return RF_OK;
}

}  // namespace rfh
