#include "params.h"

#include <cctype>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace seis {

ParamFile::ParamFile(const std::string& path) {
  std::ifstream in(path);
  if (!in) return;
  ok_ = true;
  std::string line;
  while (std::getline(in, line)) {
    const size_t c = line.find('!');
    if (c != std::string::npos) line.erase(c);
    lines_.push_back(line);
  }
}

bool ParamFile::has_key(const std::string& key) const {
  const std::string k = key + " ";
  for (const std::string& l : lines_)
    if (l.find(k) != std::string::npos) return true;
  return false;
}

bool ParamFile::find(const std::string& key, const char* cls, std::string& tok) const {
  const std::string k = key + " ";
  for (const std::string& l : lines_) {
    const size_t at = l.find(k);
    if (at == std::string::npos) continue;
    const std::string rest = l.substr(at + k.size());
    const size_t b = rest.find_first_of(cls);
    if (b == std::string::npos) continue;
    size_t e = b;
    while (e < rest.size() && !std::isspace((unsigned char)rest[e])) e++;
    tok = rest.substr(b, e - b);
    return true;
  }
  return false;
}

bool ParamFile::get_int(const std::string& key, int& val) const {
  std::string t;
  if (!find(key, "+-0123456789", t)) return false;
  try {
    val = std::stoi(t);
  } catch (...) {
    return false;
  }
  return true;
}

bool ParamFile::get_float(const std::string& key, float& val) const {
  std::string t;
  if (!find(key, "+-.0123456789", t)) return false;
  try {
    val = std::stof(t);
  } catch (...) {
    return false;
  }
  return true;
}

bool ParamFile::get_string(const std::string& key, std::string& val) const {
  std::string t;
  if (!find(key, "/qwertyuiopasdfghjklzxcvbnmQWERTYUIOPASDFGHJKLZXCVBNM", t)) return false;
  val = t;
  return true;
}

namespace {

template <class P, class G>
bool fetch(const ParamFile& f, P& p, G get, std::string& err) {
  if (get(p.key, p.val)) return true;
  if (p.opt) {
    p.val = p.def;
    return true;
  }
  err = std::string("key ") + p.key + (f.has_key(p.key) ? " value missing" : " not found");
  return false;
}

}  // namespace

bool read_params(const std::string& path, Params& p, std::string& err) {
  ParamFile f(path);
  if (!f.ok()) {
    err = "File can not open : " + path;
    return false;
  }
  auto gi = [&](const std::string& k, int& v) { return f.get_int(k, v); };
  auto gf = [&](const std::string& k, float& v) { return f.get_float(k, v); };
  auto gs = [&](const std::string& k, std::string& v) { return f.get_string(k, v); };
  for (IntParam* q : {&p.nthread, &p.ns, &p.ngmax, &p.nx, &p.nz, &p.nt, &p.npml, &p.delay, &p.delaycal, &p.mask,
                      &p.maxiter, &p.lpad, &p.rpad, &p.planeTag, &p.sbegin, &p.send})
    if (!fetch(f, *q, gi, err)) return false;
  for (FloatParam* q : {&p.dx, &p.dz, &p.dt, &p.fr, &p.velfx})
    if (!fetch(f, *q, gf, err)) return false;
  for (StringParam* q : {&p.precsg, &p.wdir, &p.v0file, &p.vfile, &p.coordfile})
    if (!fetch(f, *q, gs, err)) return false;
  if (p.mask.val) {
    p.maskfile.opt = false;
    if (!fetch(f, p.maskfile, gs, err)) return false;
  }
  if (p.planeTag.val) {
    if (!fetch(f, p.np, gi, err) || !fetch(f, p.dp, gf, err) || !fetch(f, p.pmin, gf, err)) return false;
  }
  return true;
}

std::string format_params(const Params& p) {
  std::ostringstream o;
  for (const IntParam* q : {&p.nthread, &p.ns, &p.ngmax, &p.nx, &p.nz, &p.nt, &p.npml, &p.delay, &p.delaycal,
                            &p.mask, &p.maxiter, &p.lpad, &p.rpad, &p.planeTag, &p.sbegin, &p.send, &p.np})
    o << q->key << " = " << q->val << "\n";
  char b[64];
  for (const FloatParam* q : {&p.dx, &p.dz, &p.dt, &p.fr, &p.velfx, &p.dp, &p.pmin}) {
    std::snprintf(b, sizeof(b), "%.9g", (double)q->val);  // exact for float
    o << q->key << " = " << b << "\n";
  }
  for (const StringParam* q : {&p.precsg, &p.wdir, &p.v0file, &p.vfile, &p.coordfile, &p.maskfile})
    o << q->key << " = " << q->val << "\n";
  return o.str();
}

}  // namespace seis
