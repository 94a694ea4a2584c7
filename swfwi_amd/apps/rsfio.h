// Minimal Madagascar RSF I/O and key=value parameters for the command-line
// tools: the subset the reference's tools use (sf_input/sf_output,
// sf_histint/sf_histfloat, sf_getint/sf_getfloat/sf_getstring,
// sf_floatread/sf_floatwrite; src/rsf/file.c, getpar.c of swfwi).
//
// Header: text "key=value" tokens; the samples (native little-endian float)
// live in the file named by in= ("path@" by convention), or follow the
// header after the \014\014\004 separator when in="stdin".
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace rsf {

inline void die(const std::string& m) {
  std::fprintf(stderr, "%s\n", m.c_str());
  std::exit(1);
}

inline std::string unquote(std::string v) {
  if (v.size() >= 2 && (v.front() == '"' || v.front() == '\'') && v.back() == v.front()) v = v.substr(1, v.size() - 2);
  return v;
}

// key=value pairs of argv (later ones win), like sf_init + sf_get*
class Params {
 public:
  Params(int argc, char** argv) {
    for (int i = 1; i < argc; i++) {
      const char* eq = std::strchr(argv[i], '=');
      if (eq) kv_[std::string(argv[i], eq - argv[i])] = unquote(eq + 1);
    }
  }
  bool get(const std::string& k, std::string& v) const {
    auto it = kv_.find(k);
    if (it == kv_.end()) return false;
    v = it->second;
    return true;
  }
  bool getint(const std::string& k, int& v) const {
    std::string s;
    if (!get(k, s)) return false;
    v = std::atoi(s.c_str());
    return true;
  }
  bool getfloat(const std::string& k, float& v) const {
    std::string s;
    if (!get(k, s)) return false;
    v = (float)std::atof(s.c_str());
    return true;
  }
  std::string need(const std::string& k) const {
    std::string v;
    if (!get(k, v)) die("no " + k);
    return v;
  }

 private:
  std::map<std::string, std::string> kv_;
};

class Input {
 public:
  explicit Input(const std::string& path) : path_(path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) die("cannot open " + path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const std::string sep("\014\014\004");
    size_t cut = text.find(sep);
    std::string hdr = cut == std::string::npos ? text : text.substr(0, cut);
    std::istringstream is(hdr);
    std::string tok;
    while (is >> tok) {
      size_t eq = tok.find('=');
      if (eq != std::string::npos) kv_[tok.substr(0, eq)] = unquote(tok.substr(eq + 1));
    }
    std::string in = kv_.count("in") ? kv_["in"] : "stdin";
    if (in == "stdin") {
      if (cut == std::string::npos) die("no data in " + path);
      data_path_ = path;
      offset_ = (long)(cut + sep.size());
    } else {
      data_path_ = in;
      offset_ = 0;
    }
  }
  bool histint(const std::string& k, int& v) const {
    auto it = kv_.find(k);
    if (it == kv_.end()) return false;
    v = std::atoi(it->second.c_str());
    return true;
  }
  bool histfloat(const std::string& k, float& v) const {
    auto it = kv_.find(k);
    if (it == kv_.end()) return false;
    v = (float)std::atof(it->second.c_str());
    return true;
  }
  std::vector<float> floats(size_t n) const {
    std::vector<float> v(n);
    FILE* f = std::fopen(data_path_.c_str(), "rb");
    if (!f) die("cannot open data " + data_path_);
    std::fseek(f, offset_, SEEK_SET);
    if (std::fread(v.data(), sizeof(float), n, f) != n) die("short read in " + data_path_);
    std::fclose(f);
    return v;
  }

 private:
  std::string path_, data_path_;
  long offset_ = 0;
  std::map<std::string, std::string> kv_;
};

class Output {
 public:
  explicit Output(const std::string& path) : path_(path), data_path_(path + "@") {
    data_ = std::fopen(data_path_.c_str(), "wb");
    if (!data_) die("cannot create " + data_path_);
  }
  ~Output() { close(); }
  void put(const std::string& k, const std::string& v) { kv_.push_back(k + "=" + v); }
  void putint(const std::string& k, int v) { put(k, std::to_string(v)); }
  void putfloat(const std::string& k, float v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.9g", v);
    put(k, b);
  }
  void putstring(const std::string& k, const std::string& v) { put(k, "\"" + v + "\""); }
  FILE* data() { return data_; }
  void floatwrite(const float* p, size_t n) {
    if (std::fwrite(p, sizeof(float), n, data_) != n) die("short write to " + data_path_);
  }
  void close() {
    if (!data_) return;
    std::fclose(data_);
    data_ = nullptr;
    FILE* h = std::fopen(path_.c_str(), "w");
    if (!h) die("cannot create " + path_);
    for (const auto& s : kv_) std::fprintf(h, "\t%s\n", s.c_str());
    std::fprintf(h, "\tin=\"%s\"\n\tesize=4\n\tdata_format=\"native_float\"\n", data_path_.c_str());
    std::fclose(h);
  }

 private:
  std::string path_, data_path_;
  FILE* data_ = nullptr;
  std::vector<std::string> kv_;
};

}  // namespace rsf
