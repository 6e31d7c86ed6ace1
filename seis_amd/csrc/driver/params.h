// Parameter files of the inversion driver: the key/value grammar of the
// reference's ParserFromFile (util/parser.cpp:16-168) and the typed parameter
// set paramLSRTM (lsrtm/datastr.h:51-75) with its defaults and optional flags.
#pragma once
#include <string>
#include <vector>

namespace seis {

// One "key value" lookup in a parameter file.  A line matches when it
// contains `key` followed by a space, after removing everything from '!' on;
// the value is the first token of the rest of the line that starts with a
// character of the value class (parser.cpp:29-45):
//   int    : one of "+-0123456789"      (then parsed like std::stoi)
//   float  : one of "+-.0123456789"     (then parsed like std::stof)
//   string : '/' or an ASCII letter
// The first matching line that carries a value wins.
class ParamFile {
 public:
  explicit ParamFile(const std::string& path);
  bool ok() const { return ok_; }
  // Each returns true when a value was found; false leaves `val` untouched.
  bool get_int(const std::string& key, int& val) const;
  bool get_float(const std::string& key, float& val) const;
  bool get_string(const std::string& key, std::string& val) const;
  // Whether any line mentions the key at all (for "value missing" errors).
  bool has_key(const std::string& key) const;

 private:
  bool find(const std::string& key, const char* cls, std::string& tok) const;
  std::vector<std::string> lines_;
  bool ok_ = false;
};

// The driver's parameters (datastr.h:51-75): value, default and whether the
// key may be omitted.
struct IntParam {
  const char* key;
  int val, def;
  bool opt;
};
struct FloatParam {
  const char* key;
  float val, def;
  bool opt;
};
struct StringParam {
  const char* key;
  std::string val, def;
  bool opt;
};

struct Params {
  IntParam sbegin{"sbegin", 0, 0, false}, send{"send", 0, 0, false};
  IntParam planeTag{"planeTag", 0, 0, true}, nthread{"nthread", 0, 1, true};
  IntParam ns{"ns", 0, 0, false}, ngmax{"ngmax", 0, 0, false};
  IntParam nx{"nx", 0, 0, false}, nz{"nz", 0, 0, false}, nt{"nt", 0, 0, false};
  IntParam npml{"npml", 0, 10, true}, mask{"mask", 0, 0, true}, maxiter{"maxiter", 0, 30, true};
  IntParam delay{"delay", 0, 0, true}, delaycal{"delaycal", 0, 0, true};
  IntParam lpad{"lpad", 0, 0, true}, rpad{"rpad", 0, 0, true};
  FloatParam dx{"dx", 0, 0, false}, dz{"dz", 0, 0, false}, dt{"dt", 0, 0, false}, fr{"fr", 0, 0, false};
  FloatParam velfx{"velfx", 0, 0, true};
  FloatParam pmin{"pmin", 0, 0, false}, dp{"dp", 0, 0, false};
  IntParam np{"np", 0, 0, false};
  StringParam precsg{"precsg", "", "", false}, wdir{"wdir", "", "", false};
  StringParam v0file{"v0file", "", "", false}, vfile{"vfile", "", "", true};
  StringParam maskfile{"maskfile", "", "", true}, coordfile{"coordfile", "", "", false};
};

// Reads every key in the order of Inversion::readParamFile
// (inversion.cpp:38-79): maskfile becomes mandatory when mask != 0 and
// np / dp / pmin are read only when planeTag != 0.  Returns false with a
// message ("key <k> not found", "key <k> value missing", "File can not
// open") instead of exiting as the reference does.
bool read_params(const std::string& path, Params& p, std::string& err);

// Human-readable dump of the parameters (datastr.cpp:72-104).
std::string format_params(const Params& p);

}  // namespace seis
