// ConfigBase (Utils/include/ConfigBase.h, Utils/src/ConfigBase.cpp): a
// driver's configuration class derives from it, implements parse() with
// read<T>(key[, default]) and optionally validate(), and calls load(file).
//
// File format (the reference's ConfigFile): one `key = value` per line;
// text after '#' is a comment; keys and values are trimmed of blanks; a
// later line overrides an earlier one.  Values convert with operator>>;
// strings are taken whole; booleans are false for "false", "f", "no",
// "n", "0" and "none" (any case) and true otherwise.  A missing file or a
// missing key without a default prints the reference's message and exits.
// One process reads the file (the single rank of this build).
#ifndef BCM_FACADE_CONFIGBASE_H
#define BCM_FACADE_CONFIGBASE_H

#include <cctype>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

#include "mpi.h"

// The reference's ConfigFile.h puts std::string in the global namespace,
// and drivers write read<string>(key).
using std::string;

class ConfigFile {
  std::map<std::string, std::string> items;

  static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
  }

  template <class T>
  static T convert(const std::string& s, T*) {
    T t = T();
    std::istringstream in(s);
    in >> t;
    return t;
  }
  static std::string convert(const std::string& s, std::string*) { return s; }
  static bool convert(const std::string& s, bool*) {
    std::string u;
    for (char c : s) u += char(std::toupper(static_cast<unsigned char>(c)));
    return !(u == "FALSE" || u == "F" || u == "NO" || u == "N" || u == "0" || u == "NONE");
  }

 public:
  struct file_not_found {
    std::string filename;
  };
  struct key_not_found {
    std::string key;
  };

  ConfigFile() {}
  explicit ConfigFile(const char* file) {
    std::ifstream in(file);
    if (!in) throw file_not_found{file};
    std::string line;
    while (std::getline(in, line)) {
      const size_t hash = line.find('#');
      if (hash != std::string::npos) line.erase(hash);
      const size_t eq = line.find('=');
      if (eq == std::string::npos) continue;
      const std::string key = trim(line.substr(0, eq));
      if (!key.empty()) items[key] = trim(line.substr(eq + 1));
    }
  }
  bool keyExists(const std::string& key) const { return items.count(key) != 0; }
  template <class T>
  T read(const std::string& key) const {
    auto it = items.find(key);
    if (it == items.end()) throw key_not_found{key};
    return convert(it->second, static_cast<T*>(nullptr));
  }
  template <class T>
  T read(const std::string& key, const T& value) const {
    auto it = items.find(key);
    if (it == items.end()) return value;
    return convert(it->second, static_cast<T*>(nullptr));
  }
};

class ConfigBase {
  MPI::Comm& comm;
  ConfigFile* configFile;

 public:
  ConfigBase(MPI::Comm& comm_ = MPI::COMM_WORLD) : comm(comm_), configFile(0) {}
  virtual ~ConfigBase() { delete configFile; }

  // ConfigBase::load (Utils/src/ConfigBase.cpp:34-60), one rank.
  void load(const char* file) {
    try {
      delete configFile;
      configFile = new ConfigFile(file);
      parse();
    } catch (ConfigFile::file_not_found& e) {
      std::cout << "error: cannot open configfile: " << e.filename << std::endl;
      errorExit("input config file.");
    } catch (ConfigFile::key_not_found& e) {
      std::cout << "error: cannot find key: " << e.key << std::endl;
      errorExit("input config file.");
    }
    if (!validate()) errorExit("input config file.");
  }

 private:
  virtual void parse() = 0;
  virtual bool validate() { return true; }

 protected:
  template <class T>
  T read(const std::string& key) const {
    return configFile->read<T>(key);
  }
  template <class T>
  T read(const std::string& key, const T& value) const {
    return configFile->read<T>(key, value);
  }
  void errorExit(const char* message, int code = 1) {
    std::cout << "error: " << message << std::endl;
    comm.Abort(code);
  }
};

#endif
