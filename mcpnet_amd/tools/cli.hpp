// Minimal command-line parser for the mi / mcp / mcpnet tools.  The reference parses with CLI11
// (absent from its tree); option names follow its app_parameters::config() methods
// (apps/correlation/mi.cpp:54-58, apps/filter/mcp.cpp:93-110, apps/pipeline/mcpnet.cpp:105-129)
// and the README's common options -i, -o, -t, -m (README.md:116-124).
#pragma once

#include <cstdlib>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace cli {

class Parser {
 public:
  explicit Parser(std::string prog) : prog_(std::move(prog)) {}

  // names: "-i,--input"; takes_value: option expects a value (repeatable options collect all).
  void option(const std::string& names, const std::string& help, bool takes_value = true, std::string def = "") {
    Opt o{names, help, takes_value, def, {}, false};
    std::stringstream ss(names);
    std::string n;
    while (std::getline(ss, n, ',')) index_[n] = opts_.size();
    opts_.push_back(o);
  }

  void parse(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      if (a == "-h" || a == "--help") {
        usage(std::cout);
        std::exit(0);
      }
      std::string val;
      bool has_eq = false;
      auto eq = a.find('=');
      if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
        val = a.substr(eq + 1);
        a = a.substr(0, eq);
        has_eq = true;
      }
      auto it = index_.find(a);
      if (it == index_.end()) throw std::runtime_error("unknown option " + a);
      Opt& o = opts_[it->second];
      o.seen = true;
      if (!o.takes_value) continue;
      if (has_eq) {
        o.values.push_back(val);
        continue;
      }
      // collect one or more values up to the next option
      int n = 0;
      while (i + 1 < argc && !(argv[i + 1][0] == '-' && argv[i + 1][1] != '\0' && !isdigit(argv[i + 1][1]) &&
                               argv[i + 1][1] != '.')) {
        o.values.push_back(argv[++i]);
        ++n;
      }
      if (n == 0) throw std::runtime_error("option " + a + " needs a value");
    }
  }

  bool seen(const std::string& name) const { return get(name).seen; }
  std::string str(const std::string& name) const {
    const Opt& o = get(name);
    return o.values.empty() ? o.def : o.values.back();
  }
  long integer(const std::string& name) const { return std::stol(str(name)); }
  double real(const std::string& name) const { return std::stod(str(name)); }
  std::vector<std::string> all(const std::string& name) const { return get(name).values; }

  void usage(std::ostream& os) const {
    os << "usage: " << prog_ << " [options]\n";
    for (const auto& o : opts_) {
      os << "  " << o.names << (o.takes_value ? " VALUE" : "") << "\n      " << o.help;
      if (!o.def.empty()) os << " (default " << o.def << ")";
      os << "\n";
    }
  }

 private:
  struct Opt {
    std::string names, help;
    bool takes_value;
    std::string def;
    std::vector<std::string> values;
    bool seen;
  };
  const Opt& get(const std::string& name) const {
    auto it = index_.find(name);
    if (it == index_.end()) throw std::logic_error("undeclared option " + name);
    return opts_[it->second];
  }
  std::string prog_;
  std::vector<Opt> opts_;
  std::map<std::string, size_t> index_;
};

}  // namespace cli
