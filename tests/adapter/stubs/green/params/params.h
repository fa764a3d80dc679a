// Test stand-in for green::params::params: string-keyed values with as<T>()
// and define<T>(name, description[, default]).
#pragma once
#include <map>
#include <sstream>
#include <string>
#include <type_traits>
namespace green::gpu { enum class LinearSolverType; }
namespace green::params {
  class params {
  public:
    struct value {
      std::string v;
      template <class T>
      T as() const {
        if constexpr (std::is_same_v<T, std::string>) return v;
        else if constexpr (std::is_enum_v<T>) return static_cast<T>(v == "LU" ? 1 : 0);
        else if constexpr (std::is_same_v<T, bool>) return v == "true" || v == "1";
        else { std::istringstream s(v); T x{}; s >> x; return x; }
      }
    };
    value operator[](const std::string& k) const {
      auto it = m_.find(k);
      return value{it == m_.end() ? std::string() : it->second};
    }
    void set(const std::string& k, const std::string& v) { m_[k] = v; }
    template <class T>
    void define(const std::string& k, const std::string& descr, T def) { (void)descr; (void)def; defined_[k] = true; }
    template <class T>
    void define(const std::string& k, const std::string& descr) { (void)descr; defined_[k] = true; }
    bool defined(const std::string& k) const { return defined_.count(k) != 0; }
  private:
    std::map<std::string, std::string> m_;
    std::map<std::string, bool>        defined_;
  };
}  // namespace green::params
