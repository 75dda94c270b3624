// Minimal stand-in for the parts of srrg2_core the aligner plugins use:
// Configurable, PARAM properties with a "changed" flag, and
// BOSS_REGISTER_CLASS (create-by-type-name).  srrg2_core is not vendored by
// the reference (docker/repos.yml:6-13) and is not available here; this
// header keeps the adapter's surface identical so the same class bodies
// compile against the real library with only the include swapped.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace srrg2_core {

struct Configurable {
  virtual ~Configurable() = default;
};

template <typename T>
struct Property_ {
  Property_(const std::string& name, const std::string& desc, T def, bool* changed)
      : name_(name), desc_(desc), value_(def), changed_(changed) {}
  const T& value() const { return value_; }
  void setValue(const T& v) {
    value_ = v;
    if (changed_) *changed_ = true;
  }
  const std::string& name() const { return name_; }
  const std::string& description() const { return desc_; }

 private:
  std::string name_, desc_;
  T value_;
  bool* changed_;
};

using PropertyFloat = Property_<float>;
using PropertyInt = Property_<int>;
using PropertyBool = Property_<bool>;
using PropertyString = Property_<std::string>;

template <typename C>
struct PropertyConfigurable_ : Property_<std::shared_ptr<C>> {
  using Property_<std::shared_ptr<C>>::Property_;
  C* operator->() const { return this->value().get(); }
};

// Type registry: BOSS_REGISTER_CLASS(T) makes T creatable by its name.
class TypeRegistry {
 public:
  using Factory = std::function<std::shared_ptr<Configurable>()>;
  static TypeRegistry& instance() {
    static TypeRegistry r;
    return r;
  }
  void add(const std::string& name, Factory f) { factories_[name] = std::move(f); }
  std::shared_ptr<Configurable> create(const std::string& name) const {
    auto it = factories_.find(name);
    if (it == factories_.end()) throw std::runtime_error("unknown type " + name);
    return it->second();
  }
  bool has(const std::string& name) const { return factories_.count(name) != 0; }
  std::vector<std::string> names() const {
    std::vector<std::string> v;
    for (auto& kv : factories_) v.push_back(kv.first);
    return v;
  }

 private:
  std::map<std::string, Factory> factories_;
};

}  // namespace srrg2_core

// PARAM(type, name, description, default, changed_flag_ptr) declares a member
// `param_<name>` exactly as srrg2_core does.
#define PARAM(type_, name_, desc_, default_, flag_) type_ param_##name_{#name_, desc_, default_, flag_}

#define BOSS_REGISTER_CLASS(T_) \
  srrg2_core::TypeRegistry::instance().add(#T_, [] { return std::shared_ptr<srrg2_core::Configurable>(new T_()); })
