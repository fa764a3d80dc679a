// Test stand-in for green::utils::shared_object (single process).
#pragma once
namespace green::utils {
  template <class T>
  class shared_object {
  public:
    explicit shared_object(T obj) : obj_(std::move(obj)) {}
    T&   object() { return obj_; }
    void fence() {}
  private:
    T obj_;
  };
}  // namespace green::utils
