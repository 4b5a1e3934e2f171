#pragma once

// Keyed store through which passes hand data to each other (the role of the
// reference's AnyBag, base/anybag.hpp:22-176, with the same method names:
// set / get / getPtr / take / has / erase).

#include <any>
#include <map>
#include <optional>
#include <string>
#include <type_traits>
#include <utility>

namespace wslam::compute {

class Store {
   public:
    template <typename T>
    void set(const std::string& key, T&& value) {
        data_[key] = std::make_any<std::decay_t<T>>(std::forward<T>(value));
    }

    template <typename T>
    [[nodiscard]] std::optional<T> get(const std::string& key) const {
        const T* p = raw<T>(key);
        if (!p) return std::nullopt;
        return *p;
    }

    template <typename T>
    [[nodiscard]] std::optional<const T*> getPtr(const std::string& key) const {
        const T* p = raw<T>(key);
        if (!p) return std::nullopt;
        return p;
    }

    template <typename T>
    [[nodiscard]] std::optional<T*> getPtr(const std::string& key) {
        T* p = const_cast<T*>(raw<T>(key));
        if (!p) return std::nullopt;
        return p;
    }

    // Moves the value out; erases the key unless told otherwise.
    template <typename T>
    [[nodiscard]] std::optional<T> take(const std::string& key, bool erase_after_move = true) {
        T* p = const_cast<T*>(raw<T>(key));
        if (!p) return std::nullopt;
        std::optional<T> out(std::move(*p));
        if (erase_after_move) data_.erase(key);
        return out;
    }

    [[nodiscard]] bool has(const std::string& key) const { return data_.count(key) != 0; }
    bool erase(const std::string& key) { return data_.erase(key) > 0; }

   private:
    template <typename T>
    const T* raw(const std::string& key) const {
        const auto it = data_.find(key);
        if (it == data_.end()) return nullptr;
        return std::any_cast<T>(&it->second);
    }

    std::map<std::string, std::any> data_;
};

}  // namespace wslam::compute
