// props.h -- typed property bag, the subset of sys/props.h + sys/value.h
// (Props::get/set/exists, Value::get<T>) that reaches the Device API.
#pragma once

#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace prt {

struct Vec2i { int x = 0, y = 0; };
struct Vec2f { float x = 0, y = 0; };
struct Vec3f { float x = 0, y = 0, z = 0; };
struct Box3f { Vec3f low, high; };
struct Basis3f { Vec3f U, V, N; };

// A value holds either a string or up to 9 numbers; get<T>() converts like
// Value::get (sys/value.h:140-350): numbers to numbers, strings parsed with
// ',' between components (math/vec2.h:446-458).
class Value {
public:
    Value() = default;
    Value(const std::string& s) : str_(s), is_str_(true) {}
    Value(const char* s) : str_(s), is_str_(true) {}
    Value(int v) : nums_{(double)v} {}
    Value(int64_t v) : nums_{(double)v} {}
    Value(float v) : nums_{(double)v} {}
    Value(double v) : nums_{v} {}
    Value(bool v) : nums_{v ? 1.0 : 0.0} {}
    Value(const Vec2i& v) : nums_{(double)v.x, (double)v.y} {}
    Value(const Vec2f& v) : nums_{v.x, v.y} {}
    Value(const Vec3f& v) : nums_{v.x, v.y, v.z} {}
    Value(const Basis3f& b) : nums_{b.U.x, b.U.y, b.U.z, b.V.x, b.V.y, b.V.z, b.N.x, b.N.y, b.N.z} {}

    bool empty() const { return is_str_ ? str_.empty() : nums_.empty(); }

    template <class T> T get() const;

    std::string text() const
    {
        if (is_str_) return str_;
        std::ostringstream sm;
        sm.precision(9);
        for (size_t i = 0; i < nums_.size(); ++i) sm << (i ? "," : "") << nums_[i];
        return sm.str();
    }

private:
    std::vector<double> numbers() const
    {
        if (!is_str_) return nums_;
        std::vector<double> out;
        std::string s = str_;
        for (char& c : s)
            if (c == ',' || c == 'x') c = ' ';
        std::istringstream sm(s);
        double d;
        while (sm >> d) out.push_back(d);
        if (out.empty()) throw std::invalid_argument("not a number: '" + str_ + "'");
        return out;
    }

    std::string str_;
    std::vector<double> nums_;
    bool is_str_ = false;
};

template <> inline std::string Value::get<std::string>() const { return text(); }
template <> inline double Value::get<double>() const { return numbers().at(0); }
template <> inline float Value::get<float>() const { return (float)numbers().at(0); }
template <> inline int Value::get<int>() const { return (int)numbers().at(0); }
template <> inline int64_t Value::get<int64_t>() const { return (int64_t)numbers().at(0); }
template <> inline Vec2i Value::get<Vec2i>() const
{
    auto n = numbers();
    return Vec2i{(int)n[0], (int)(n.size() > 1 ? n[1] : n[0])};
}
template <> inline Vec3f Value::get<Vec3f>() const
{
    auto n = numbers();
    if (n.size() == 1) return Vec3f{(float)n[0], (float)n[0], (float)n[0]};
    return Vec3f{(float)n.at(0), (float)n.at(1), (float)n.at(2)};
}
template <> inline Basis3f Value::get<Basis3f>() const
{
    auto n = numbers();
    if (n.size() != 9) throw std::invalid_argument("basis needs 9 numbers");
    return Basis3f{{(float)n[0], (float)n[1], (float)n[2]}, {(float)n[3], (float)n[4], (float)n[5]},
                   {(float)n[6], (float)n[7], (float)n[8]}};
}

// Insertion-ordered property bag.
class Props {
public:
    void set(const std::string& name, const Value& v)
    {
        for (auto& kv : items_)
            if (kv.first == name) { kv.second = v; return; }
        items_.emplace_back(name, v);
    }
    bool exists(const std::string& name) const { return find(name) != nullptr; }
    template <class T> T get(const std::string& name) const
    {
        const Value* v = find(name);
        if (!v) throw std::invalid_argument("missing property: " + name);
        return v->get<T>();
    }
    template <class T> T get(const std::string& name, const T& def) const
    {
        const Value* v = find(name);
        return (v && !v->empty()) ? v->get<T>() : def;
    }
    std::string get(const std::string& name, const char* def) const { return get<std::string>(name, std::string(def)); }
    const std::vector<std::pair<std::string, Value>>& items() const { return items_; }
    void clear() { items_.clear(); }
    bool empty() const { return items_.empty(); }

private:
    const Value* find(const std::string& name) const
    {
        for (const auto& kv : items_)
            if (kv.first == name) return &kv.second;
        return nullptr;
    }
    std::vector<std::pair<std::string, Value>> items_;
};

}  // namespace prt
