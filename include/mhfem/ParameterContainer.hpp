// ParameterContainer.hpp -- the option set the solver reads (configSetup.h:9-63):
// a string map with typed getters, filled from code or from "key = value" INI
// lines (config.ini of the reference's HeatEquation example).
#pragma once

#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace mhfem {

class ParameterContainer {
public:
    void setParameter(const std::string &key, const std::string &value) { values_[key] = value; }
    template <typename T>
    void setParameter(const std::string &key, const T &value)
    {
        std::ostringstream os;
        os.precision(17);
        os << value;
        values_[key] = os.str();
    }
    bool checkParameter(const std::string &key) const { return values_.count(key) != 0; }

    template <typename T>
    T getParameter(const std::string &key) const
    {
        auto it = values_.find(key);
        if (it == values_.end()) throw std::runtime_error("missing parameter: " + key);
        std::istringstream is(it->second);
        T v{};
        is >> v;
        if (is.fail()) throw std::runtime_error("bad value for parameter " + key + ": " + it->second);
        return v;
    }
    template <typename T>
    T getParameter(const std::string &key, const T &fallback) const
    {
        return checkParameter(key) ? getParameter<T>(key) : fallback;
    }

    // "key = value" lines, '#' comments (the INI format of config.ini)
    void parseINI(const std::string &path)
    {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("cannot open " + path);
        std::string line;
        while (std::getline(in, line)) {
            const auto hash = line.find('#');
            if (hash != std::string::npos) line.resize(hash);
            const auto eq = line.find('=');
            if (eq == std::string::npos) continue;
            auto trim = [](std::string s) {
                const auto a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
                return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
            };
            values_[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
        }
    }

private:
    std::map<std::string, std::string> values_;
};

template <>
inline std::string ParameterContainer::getParameter<std::string>(const std::string &key) const
{
    auto it = values_.find(key);
    if (it == values_.end()) throw std::runtime_error("missing parameter: " + key);
    return it->second;
}

}  // namespace mhfem
