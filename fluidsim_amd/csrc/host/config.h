// config.h -- flat `key = value` reader for config/input.ini.
//
// Same reading rules as the reference Config class (include/config.h:14-200,
// src/config.cpp:3-150): '#' starts a comment, lines without '=' are ignored
// unless they directly continue the value of the previous key, `[sections]`
// carry no meaning, a repeated key overwrites, keys and values are trimmed,
// and typed reads go through std::istringstream (bools accept
// false/f/no/n/0/none as false, anything else as true).
#pragma once

#include <istream>
#include <map>
#include <sstream>
#include <string>

namespace fsa_host {

class Config {
public:
    struct File_not_found {
        std::string filename;
    };
    struct Key_not_found {
        std::string key;
    };

    Config() = default;
    explicit Config(const std::string &filename);

    // parse an input stream; keys already present are overwritten
    void Load(std::istream &in);

    bool KeyExists(const std::string &key) const { return values_.count(key) != 0; }
    const std::map<std::string, std::string> &Contents() const { return values_; }

    template <class T> T Read(const std::string &key) const {
        auto it = values_.find(key);
        if (it == values_.end()) throw Key_not_found{key};
        return Convert<T>(it->second);
    }
    template <class T> T Read(const std::string &key, const T &fallback) const {
        auto it = values_.find(key);
        return it == values_.end() ? fallback : Convert<T>(it->second);
    }

    static void Trim(std::string &s);

private:
    template <class T> static T Convert(const std::string &s) {
        T value{};
        std::istringstream in(s);
        in >> value;
        return value;
    }
    std::map<std::string, std::string> values_;
};

template <> inline std::string Config::Convert<std::string>(const std::string &s) { return s; }
template <> bool Config::Convert<bool>(const std::string &s);

}  // namespace fsa_host
