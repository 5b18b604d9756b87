// config.cpp -- see config.h.  Reading rules follow src/config.cpp:59-129.
#include "config.h"

#include <cctype>
#include <fstream>

namespace fsa_host {

namespace {
const char kDelimiter = '=';
const char kComment = '#';

std::string strip_comment(const std::string &line) {
    const auto pos = line.find(kComment);
    return pos == std::string::npos ? line : line.substr(0, pos);
}
}  // namespace

void Config::Trim(std::string &s) {
    static const char *ws = " \n\t\v\r\f";
    s.erase(0, s.find_first_not_of(ws));
    s.erase(s.find_last_not_of(ws) + 1U);
}

Config::Config(const std::string &filename) {
    std::ifstream in(filename);
    if (!in) throw File_not_found{filename};
    Load(in);
}

void Config::Load(std::istream &in) {
    // A key line may be followed by continuation lines: non-blank lines
    // without a delimiter are appended (joined by '\n') until a blank line,
    // a line holding a delimiter, or the end of the stream.  The line that
    // ends a value is parsed in the next round.
    std::string pending;
    bool have_pending = false;
    for (;;) {
        std::string line;
        if (have_pending) {
            line = pending;
            have_pending = false;
        } else if (!std::getline(in, line)) {
            break;
        }
        line = strip_comment(line);
        const auto eq = line.find(kDelimiter);
        if (eq == std::string::npos) continue;

        std::string key = line.substr(0, eq);
        std::string value = line.substr(eq + 1);
        std::string next;
        while (std::getline(in, next)) {
            std::string probe = next;
            Trim(probe);
            if (probe.empty()) {
                // a blank line ends the value; a whitespace-only line is
                // handed to the next round like any other line
                if (!next.empty()) {
                    pending = next;
                    have_pending = true;
                }
                break;
            }
            next = strip_comment(next);
            if (next.find(kDelimiter) != std::string::npos) {
                pending = next;
                have_pending = true;
                break;
            }
            probe = next;
            Trim(probe);
            if (!probe.empty()) value += "\n";
            value += next;
        }
        Trim(key);
        Trim(value);
        values_[key] = value;
    }
}

template <> bool Config::Convert<bool>(const std::string &s) {
    std::string up = s;
    for (auto &c : up) c = (char)std::toupper((unsigned char)c);
    return !(up == "FALSE" || up == "F" || up == "NO" || up == "N" || up == "0" || up == "NONE");
}

}  // namespace fsa_host
