// ini_reader.cpp -- INI parser with the syntax of the reader the reference
// vendors (CUDA_week2/INIReader.cuh:121-454, inih r-series):
//   * optional UTF-8 BOM on the first line
//   * lines are stripped of surrounding whitespace; a line starting with ';'
//     or '#' is a comment
//   * "[name]" opens a section (text after ']' is ignored)
//   * "name = value" or "name : value" (the first '=' or ':' splits)
//   * ';' starts an inline comment only when preceded by whitespace
//   * an indented non-blank line continues the previous key: its text is
//     appended to the value after a '\n'
//   * a key given twice keeps both values joined by '\n'
//   * section and key names are case-insensitive
//   * a line that is none of the above records its line number as the parse
//     error (the first such line), and parsing continues
// GetInteger / GetReal parse the longest numeric prefix (strtol base 0 /
// strtod) and fall back to the default when there is none.
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "fd_host.hpp"

namespace {

bool is_space(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }

std::string strip(const std::string& s)
{
    size_t b = 0, e = s.size();
    while (b < e && is_space(s[b])) b++;
    while (e > b && is_space(s[e - 1])) e--;
    return s.substr(b, e - b);
}

// position of the first character of `stop` or of an inline ';' comment
// (a ';' right after whitespace), or s.size()
size_t scan(const std::string& s, size_t from, const char* stop)
{
    bool after_space = false;
    for (size_t i = from; i < s.size(); i++) {
        const char c = s[i];
        if (stop && std::strchr(stop, c)) return i;
        if (after_space && c == ';') return i;
        after_space = is_space(c);
    }
    return s.size();
}

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

std::string key_of(const std::string& section, const std::string& name) { return lower(section + "=" + name); }

}  // namespace

INIReader::INIReader(const std::string& filename)
{
    std::ifstream in(filename, std::ios::binary);
    if (!in) {
        error_ = -1;
        return;
    }
    std::string raw, section, prev_name;
    int lineno = 0;
    auto add = [&](const std::string& name, const std::string& value) {
        std::string& v = values_[key_of(section, name)];
        if (!v.empty()) v += "\n";
        v += value;
        sections_.insert(section);
    };
    while (std::getline(in, raw)) {
        lineno++;
        if (lineno == 1 && raw.size() >= 3 && (unsigned char)raw[0] == 0xEF && (unsigned char)raw[1] == 0xBB &&
            (unsigned char)raw[2] == 0xBF)
            raw.erase(0, 3);
        const bool indented = !raw.empty() && is_space(raw[0]);
        const std::string line = strip(raw);
        if (line.empty()) continue;
        if (line[0] == ';' || line[0] == '#') continue;
        if (!prev_name.empty() && indented) {
            // continuation of the previous key's value
            add(prev_name, strip(line.substr(0, scan(line, 0, nullptr))));
            continue;
        }
        if (line[0] == '[') {
            const size_t e = scan(line, 1, "]");
            if (e < line.size() && line[e] == ']') {
                section = line.substr(1, e - 1);
                prev_name.clear();
            } else if (!error_) {
                error_ = lineno;
            }
            continue;
        }
        const size_t e = scan(line, 0, "=:");
        if (e < line.size() && (line[e] == '=' || line[e] == ':')) {
            const std::string name = strip(line.substr(0, e));
            std::string value = line.substr(e + 1);
            size_t b = 0;
            while (b < value.size() && is_space(value[b])) b++;
            value = value.substr(b);
            value = strip(value.substr(0, scan(value, 0, nullptr)));
            prev_name = name;
            add(name, value);
        } else if (!error_) {
            error_ = lineno;
        }
    }
}

std::string INIReader::Get(const std::string& section, const std::string& name,
                           const std::string& default_value) const
{
    const auto it = values_.find(key_of(section, name));
    return it == values_.end() ? default_value : it->second;
}

bool INIReader::Has(const std::string& section, const std::string& name) const
{
    return values_.count(key_of(section, name)) != 0;
}

long INIReader::GetInteger(const std::string& section, const std::string& name, long default_value) const
{
    const std::string v = Get(section, name, "");
    char* end = nullptr;
    const long n = std::strtol(v.c_str(), &end, 0);
    return end > v.c_str() ? n : default_value;
}

double INIReader::GetReal(const std::string& section, const std::string& name, double default_value) const
{
    const std::string v = Get(section, name, "");
    char* end = nullptr;
    const double n = std::strtod(v.c_str(), &end);
    return end > v.c_str() ? n : default_value;
}

bool INIReader::GetBoolean(const std::string& section, const std::string& name, bool default_value) const
{
    const std::string v = lower(Get(section, name, ""));
    if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
    if (v == "false" || v == "no" || v == "off" || v == "0") return false;
    return default_value;
}
