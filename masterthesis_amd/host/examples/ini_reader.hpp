// ini_reader.hpp -- the small ini-file reader of the example programs
// (in place of boost::property_tree::ini_parser, which the reference's
// examples use): "[section]" headers and "key = value" lines, ';' and '#'
// comments; values are looked up as "section.key".
#ifndef SCAFES_EXAMPLE_INI_READER_HPP_
#define SCAFES_EXAMPLE_INI_READER_HPP_

#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

class Ini {
public:
    explicit Ini(const std::string& path)
    {
        std::ifstream in(path.c_str());
        if (!in) throw std::runtime_error("cannot open config file \"" + path + "\" (use --configfile=<name>)");
        std::string line, sec;
        while (std::getline(in, line)) {
            const std::size_t c = line.find_first_of(";#");
            if (c != std::string::npos) line = line.substr(0, c);
            line = trim(line);
            if (line.empty()) continue;
            if (line[0] == '[') {
                sec = trim(line.substr(1, line.find(']') - 1));
                continue;
            }
            const std::size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            mKv[sec + "." + trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
        }
    }
    const std::string& get(const std::string& key) const
    {
        std::map<std::string, std::string>::const_iterator it = mKv.find(key);
        if (it == mKv.end()) throw std::runtime_error("config key " + key + " missing");
        return it->second;
    }
    double real(const std::string& key) const { return std::stod(get(key)); }
    // "AxBxC" or "a b c ..." lists
    std::vector<double> list(const std::string& key, char sep) const
    {
        std::vector<double> v;
        std::stringstream ss(get(key));
        std::string tok;
        while (std::getline(ss, tok, sep))
            if (!trim(tok).empty()) v.push_back(std::stod(tok));
        return v;
    }

private:
    static std::string trim(const std::string& s)
    {
        const std::size_t a = s.find_first_not_of(" \t\r\n");
        if (a == std::string::npos) return "";
        return s.substr(a, s.find_last_not_of(" \t\r\n") - a + 1);
    }
    std::map<std::string, std::string> mKv;
};

#endif
