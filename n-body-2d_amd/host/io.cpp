// io.cpp — universe CSV reader/writer (InputOutput.cpp:18-72) and Bodies
// validation (Body.cpp:99-113).
#include <charconv>
#include <fstream>
#include <unordered_set>

#include "nbody2d_host.hpp"

namespace nbody2d {

bool Bodies::validate(std::string* why) const {
    const uint64_t count = n();
    if (id.size() != count || pos.size() != 2 * count || vel.size() != 2 * count) {
        if (why) *why = "inconsistent body array sizes";
        return false;
    }
    std::unordered_set<std::string> seen;
    seen.reserve(count);
    for (uint64_t i = 0; i < count; i++) {
        if (!seen.insert(id[i]).second) {
            if (why) *why = "Body #" + std::to_string(i) + " duplicate ID `" + id[i] + "`";
            return false;
        }
    }
    return true;
}

namespace io {
namespace {

// One CSV record: comma-separated fields, double quotes group a field and
// "" inside quotes is a literal quote.
std::vector<std::string> split_record(const std::string& line) {
    std::vector<std::string> out;
    std::string cur;
    bool quoted = false;
    for (size_t k = 0; k < line.size(); k++) {
        const char c = line[k];
        if (quoted) {
            if (c == '"') {
                if (k + 1 < line.size() && line[k + 1] == '"') {
                    cur += '"';
                    k++;
                } else {
                    quoted = false;
                }
            } else {
                cur += c;
            }
        } else if (c == '"') {
            quoted = true;
        } else if (c == ',') {
            out.push_back(cur);
            cur.clear();
        } else if (c != '\r') {
            cur += c;
        }
    }
    out.push_back(cur);
    return out;
}

// std::stod's rule: leading whitespace skipped, the longest valid prefix
// parsed ("2.97-e4" reads 2.97), no digits at all is an error.
double stod_prefix(const std::string& s, const std::string& where) {
    try {
        return std::stod(s);
    } catch (const std::exception&) {
        throw std::runtime_error("cannot parse number `" + s + "` in " + where);
    }
}

}  // namespace

Bodies parse_csv(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to parse: `" + path + "` (cannot open)");
    Bodies b;
    std::string line;
    bool header = true;
    uint64_t row = 0;
    try {
        while (std::getline(in, line)) {
            if (header) {  // the first row names the columns; they are positional
                header = false;
                continue;
            }
            row++;
            if (line.empty() || line == "\r") continue;
            const std::vector<std::string> f = split_record(line);
            if (f.size() < 6)
                throw std::runtime_error("row " + std::to_string(row) + " has " +
                                         std::to_string(f.size()) + " fields, expected 6");
            const std::string where = "row " + std::to_string(row);
            b.id.push_back(f[0]);
            b.mass.push_back(stod_prefix(f[1], where));
            b.pos.push_back(stod_prefix(f[2], where));
            b.pos.push_back(stod_prefix(f[3], where));
            b.vel.push_back(stod_prefix(f[4], where));
            b.vel.push_back(stod_prefix(f[5], where));
        }
    } catch (const std::exception& e) {
        throw std::runtime_error("Failed to parse: `" + path + "`: " + e.what());
    }
    std::string why;
    if (!b.validate(&why)) throw std::runtime_error("Failed to validate: `" + path + "`: " + why);
    return b;
}

std::string format_double(double v) {
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof(buf), v);  // shortest round trip
    return std::string(buf, r.ptr);
}

void write_csv(const std::string& path, const Bodies& b) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("Failed to write: `" + path + "`");
    out << "id,mass,x,y,vel_x,vel_y\n";
    for (uint64_t i = 0; i < b.n(); i++) {
        out << b.id[i] << ',' << format_double(b.mass[i]) << ',' << format_double(b.pos[2 * i])
            << ',' << format_double(b.pos[2 * i + 1]) << ',' << format_double(b.vel[2 * i])
            << ',' << format_double(b.vel[2 * i + 1]) << '\n';
    }
    if (!out) throw std::runtime_error("Failed to write: `" + path + "`");
}

}  // namespace io
}  // namespace nbody2d
