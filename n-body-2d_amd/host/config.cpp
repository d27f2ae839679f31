// config.cpp — config.json reader (Config.cpp:17-64) with a small JSON parser
// (the reference uses nlohmann_json, absent here), and the range checks of
// Config::Simulation::validate (Config.cpp:224-253).
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>

#include "nbody2d_host.hpp"

namespace nbody2d {
namespace {

struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    const Json& at(const std::string& k) const {
        if (kind != OBJ) throw std::runtime_error("expected an object around key `" + k + "`");
        auto it = obj.find(k);
        if (it == obj.end()) throw std::runtime_error("missing key `" + k + "`");
        return it->second;
    }
    bool has(const std::string& k) const { return kind == OBJ && obj.count(k); }
    double number() const {
        if (kind != NUM) throw std::runtime_error("expected a number");
        return num;
    }
    bool boolean() const {
        if (kind != BOOL) throw std::runtime_error("expected a boolean");
        return b;
    }
    const std::string& string() const {
        if (kind != STR) throw std::runtime_error("expected a string");
        return str;
    }
};

class Parser {
public:
    explicit Parser(const std::string& s) : s_(s) {}
    Json parse() {
        Json v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const std::string& m) const {
        throw std::runtime_error("JSON parse error at offset " + std::to_string(i_) + ": " + m);
    }
    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) i_++;
    }
    bool eat(char c) {
        ws();
        if (i_ < s_.size() && s_[i_] == c) {
            i_++;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail(std::string("expected `") + c + "`");
    }
    Json value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        const char c = s_[i_];
        Json v;
        if (c == '{') {
            i_++;
            v.kind = Json::OBJ;
            if (eat('}')) return v;
            do {
                ws();
                Json k = value();
                if (k.kind != Json::STR) fail("object key must be a string");
                expect(':');
                v.obj[k.str] = value();
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            i_++;
            v.kind = Json::ARR;
            if (eat(']')) return v;
            do v.arr.push_back(value());
            while (eat(','));
            expect(']');
        } else if (c == '"') {
            i_++;
            v.kind = Json::STR;
            while (i_ < s_.size() && s_[i_] != '"') {
                char ch = s_[i_++];
                if (ch == '\\') {
                    if (i_ >= s_.size()) fail("bad escape");
                    const char e = s_[i_++];
                    switch (e) {
                    case 'n': ch = '\n'; break;
                    case 't': ch = '\t'; break;
                    case 'r': ch = '\r'; break;
                    case 'b': ch = '\b'; break;
                    case 'f': ch = '\f'; break;
                    case 'u': {
                        if (i_ + 4 > s_.size()) fail("bad \\u escape");
                        const unsigned cp = std::stoul(s_.substr(i_, 4), nullptr, 16);
                        i_ += 4;
                        if (cp < 0x80) {
                            ch = (char)cp;
                        } else if (cp < 0x800) {
                            v.str += (char)(0xC0 | (cp >> 6));
                            ch = (char)(0x80 | (cp & 0x3F));
                        } else {
                            v.str += (char)(0xE0 | (cp >> 12));
                            v.str += (char)(0x80 | ((cp >> 6) & 0x3F));
                            ch = (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: ch = e;
                    }
                }
                v.str += ch;
            }
            if (i_ >= s_.size()) fail("unterminated string");
            i_++;
        } else if (s_.compare(i_, 4, "true") == 0) {
            i_ += 4;
            v.kind = Json::BOOL;
            v.b = true;
        } else if (s_.compare(i_, 5, "false") == 0) {
            i_ += 5;
            v.kind = Json::BOOL;
        } else if (s_.compare(i_, 4, "null") == 0) {
            i_ += 4;
        } else {
            const char* b = s_.c_str() + i_;
            char* e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("unexpected character");
            i_ += (size_t)(e - b);
            v.kind = Json::NUM;
        }
        return v;
    }
};

std::string lower(const std::string& s) {
    std::string o;
    for (char c : s) o += (char)std::tolower((unsigned char)c);
    return o;
}

template <class T>
bool in_range(T v, T lo, T hi) {
    return v >= lo && v <= hi;
}

}  // namespace

bool SimConfig::validate(std::vector<std::string>* errors) {
    bool ok = true;
    auto err = [&](const std::string& m) {
        ok = false;
        if (errors) errors->push_back(m);
    };
    if (!in_range(timestep, kTimestepMin, kTimestepMax))
        err("Simulation.timestep " + io::format_double(timestep) + " not within [1e-09, 3.1556952e+16]");
    const std::string a = lower(algorithm);
    if (a == "all pairs")
        simtype = SimType::ALL_PAIRS;
    else if (a == "barnes-hut")
        simtype = SimType::BARNES_HUT;
    else if (a == "barnes-hut gpu")
        simtype = SimType::BARNES_HUT_GPU;
    else
        err("Simulation.algorithm `" + algorithm +
            "` is not one of `All Pairs`, `Barnes-Hut`, `Barnes-Hut GPU`");
    if (!in_range(theta, 0.0, kThetaMax)) err("Simulation.theta not within [0, 100]");
    if (!in_range(softening_factor, 0.0, kSofteningFactorMax))
        err("Simulation.softening_factor not within [0, 0.2]");
    if (!in_range(threads, kThreadsMin, kThreadsMax)) err("Simulation.threads not within [1, 256]");
    if (has_softening && !(softening >= 0.0 && std::isfinite(softening)))
        err("Simulation.softening must be finite and >= 0");
    if (has_G && !std::isfinite(G)) err("Simulation.G must be finite");
    return ok;
}

Config Config::load(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to parse config: `" + path + "` (cannot open)");
    std::stringstream ss;
    ss << in.rdbuf();
    Config cfg;
    try {
        const Json j = Parser(ss.str()).parse();
        const Json& jio = j.at("IO");
        cfg.io.echo_config = jio.at("echo_config").boolean();
        cfg.io.universe_infile = jio.at("universe_infile").string();
        cfg.io.universe_outfile = jio.at("universe_outfile").string();
        cfg.io.echo_bodies = jio.at("echo_bodies").boolean();
        const Json& js = j.at("Simulation");
        SimConfig& s = cfg.sim;
        s.timestep = js.at("timestep").number();
        const double it = js.at("iterations").number();
        if (!(it >= 0.0) || it > 1.8e19) throw std::runtime_error("iterations out of range");
        s.iterations = (uint64_t)it;
        s.algorithm = js.at("algorithm").string();
        s.theta = js.at("theta").number();
        s.softening_factor = js.at("softening_factor").number();
        const double th = js.at("threads").number();
        s.threads = (th >= -1e9 && th <= 1e9) ? (int)th : -1;
        if (js.has("G")) {
            s.has_G = true;
            s.G = js.at("G").number();
        }
        if (js.has("softening")) {
            s.has_softening = true;
            s.softening = js.at("softening").number();
        }
        if (js.has("traversal")) {
            const std::string t = lower(js.at("traversal").string());
            if (t == "exact")
                s.traversal = NB_TRAV_EXACT;
            else if (t == "group")
                s.traversal = NB_TRAV_GROUP;
            else
                throw std::runtime_error("Simulation.traversal must be `exact` or `group`");
        }
        if (js.has("softening_seed")) s.softening_seed = (uint64_t)js.at("softening_seed").number();
        // The Graphics section is read by the reference's window only; a
        // headless drop-in accepts and ignores it.
    } catch (const std::exception& e) {
        throw std::runtime_error("Failed to parse config: `" + path + "`: " + e.what());
    }
    std::vector<std::string> errors;
    bool ok = cfg.sim.validate(&errors);
    namespace fs = std::filesystem;
    std::error_code ec;
    if (!fs::is_regular_file(cfg.io.universe_infile, ec)) {
        ok = false;
        errors.push_back("IO.universe_infile `" + cfg.io.universe_infile +
                         "` is not a regular file");
    }
    const fs::path parent = fs::path(cfg.io.universe_outfile).parent_path();
    if (!parent.empty() && !fs::is_directory(parent, ec)) {
        ok = false;
        errors.push_back("IO.universe_outfile parent `" + parent.string() +
                         "` is not a directory");
    }
    if (!ok) {
        std::string m = "Failed to validate config: `" + path + "`";
        for (const auto& e : errors) m += "\n  " + e;
        throw std::runtime_error(m);
    }
    return cfg;
}

}  // namespace nbody2d
