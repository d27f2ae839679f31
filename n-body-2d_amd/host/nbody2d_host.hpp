// nbody2d_host.hpp — C++ host side of the drop-in: the reference's Bodies /
// Config::Simulation / IO / Simulation interfaces, backed by the gfx950 C ABI
// (include/nbody.h).
//
//   Bodies               src/common_lib/Body/include/Body/Body.hpp:10-40
//   SimConfig, Config    src/common_lib/Config/include/Config/Config.hpp:11-59
//   io::parse_csv/write  src/common_lib/InputOutput/include/InputOutput/InputOutput.hpp:11-12
//   Simulation           src/Simulation/include/Simulation/Simulation.hpp:13-67
//   create_sim           src/main.cpp:26-38
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "nbody.h"

namespace nbody2d {

// Reference constants (Constants.hpp:14-29).
constexpr double kReferenceG = 6.67430e-11;
constexpr double kTimestepMin = 1e-9, kTimestepMax = 31'556'952'000'000'000.0;
constexpr double kSofteningFactorMax = 0.2;
constexpr int kThreadsMin = 1, kThreadsMax = 256;
constexpr double kThetaMax = 100.0;
constexpr uint64_t kMaxPairwiseSofteningComputations = 1'000'000;

// Body state in the caller's order: interleaved x,y FP64 like
// std::vector<sf::Vector2<double>> (Body.hpp:34-37).
struct Bodies {
    std::vector<std::string> id;
    std::vector<double> mass;
    std::vector<double> pos;  // x0, y0, x1, y1, ...
    std::vector<double> vel;
    uint64_t n() const { return mass.size(); }
    bool validate(std::string* why = nullptr) const;  // unique ids (Body.cpp:99-113)
};

enum class SimType { ALL_PAIRS, BARNES_HUT, BARNES_HUT_GPU };

// Simulation section of config.json (Config.cpp:23-30) plus the keys this
// build adds: "G" (the reference hard-codes it), "softening" (absolute
// epsilon; overrides "softening_factor"), "traversal" ("exact" | "group").
struct SimConfig {
    double timestep = 1e-3;
    uint64_t iterations = 0;
    std::string algorithm = "barnes-hut gpu";
    SimType simtype = SimType::BARNES_HUT_GPU;
    double theta = 0.5;
    double softening_factor = 0.0;
    int threads = 1;
    bool has_G = false;
    double G = kReferenceG;
    bool has_softening = false;
    double softening = 0.0;
    int traversal = NB_TRAV_EXACT;
    uint64_t softening_seed = 0x5eed;  // seed of the sampled epsilon branch

    // Range checks of Config::Simulation::validate (Config.cpp:224-253);
    // appends messages to *errors and returns false on any failure.
    bool validate(std::vector<std::string>* errors);
};

struct IOConfig {
    std::string universe_infile, universe_outfile;
    bool echo_config = false, echo_bodies = false;
};

struct Config {
    IOConfig io;
    SimConfig sim;
    static Config load(const std::string& path);  // throws std::runtime_error
};

namespace io {
// IO::parse_csv (InputOutput.cpp:18-52): header row, positional columns
// id,mass,x,y,vel_x,vel_y, numbers read with std::stod's prefix rule.
Bodies parse_csv(const std::string& path);
// IO::write_csv (InputOutput.cpp:54-72): same columns, shortest round-trip numbers.
void write_csv(const std::string& path, const Bodies& bodies);
std::string format_double(double v);
}  // namespace io

// Simulation::compute_plummer_softening (Simulation.cpp:82-126): epsilon =
// factor x mean pairwise distance. Exact branch (n^2 <= max_samples) divides
// the pair sum by n, as the reference does; otherwise max_samples random
// pairs from a seeded generator (the reference seeds from the clock).
double compute_plummer_softening(const Bodies& bodies, double factor, uint64_t seed,
                                 double max_samples = (double)kMaxPairwiseSofteningComputations);

// Simulation (Simulation.hpp:13-67): run() starts a stepping thread, pause()
// stops and joins it, set_timestep() is latched after the current step,
// Bodies is updated in original order after every step.
class Simulation {
public:
    enum class State : bool { PAUSED, RUNNING };
    struct Stats {
        uint64_t iteration = 0;
        float ips = 0.0f;
        double real_elapsed_s = 0.0;
        double simulated_elapsed_s = 0.0;
    };

    Simulation(const SimConfig& cfg, Bodies& bodies);
    virtual ~Simulation();
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    State get_state();
    Stats get_stats();
    bool is_finished() const { return finished_; }
    void run();
    void pause();
    void set_timestep(double dt) { requested_dt_.store(dt); }
    double epsilon() const { return epsilon_; }
    // Runs steps on the calling thread until is_finished() (headless use).
    void run_to_completion();

protected:
    Bodies& bodies_;
    SimConfig cfg_;
    double epsilon_;
    nb_ctx* ctx_ = nullptr;
    std::atomic<double> requested_dt_;
    double dt_;
    uint64_t iteration_ = 0;
    std::atomic<bool> finished_{false};
    std::atomic<bool> stop_{false};

private:
    std::mutex state_mtx_, stats_mtx_;
    State state_ = State::PAUSED;
    Stats stats_{};
    std::thread thread_;
    double run_started_ = 0.0, run_accum_ = 0.0;
    double sim_time_ = 0.0;

    bool should_stop();
    void step_once();
    void loop();
    double elapsed() const;
    void check(int status, const char* what) const;
};

// create_sim (main.cpp:26-38). Every algorithm string maps to the GPU with
// the physics of the reference path it names: "all pairs" -> direct sum,
// drift then kick, the CPU force law (AllPairs.cpp, Simulation.cpp:74-80);
// "barnes-hut" -> the CPU quadtree (rectangular cells, `else if` box), the
// diagonal opening test and the CPU force law, kick then drift
// (BarnesHut.cpp, Quadtree.cpp); "barnes-hut gpu" -> square Morton cells, the
// side test and the Plummer law, kick then drift (BarnesHutCuda.cu).
std::unique_ptr<Simulation> create_sim(const SimConfig& cfg, Bodies& bodies);

}  // namespace nbody2d
