// simulation.cpp — the reference's Simulation state machine (Simulation.cpp:
// 11-72) over the gfx950 C ABI, plus epsilon (Simulation.cpp:82-126) and the
// factory (main.cpp:26-38).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "nbody2d_host.hpp"

namespace nbody2d {

double compute_plummer_softening(const Bodies& bodies, double factor, uint64_t seed,
                                 double max_samples) {
    const uint64_t n = bodies.n();
    if (factor == 0.0 || n <= 1) return 0.0;
    auto dist = [&](uint64_t i, uint64_t j) {
        const double dx = bodies.pos[2 * i] - bodies.pos[2 * j];
        const double dy = bodies.pos[2 * i + 1] - bodies.pos[2 * j + 1];
        return std::sqrt(dx * dx + dy * dy);
    };
    double avg;
    if ((double)(n * n) <= max_samples) {
        double sum = 0.0;
        for (uint64_t i = 0; i < n; i++)
            for (uint64_t j = i + 1; j < n; j++) sum += dist(i, j);
        avg = sum / (double)n;  // the reference divides by n (Simulation.cpp:99)
    } else {
        std::mt19937_64 rng(seed);
        std::uniform_int_distribution<uint64_t> uni(0, n - 1);
        const uint64_t samples = (uint64_t)max_samples;
        double sum = 0.0;
        for (uint64_t s = 0; s < samples; s++) {
            const uint64_t i = uni(rng);
            uint64_t j = uni(rng);
            while (i == j) j = uni(rng);
            sum += dist(i, j);
        }
        avg = sum / max_samples;
    }
    return avg * factor;
}

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

Simulation::Simulation(const SimConfig& cfg, Bodies& bodies)
    : bodies_(bodies), cfg_(cfg), requested_dt_(cfg.timestep), dt_(cfg.timestep) {
    std::string why;
    if (!bodies.validate(&why)) throw std::runtime_error("invalid bodies: " + why);
    epsilon_ = cfg.has_softening
                   ? cfg.softening
                   : compute_plummer_softening(bodies, cfg.softening_factor, cfg.softening_seed);
    nb_params p;
    nb_params_default(&p);
    p.G = cfg.has_G ? cfg.G : kReferenceG;
    p.dt = cfg.timestep;
    p.theta = cfg.theta;
    p.softening = epsilon_;
    p.traversal = cfg.traversal;
    switch (cfg.simtype) {
    case SimType::ALL_PAIRS:  // AllPairsSim: drift then kick, Simulation::force
        p.algorithm = NB_ALGO_ALL_PAIRS;
        p.integrator = NB_INTEG_DK;  // AllPairs.cpp:25-26
        p.force_kernel = NB_FORCE_REF_CPU;  // Simulation.cpp:74-80
        break;
    case SimType::BARNES_HUT:  // BarnesHut: the CPU quadtree, diagonal test, CPU force
        p.algorithm = NB_ALGO_BARNES_HUT;
        p.integrator = NB_INTEG_KD;  // BarnesHut.cpp:78-79
        p.force_kernel = NB_FORCE_REF_CPU;  // BarnesHut.cpp:130-135
        p.mac = NB_MAC_DIAG;                // BarnesHut.cpp:113-115
        p.cells = NB_CELLS_REF_CPU;         // Quadtree.cpp:9-152
        break;
    case SimType::BARNES_HUT_GPU:  // BarnesHutCuda: Morton cells, side test, Plummer
        p.algorithm = NB_ALGO_BARNES_HUT;
        p.integrator = NB_INTEG_KD;  // BarnesHutCuda.cpp:35-36
        break;
    }
    const int st = nb_create(&p, bodies.n(), bodies.mass.data(), bodies.pos.data(),
                             bodies.vel.data(), &ctx_);
    if (st != NB_OK)
        throw std::runtime_error(std::string("nb_create: ") + nb_status_string(st) + ": " +
                                 nb_last_error(nullptr));
}

Simulation::~Simulation() {
    if (thread_.joinable()) {
        stop_ = true;
        thread_.join();
    }
    nb_destroy(ctx_);
}

void Simulation::check(int status, const char* what) const {
    if (status != NB_OK)
        throw std::runtime_error(std::string(what) + ": " + nb_status_string(status) + ": " +
                                 nb_last_error(ctx_));
}

Simulation::State Simulation::get_state() {
    std::lock_guard<std::mutex> g(state_mtx_);
    return state_;
}

Simulation::Stats Simulation::get_stats() {
    std::lock_guard<std::mutex> g(stats_mtx_);
    return stats_;
}

double Simulation::elapsed() const {
    return run_accum_ + (state_ == State::RUNNING ? now_s() - run_started_ : 0.0);
}

void Simulation::run() {
    std::lock_guard<std::mutex> g(state_mtx_);
    if (state_ != State::PAUSED) {
        std::fprintf(stderr, "warning: cannot run simulation, it is not paused\n");
        return;
    }
    state_ = State::RUNNING;
    stop_ = false;
    run_started_ = now_s();
    thread_ = std::thread(&Simulation::loop, this);
}

void Simulation::pause() {
    std::lock_guard<std::mutex> g(state_mtx_);
    if (state_ != State::RUNNING) {
        std::fprintf(stderr, "warning: cannot pause simulation, it is not running\n");
        return;
    }
    stop_ = true;
    if (thread_.joinable()) thread_.join();
    run_accum_ += now_s() - run_started_;
    state_ = State::PAUSED;
}

bool Simulation::should_stop() {
    if (iteration_ >= cfg_.iterations) {
        finished_ = true;
        return true;
    }
    return stop_;
}

// One iteration of the reference's loop body: step, refresh Bodies in the
// caller's order (BarnesHutCuda.cu:632-646), then post_iteration
// (Simulation.cpp:68-72): stats, dt latch, iteration++. The un-permute runs
// on the GPU and the copy lands in pinned memory (a snapshot slot).
void Simulation::step_once() {
    check(nb_step(ctx_, 1), "nb_step");
    uint64_t ticket = 0;
    const double *p = nullptr, *v = nullptr;
    check(nb_snapshot_begin(ctx_, &ticket), "nb_snapshot_begin");
    check(nb_snapshot_wait(ctx_, ticket, &p, &v, nullptr), "nb_snapshot_wait");
    const uint64_t n = bodies_.n();
    if (n) {
        std::memcpy(bodies_.pos.data(), p, 2 * n * sizeof(double));
        std::memcpy(bodies_.vel.data(), v, 2 * n * sizeof(double));
    }
    sim_time_ += dt_;
    iteration_++;
    {
        std::lock_guard<std::mutex> g(stats_mtx_);
        const double t = elapsed();
        const double dt_wall = t - stats_.real_elapsed_s;
        if (dt_wall > 0) stats_.ips = (float)((double)(iteration_ - stats_.iteration) / dt_wall);
        stats_.iteration = iteration_;
        stats_.real_elapsed_s = t;
        stats_.simulated_elapsed_s = sim_time_;
    }
    const double want = requested_dt_.load();
    if (want != dt_) {
        check(nb_set_dt(ctx_, want), "nb_set_dt");
        dt_ = want;
    }
}

void Simulation::loop() {
    try {
        while (!should_stop()) step_once();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        finished_ = true;
    }
}

void Simulation::run_to_completion() {
    {
        std::lock_guard<std::mutex> g(state_mtx_);
        if (state_ != State::PAUSED) throw std::runtime_error("simulation is running");
        state_ = State::RUNNING;
        run_started_ = now_s();
    }
    // Nobody reads Bodies between steps here, so steps go in batches and
    // Bodies is refreshed once per batch; dt cannot change meanwhile.
    while (iteration_ < cfg_.iterations) {
        const uint64_t k = std::min<uint64_t>(cfg_.iterations - iteration_, 1000);
        check(nb_step(ctx_, k), "nb_step");
        iteration_ += k;
        sim_time_ += (double)k * dt_;
    }
    check(nb_get_state(ctx_, bodies_.pos.data(), bodies_.vel.data()), "nb_get_state");
    finished_ = true;
    std::lock_guard<std::mutex> g(state_mtx_);
    run_accum_ += now_s() - run_started_;
    state_ = State::PAUSED;
    std::lock_guard<std::mutex> g2(stats_mtx_);
    stats_.iteration = iteration_;
    stats_.real_elapsed_s = run_accum_;
    stats_.simulated_elapsed_s = sim_time_;
    stats_.ips = run_accum_ > 0 ? (float)((double)iteration_ / run_accum_) : 0.f;
}

std::unique_ptr<Simulation> create_sim(const SimConfig& cfg, Bodies& bodies) {
    return std::make_unique<Simulation>(cfg, bodies);
}

}  // namespace nbody2d
