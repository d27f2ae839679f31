// nbody2d_run.cpp — headless driver: the reference's main (main.cpp:40-61)
// without the window. Reads config.json and the universe CSV, runs the
// configured number of iterations on the GPU, writes the final universe CSV.
//
//   nbody2d-run <config.json>             run and write IO.universe_outfile
//   nbody2d-run <config.json> --check     parse + validate config and CSV only
//   nbody2d-run --echo-csv <in> <out>     parse a universe CSV and write it back
//   nbody2d-run --softening <in> <factor> epsilon of compute_plummer_softening
//   nbody2d-run <config.json> --run-pause run() / pause() / set_timestep(2 dt) /
//                                         run() from a second thread, as the
//                                         reference's controller does
#include <chrono>
#include <cstdio>
#include <thread>
#include <cstring>
#include <string>

#include "nbody2d_host.hpp"

using namespace nbody2d;

static int usage() {
    std::fprintf(stderr,
                 "usage: nbody2d-run <config.json> [--check]\n"
                 "       nbody2d-run --echo-csv <in.csv> <out.csv>\n"
                 "       nbody2d-run --softening <in.csv> <factor>\n");
    return 2;
}

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::strcmp(argv[1], "--echo-csv") == 0) {
            if (argc != 4) return usage();
            Bodies b = io::parse_csv(argv[2]);
            io::write_csv(argv[3], b);
            std::printf("{\"bodies\": %llu}\n", (unsigned long long)b.n());
            return 0;
        }
        if (argc >= 2 && std::strcmp(argv[1], "--softening") == 0) {
            if (argc != 4) return usage();
            Bodies b = io::parse_csv(argv[2]);
            const double eps = compute_plummer_softening(b, std::stod(argv[3]), 0x5eed);
            std::printf("{\"epsilon\": %s}\n", io::format_double(eps).c_str());
            return 0;
        }
        if (argc < 2) return usage();
        const Config cfg = Config::load(argv[1]);
        Bodies bodies = io::parse_csv(cfg.io.universe_infile);
        if (argc >= 3 && std::strcmp(argv[2], "--check") == 0) {
            std::printf("{\"bodies\": %llu, \"iterations\": %llu, \"algorithm\": \"%s\"}\n",
                        (unsigned long long)bodies.n(), (unsigned long long)cfg.sim.iterations,
                        cfg.sim.algorithm.c_str());
            return 0;
        }
        std::unique_ptr<Simulation> sim = create_sim(cfg.sim, bodies);
        if (argc >= 3 && std::strcmp(argv[2], "--run-pause") == 0) {
            sim->run();
            while (sim->get_stats().iteration < 1 && !sim->is_finished())
                std::this_thread::yield();
            sim->pause();
            const Simulation::Stats first = sim->get_stats();
            sim->set_timestep(2.0 * cfg.sim.timestep);
            sim->run();
            while (!sim->is_finished()) std::this_thread::sleep_for(std::chrono::milliseconds(1));
            sim->pause();
            const Simulation::Stats st = sim->get_stats();
            io::write_csv(cfg.io.universe_outfile, bodies);
            std::printf("{\"paused_at\": %llu, \"iteration\": %llu, \"simulated_elapsed_s\": %s, "
                        "\"finished\": %s}\n",
                        (unsigned long long)first.iteration, (unsigned long long)st.iteration,
                        io::format_double(st.simulated_elapsed_s).c_str(),
                        sim->is_finished() ? "true" : "false");
            return 0;
        }
        sim->run_to_completion();
        io::write_csv(cfg.io.universe_outfile, bodies);
        const Simulation::Stats st = sim->get_stats();
        std::printf("{\"bodies\": %llu, \"iterations\": %llu, \"epsilon\": %s, "
                    "\"real_elapsed_s\": %.6f, \"simulated_elapsed_s\": %s}\n",
                    (unsigned long long)bodies.n(), (unsigned long long)st.iteration,
                    io::format_double(sim->epsilon()).c_str(), st.real_elapsed_s,
                    io::format_double(st.simulated_elapsed_s).c_str());
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
