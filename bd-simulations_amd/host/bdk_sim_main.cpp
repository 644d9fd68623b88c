// Driver with the run loop of main.cpp:130-300 over the GPU timestep:
// for each run, init, obs_point(0), then obs_anzahl-1 checkpoint intervals
// of zeitschritte_bis_obs + obs_point, writing pos1_<run+1>.txt and
// pos2_<run+1>.txt.
#include <cstdio>
#include <ctime>
#include <iostream>

#include "run_zustand.hpp"

int main() {
    RunConfig cfg = config_from_parameter_h();
    if (cfg.krafttestmode) {
        std::cout << "krafttestmode: use the force entry points (bdk_compute_wca / bdk_compute_capillary)"
                  << std::endl;
        return 0;
    }
    const time_t started = std::time(nullptr);
    for (int run = 0; run < cfg.runs; ++run) {
        RunZustand theRun(cfg);
        theRun.init(run);
        theRun.obs_point(0);
        std::cout << "Run " << run << ": initial positions written, starting timesteps." << std::endl;
        time_t last_write = std::time(nullptr);
        for (int obs_nr = 1; obs_nr < cfg.obs_anzahl; ++obs_nr) {
            theRun.zeitschritte_bis_obs(last_write);
            const int numSteps = theRun.schritte_seit_obs;
            const time_t now = std::time(nullptr);
            const double sps = (now != last_write) ? numSteps / double(now - last_write) : 9999.0;
            std::cout << "Run " << run << ": recording obs-Point " << theRun.obs_nr + 1 << " / " << cfg.obs_anzahl
                      << ". " << numSteps << " steps since CP. Performance: " << sps << " steps/s." << std::endl;
            theRun.obs_point(obs_nr);
            last_write = std::time(nullptr);
        }
    }
    std::cout << "done after " << std::time(nullptr) - started << " s" << std::endl;
    return 0;
}
