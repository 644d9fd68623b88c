// Driver with the run loop of main.cpp:130-300 over the GPU timestep (the
// parameter.h switches can be overridden by environment variables, see
// params_from_header.cpp):
// for each run, init, obs_point(0), then obs_anzahl-1 checkpoint intervals
// of zeitschritte_bis_obs + obs_point, writing pos1_<run+1>.txt and
// pos2_<run+1>.txt; or, as in main.cpp, the mrd sweep (debugmode) or the
// pair-force scan (krafttestmode) on the first run's start.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iostream>

#include "run_zustand.hpp"

namespace {

// debugmode (main.cpp:156-216): for travel caps mrd = 1, 0.95, ... > 0.01, a
// fresh start, then the mean and largest |F_WCA|max and the mean dt over 50
// steps early (mrd_data1.txt) and over 50 steps after max(50, 1.5/mrd)
// ordinary steps (mrd_data2.txt).
void mrd_sweep(RunZustand &theRun) {
    RunZustand::RunDynamik::TimestepInfo info;
    FILE *f1 = std::fopen("mrd_data1.txt", "w");
    FILE *f2 = std::fopen("mrd_data2.txt", "w");
    if (!f1 || !f2) {
        std::cout << "cannot open mrd_data1.txt / mrd_data2.txt" << std::endl;
        std::exit(1);
    }
    std::fprintf(f1, "# means over the first 50 steps after the start\n");
    std::fprintf(f2, "# means over 50 steps after the skipped ones\n");
    std::fprintf(f1, "# Format: mrd TAB <F_WCA_max> TAB |F_WCA_highest| TAB dt \n\n");
    std::fprintf(f2, "# Format: mrd TAB <F_WCA_max> TAB |F_WCA_highest| TAB dt \n\n");
    auto window = [&](double mrd, FILE *f) {
        double F_mean = 0.0, F_high = 0.0, dt_mean = 0.0;
        for (int k = 0; k < 50; ++k) {
            theRun.dyn.zeitschritt_debug(mrd, info, true);
            dt_mean += info.dt / 50.0;
            const double F = std::sqrt(info.maxWCAx * info.maxWCAx + info.maxWCAy * info.maxWCAy);
            F_mean += F / 50.0;
            if (F > F_high) F_high = F;
        }
        std::fprintf(f, "%g \t %g \t %g \t %g \n", mrd, F_mean, F_high, dt_mean);
    };
    for (double mrd = 1.0; mrd > 0.01; mrd *= 0.95) {
        std::cout << "mrd=" << mrd << std::endl;
        theRun.init(0);
        window(mrd, f1);
        int skip = (int)(1.5 / mrd);
        if (skip < 50) skip = 50;
        for (int k = 0; k < skip; ++k) theRun.dyn.zeitschritt(0.1);
        window(mrd, f2);
    }
    std::fclose(f1);
    std::fclose(f2);
}

// krafttestmode (main.cpp:220-248): particle 1 of species 1 moved to
// y = x_0 + t for t = 0.001, 0.011, ... < 5 (x_0 the x coordinate of
// particle 0, as the reference takes its origin; x of particle 1 kept), the
// forces recomputed, and F1kap[1][1] and F1kap[1][1] + F1_WCA[1][1] written
// to kapkraft.txt.  y is wrapped into the box (the reference indexes past the
// cell range there).
void force_scan(RunZustand &theRun, const RunConfig &cfg) {
    double der[6];
    bdk_params_derived(&cfg.p, der);
    const double B = der[2];
    FILE *out = std::fopen("kapkraft.txt", "w");
    if (!out) {
        std::cout << "cannot open kapkraft.txt" << std::endl;
        std::exit(1);
    }
    std::fprintf(out, "#Format: r Fkap Ftotal\n\n");
    for (double t = 0.001; t < 5.0; t += 0.01) {
        const double origin = theRun.r1_git[0][0] * B + theRun.r1_rel[0][0];
        const double y = std::fmod(origin + t, cfg.p.L);
        theRun.r1_git[1][1] = (int)(y / B);
        theRun.r1_rel[1][1] = y - theRun.r1_git[1][1] * B;
        // the shim uploads the edited arrays before the next force call
        theRun.dyn.berechne_kapkraefte();
        theRun.dyn.berechne_WCA11();
        std::fprintf(out, "%g \t %g \t %g \n", t, theRun.dyn.getFKap(), theRun.dyn.getF());
    }
    std::fclose(out);
}

}  // namespace

int main() {
    const RunConfig cfg = config_from_parameter_h();  // the loop bounds and modes of the runs below
    const time_t started = std::time(nullptr);
    for (int run = 0; run < cfg.runs; ++run) {
        RunZustand theRun;  // main.cpp:132: configured from parameter.h
        theRun.ausgabe_jeansgroessen();
        theRun.init(run);
        theRun.obs_point(0);
        if (cfg.debugmode) {  // main.cpp:156-216
            theRun.obs_point(0);
            mrd_sweep(theRun);
            return 0;
        }
        if (cfg.krafttestmode) {  // main.cpp:220-248
            force_scan(theRun, cfg);
            return 0;
        }
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
        theRun.schreibe_obs();  // main.cpp:296
    }
    std::cout << "done after " << std::time(nullptr) - started << " s" << std::endl;
    return 0;
}
