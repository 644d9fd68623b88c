// C++ host shim with the interface of RunZustand / RunZustand::RunDynamik
// (signaturen2.h:29-434), running the timestep on the GPU through the C-ABI
// of include/bdk.h.  Drivers written against the reference (main.cpp,
// main_profile.cpp) keep their calls: init(run), obs_point(nr),
// zeitschritte_bis_obs(...), dyn.zeitschritt(tmax), dyn.berechne_*(),
// dyn.getFKap(), dyn.getF().
//
// Positions keep the reference's layout, int** r*_git and double** r*_rel
// of shape [N][2] (s2:380-383).  They are the truth at init() and are
// refreshed from the device at every obs_point() and by sync_positions().
// A driver may also write them directly, as the reference's force test does
// (main.cpp:220-248, which moves a particle and recomputes the forces): the
// next force or step call sees that the arrays differ from what the device
// holds and uploads them first.
#pragma once

#include <ctime>
#include <cstdio>
#include <string>
#include <vector>

#include "bdk.h"

// Run configuration: the constants of a reference-format parameter.h,
// gathered by params_from_header.cpp (compiled against that header).
struct RunConfig {
    bd_params p;
    int runs;
    double obs_dt;
    int obs_anzahl;
    int startpos_methode;
    int max_write_interval;
    bool auswerten_animation;
    bool krafttestmode;
    bool debugmode;
    bool korrfunk, korrfunk_mixed;  // auswerten_korrfunk(_mixed)
    int korr_bins;
    double korr_dr;
    bool rhofftw;             // auswerten_rhoviaFFTW
    bool abstand;             // auswerten_abstand
    bool rhovonk;             // auswerten_rhovonk
    int ftrho_qbins;
    double ftrho_dq, dq;
    int rhofftw_bins;         // rhoFFTW_bins
    double dq_rhofftw, dq_dqft;
    uint64_t seed;      // MT seed (the reference seeds with time(NULL), zufall.cpp:18)
    int noise_source;   // 0: device Philox disc noise; 1: host MT stream uploaded per step
};

RunConfig config_from_parameter_h();

class RunZustand {
  public:
    class RunDynamik {
      public:
        RunDynamik() = default;
        ~RunDynamik();
        RunDynamik(const RunDynamik &) = delete;
        RunDynamik &operator=(const RunDynamik &) = delete;

        // The counters of one step (signaturen2.h:71-145), filled by
        // zeitschritt_debug.
        struct TimestepInfo {
            int num_WCA_berechnet = -1;
            double maxWCAx = 0.0, maxWCAy = 0.0;
            int maxWCAtype = -1, maxWCAindex = -1;
            int maxWCA_numWW1 = -1, maxWCA_numWW2 = -1;
            double maxKapx = 0.0, maxKapy = 0.0;
            int maxKaptype = -1, maxKapindex = -1;
            double maxRNDx = 0.0, maxRNDy = 0.0;
            int maxRNDtype = -1, maxRNDindex = -1;
            double dt = -1.0;
            int num_Zellwechsel1 = -1, num_Zellwechsel2 = -1;
            void ausgabe() const;
        };

        // dm:74-163
        double zeitschritt(double tmax);
        // dm:168-211: one step with travel cap mrd, counters into info
        // (print: info.ausgabe())
        void zeitschritt_debug(double mrd, TimestepInfo &info, bool print);
        // am:361-364: lists from the current positions, spectral tables
        void init();
        void berechne_kapkraefte();
        // The device computes the 11, 22 and 12 WCA forces in one pass;
        // berechne_WCA11 runs it, berechne_WCA22 and addiere_WCA12 keep the
        // call sequence of dm:86-88 and are no-ops.
        void berechne_WCA11();
        void berechne_WCA22() {}
        void addiere_WCA12() {}
        void refresh_erwNachbar();
        void berechne_zufallskraefte();
        double optimaler_zeitschritt();
        double getFKap() { return F1kap.size() > 3 ? F1kap[3] : 0.0; }                 // s2:182-185
        double getF() { return getFKap() + (F1_WCA.size() > 3 ? F1_WCA[3] : 0.0); }   // s2:186-188

        // pull device positions into the aliased host arrays
        void sync_positions();
        // push the aliased host arrays to the device (after an external change)
        void push_positions();
        // push_positions() if the aliased arrays changed since the last
        // push or sync (a driver wrote them)
        void push_if_changed();

        int **r1_git = nullptr, **r2_git = nullptr;
        double **r1_rel = nullptr, **r2_rel = nullptr;

        bdk_t *h = nullptr;
        const RunConfig *cfg = nullptr;
        bdk_mt_t *mt = nullptr;  // MT stream for the upload noise source
        std::vector<double> F1kap, F2kap, F1_WCA, F2_WCA;

      private:
        // what the device holds, species 1 then 2, as of the last push or sync
        std::vector<int32_t> dev_git_;
        std::vector<double> dev_rel_;
        void remember_device_positions();
    };

    RunDynamik dyn;

    class RunObs {
      public:
        int **r1_git = nullptr, **r2_git = nullptr;
        double **r1_rel = nullptr, **r2_rel = nullptr;
        void obs_init(int run, const RunConfig &cfg);  // au:52-75
        void schreibe_pos(int nr);                     // au:631-648
        // g11, g12, g22 of checkpoint nr (record_korrelationsfunktion11/12/22,
        // au:110-188) and their file at the end of the run (au:934-975)
        void record_korrelationsfunktion(bdk_t *h, int nr);
        void schreibe_korrfunk(int run) const;
        std::vector<std::vector<double>> g11, g12, g22;
        // rho(k) shells of each species at checkpoint nr (record_rhoFFTW,
        // au:527-626) and their file at the end of the run (au:994-1014)
        void record_rhoFFTW(bdk_t *h, int nr);
        void schreibe_rhoFFT(int run) const;
        // mean distance of the species-2 particles from the box centre
        // (schreibe_abstaende, au:651-670) into abstaende2.txt
        void schreibe_abstaende(int nr);
        // rho(k) by direct sums on the (i, j) grid, shells by i^2 + j^2
        // (init_ftrho au:262-300, record_ftrho_unkorrigiert au:307-360) and
        // their file at the end of the run (schreibe_ftrho, au:977-992)
        void record_ftrho(bdk_t *h, int nr);
        void schreibe_ftrho(int run) const;
        std::vector<int> ftrho_order, ftrho_qabs2;
        std::vector<std::vector<double>> ft1_re, ft1_im, ft2_re, ft2_im;
        FILE *ab2 = nullptr;
        std::vector<std::vector<double>> rho1_re, rho1_im, rho2_re, rho2_im;
        void close();
        std::string path1, path2;
        const RunConfig *cfg = nullptr;
        double B = 0.0;
    };
    RunObs obs;

    // The reference's `RunZustand theRun;` (main.cpp:132): the run
    // configuration comes from the parameter.h that params_from_header.cpp
    // was compiled against.
    RunZustand();
    explicit RunZustand(const RunConfig &cfg) : cfg_(cfg) {}
    ~RunZustand();
    RunZustand(const RunZustand &) = delete;
    RunZustand &operator=(const RunZustand &) = delete;

    const RunConfig &config() const { return cfg_; }

    int nr = 0;
    long double t = 0.0L;
    int obs_nr = 0;
    int schritte_seit_obs = 0;

    int **r1_git = nullptr, **r2_git = nullptr;
    double **r1_rel = nullptr, **r2_rel = nullptr;

    void init(int nr);                            // am:256-320
    void zeitschritte_bis_obs(time_t last_write);  // dm:444-455
    void obs_point(int nr_obspoint);              // au:78-93
    void schreibe_obs();                          // au:912-928
    // Jeans time and wave number of the run's mean density (am:324-348),
    // printed to stdout
    void ausgabe_jeansgroessen() const;

  private:
    RunConfig cfg_;
    std::vector<int> git_store1_, git_store2_;
    std::vector<double> rel_store1_, rel_store2_;
    std::vector<int *> git_ptr1_, git_ptr2_;
    std::vector<double *> rel_ptr1_, rel_ptr2_;
    void init_zufallspos();  // dm:1647-1678
};
