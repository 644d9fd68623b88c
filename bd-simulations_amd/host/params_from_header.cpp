// Gathers the constants of a reference-format parameter.h (parameter.h:14-99)
// into a RunConfig.  Compile this file with -I<dir of your parameter.h>, as
// the reference compiles every translation unit against its parameter.h.
// Environment overrides: BDK_SEED (MT seed; the reference uses time(NULL)),
// BDK_NOISE=mt|disc (host MT stream uploaded per step, or device Philox
// disc noise), BDK_PRECISION=fp64|fp32, BDK_MAX_TRAVEL (max_reisedistanz),
// BDK_MODE=run|debug|krafttest (sets debugmode / krafttestmode),
// BDK_KORRFUNK=0|1|mixed (auswerten_korrfunk, auswerten_korrfunk_mixed),
// BDK_RHOFFTW=0|1 (auswerten_rhoviaFFTW), BDK_ABSTAND=0|1 (auswerten_abstand),
// BDK_RHOVONK=0|1 (auswerten_rhovonk).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>

#include "parameter.h"
#include "run_zustand.hpp"

RunConfig config_from_parameter_h() {
    RunConfig c;
    bdk_params_default(&c.p);
    c.p.n1 = N1;
    c.p.n2 = N2;
    c.p.L = L;
    c.p.scheme = densGrid_Schema;
    c.p.Z = FFT_samples;
    c.p.lambda = lambda_kapillar;
    c.p.kap = kapillar_vorfaktor;
    c.p.T = T;
    c.p.dt_max = dt_max;
    c.p.max_travel = max_reisedistanz;
    c.p.sigma11_22 = sigma11_22;
    c.p.sigma11_12 = sigma11_12;
    c.p.eps22_11 = eps22_11;
    c.p.eps12_11 = eps12_11;
    c.p.f2_f1 = f2_f1;
    c.p.max_pair_wca = maxPairWCA;
    c.p.idgas_only = idgas_only;
    c.p.no_wca = noWCA;
    c.p.restrict_radial = restrictRadial;
    c.runs = runs;
    c.obs_dt = obs_dt;
    c.obs_anzahl = obs_anzahl;
    c.startpos_methode = startpos_methode;
    c.max_write_interval = max_write_interval;
    c.auswerten_animation = auswerten_animation;
    c.krafttestmode = krafttestmode;
    c.debugmode = debugmode;
    c.korrfunk = auswerten_korrfunk;
    c.korrfunk_mixed = auswerten_korrfunk_mixed;
    c.korr_bins = korr_bins;
    c.korr_dr = korr_dr;
    c.rhofftw = auswerten_rhoviaFFTW;
    c.rhofftw_bins = rhoFFTW_bins;
    c.dq_rhofftw = dq_rhoFFTW;
    c.dq_dqft = dq_dqFT;
    if (const char *r = std::getenv("BDK_RHOFFTW")) c.rhofftw = std::strcmp(r, "0") != 0;
    c.abstand = auswerten_abstand;
    c.rhovonk = auswerten_rhovonk;
    c.ftrho_qbins = ftrho_qbins;
    c.ftrho_dq = ftrho_dq;
    c.dq = dq;
    if (const char *v = std::getenv("BDK_RHOVONK")) c.rhovonk = std::strcmp(v, "0") != 0;
    if (const char *a = std::getenv("BDK_ABSTAND")) c.abstand = std::strcmp(a, "0") != 0;
    if (const char *k = std::getenv("BDK_KORRFUNK")) {  // 1: g11, g22; mixed: also g12
        c.korrfunk = std::strcmp(k, "0") != 0;
        c.korrfunk_mixed = std::strcmp(k, "mixed") == 0;
    }
    if (const char *m = std::getenv("BDK_MODE")) {
        c.debugmode = std::strcmp(m, "debug") == 0;
        c.krafttestmode = std::strcmp(m, "krafttest") == 0;
    }
    const char *s = std::getenv("BDK_SEED");
    c.seed = s ? std::strtoull(s, nullptr, 10) : (uint64_t)std::time(nullptr);
    const char *nz = std::getenv("BDK_NOISE");
    c.noise_source = (nz && std::strcmp(nz, "mt") == 0) ? 1 : 0;
    c.p.noise_mode = c.noise_source ? BDK_NOISE_UPLOAD : BDK_NOISE_DISC;
    c.p.seed = c.seed;
    const char *pr = std::getenv("BDK_PRECISION");
    c.p.precision = (pr && std::strcmp(pr, "fp64") == 0) ? BDK_FP64 : BDK_FP32;
    if (const char *mt = std::getenv("BDK_MAX_TRAVEL")) c.p.max_travel = std::strtod(mt, nullptr);
    return c;
}
