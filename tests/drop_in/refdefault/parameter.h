// Run constants in the format of the reference's parameter.h (same names,
// same meaning, parameter.h:14-99 and 167-191).  Values: the reference's own
// defaults (parameter.h:9-101) -- L = 300, N1 = N2 = 2000, TSC on 384^2,
// lambda = 25, kapillar_vorfaktor = 1.111111111, q2/q1 = -1, kT = 24.3055555,
// dt_max = 0.005 with max_reisedistanz = 0.24 (adaptive dt), 12 checkpoints
// of 0.2.
#pragma once

#define _USE_MATH_DEFINES
#include <math.h>
#include <string>

using std::string;

const int N1 = 2000;
const int N2 = 2000;
const double L = 300.0;
const int densGrid_Schema = 2;      // 0 NGP, 1 CIC, 2 TSC
const int FFT_samples = 384;        // 2^7 * 3
const double lambda_kapillar = 25.0;
const double kapillar_vorfaktor = 1.111111111;
const double T = 24.3055555;
const double dt_max = 0.005;
const double max_reisedistanz = 0.24;
const double sigma11_22 = 1.0;
const double sigma11_12 = 0.5 * (1.0 + sigma11_22);
const double Gamma2_1 = sigma11_22;
const double eps22_11 = 1.0;
const double eps12_11 = 1.0;
const double f2_f1 = -1.0;
const bool idgas_only = false;
const bool noWCA = false;
const bool noRNG = false;
const bool quickInit = true;
const bool restrictRadial = false;
const bool debugmode = false;
const bool krafttestmode = false;
const double maxPairWCA = 1.0e5;
const int max_write_interval = 900;
const int runs = 1;
const double obs_dt = 0.2;
const int obs_anzahl = 12;
const int startpos_methode = 1;
const bool auswerten_animation = true;
const bool auswerten_korrfunk = false;        // g11, g22 at every checkpoint
const bool auswerten_korrfunk_mixed = false;  // and g12
const double korr_dr = 0.2;
const double korr_rmax = L * 0.5;
const int korr_bins = (int)(korr_rmax / korr_dr) + 1;
const bool auswerten_rhoviaFFTW = false;  // rho(k) shells at every checkpoint
const bool auswerten_abstand = false;     // mean distance of species 2 from the box centre
const bool auswerten_rhovonk = false;     // rho(k) by direct sums on a k grid
const int ftrho_qbins = 30;

const double zweihoch1_6 = pow(2.0, 1.0 / 6.0);
const int nachList_Zellen = (int)floor(L / zweihoch1_6);
const double nachList_Breite = L / ((double)nachList_Zellen);
const double densGrid_Breite = L / FFT_samples;
const int densGrid_Zellen = FFT_samples;
const double dq = 2.0 * M_PI / L;
const double dq_rhoFFTW = 2.0 * dq;
const double dq_dqFT = dq / dq_rhoFFTW;
const int rhoFFTW_bins = 0.5 * FFT_samples * dq / dq_rhoFFTW;
const double ftrho_dq = 4.0 * M_PI / L;
