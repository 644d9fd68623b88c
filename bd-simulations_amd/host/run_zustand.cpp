// RunZustand / RunDynamik over the C-ABI (see run_zustand.hpp).
#include "run_zustand.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>

namespace {

void check(bdk_t *h, int rc, const char *what) {
    if (rc != BDK_OK) {
        std::cerr << "bdk: " << what << " failed (" << rc << "): " << bdk_last_error(h) << std::endl;
        std::exit(1);
    }
}

void gather_pos(int n, int **git, double **rel, std::vector<int32_t> &g, std::vector<double> &r) {
    g.resize(2 * (size_t)n);
    r.resize(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        g[2 * i] = git[i][0];
        g[2 * i + 1] = git[i][1];
        r[2 * i] = rel[i][0];
        r[2 * i + 1] = rel[i][1];
    }
}

}  // namespace

RunZustand::RunDynamik::~RunDynamik() {
    if (h) bdk_destroy(h);
    if (mt) bdk_mt_destroy(mt);
}

void RunZustand::RunDynamik::push_positions() {
    std::vector<int32_t> g;
    std::vector<double> r;
    gather_pos(cfg->p.n1, r1_git, r1_rel, g, r);
    check(h, bdk_set_positions(h, 1, cfg->p.n1, g.data(), r.data()), "set_positions(1)");
    gather_pos(cfg->p.n2, r2_git, r2_rel, g, r);
    check(h, bdk_set_positions(h, 2, cfg->p.n2, g.data(), r.data()), "set_positions(2)");
    remember_device_positions();
}

void RunZustand::RunDynamik::remember_device_positions() {
    std::vector<int32_t> g2;
    std::vector<double> r2;
    gather_pos(cfg->p.n1, r1_git, r1_rel, dev_git_, dev_rel_);
    gather_pos(cfg->p.n2, r2_git, r2_rel, g2, r2);
    dev_git_.insert(dev_git_.end(), g2.begin(), g2.end());
    dev_rel_.insert(dev_rel_.end(), r2.begin(), r2.end());
}

void RunZustand::RunDynamik::push_if_changed() {
    if (!h) return;
    const int n1 = cfg->p.n1, n2 = cfg->p.n2;
    if (dev_git_.size() != 2 * (size_t)(n1 + n2)) {
        push_positions();
        return;
    }
    for (int sp = 0; sp < 2; ++sp) {
        const int n = sp ? n2 : n1;
        const size_t off = sp ? 2 * (size_t)n1 : 0;
        int **git = sp ? r2_git : r1_git;
        double **rel = sp ? r2_rel : r1_rel;
        for (int i = 0; i < n; ++i) {
            const size_t k = off + 2 * (size_t)i;
            // bitwise comparison: a NaN written by the driver counts as a change
            if (git[i][0] != dev_git_[k] || git[i][1] != dev_git_[k + 1] ||
                std::memcmp(rel[i], &dev_rel_[k], 2 * sizeof(double)) != 0) {
                push_positions();
                return;
            }
        }
    }
}

void RunZustand::RunDynamik::sync_positions() {
    for (int sp = 1; sp <= 2; ++sp) {
        const int n = sp == 1 ? cfg->p.n1 : cfg->p.n2;
        std::vector<int32_t> g(2 * (size_t)n);
        std::vector<double> r(2 * (size_t)n);
        check(h, bdk_get_positions(h, sp, n, g.data(), r.data()), "get_positions");
        int **git = sp == 1 ? r1_git : r2_git;
        double **rel = sp == 1 ? r1_rel : r2_rel;
        for (int i = 0; i < n; ++i) {
            git[i][0] = g[2 * i];
            git[i][1] = g[2 * i + 1];
            rel[i][0] = r[2 * i];
            rel[i][1] = r[2 * i + 1];
        }
    }
    remember_device_positions();
}

void RunZustand::RunDynamik::init() {
    if (!h) check(nullptr, bdk_create(&cfg->p, &h), "create");
    push_positions();
}

void RunZustand::RunDynamik::berechne_WCA11() {
    push_if_changed();
    F1_WCA.assign(2 * (size_t)cfg->p.n1, 0.0);
    F2_WCA.assign(2 * (size_t)cfg->p.n2, 0.0);
    check(h, bdk_compute_wca(h, F1_WCA.data(), F2_WCA.data()), "compute_wca");
}

void RunZustand::RunDynamik::berechne_kapkraefte() {
    push_if_changed();
    F1kap.assign(2 * (size_t)cfg->p.n1, 0.0);
    F2kap.assign(2 * (size_t)cfg->p.n2, 0.0);
    check(h, bdk_compute_capillary(h, F1kap.data(), F2kap.data()), "compute_capillary");
}

void RunZustand::RunDynamik::berechne_zufallskraefte() {
    if (cfg->p.noise_mode != BDK_NOISE_UPLOAD) return;  // device noise is drawn inside the step
    std::vector<double> a(2 * (size_t)cfg->p.n1), b(2 * (size_t)cfg->p.n2);
    bdk_mt_disc_noise(mt, cfg->p.n1, cfg->p.n2, a.data(), b.data());
    check(h, bdk_set_noise(h, a.data(), b.data()), "set_noise");
}

double RunZustand::RunDynamik::optimaler_zeitschritt() {
    push_if_changed();
    double dt = 0.0;
    check(h, bdk_optimal_dt(h, &dt), "optimal_dt");
    return dt;
}

void RunZustand::RunDynamik::refresh_erwNachbar() {
    push_if_changed();
    check(h, bdk_refresh(h), "refresh");
}

double RunZustand::RunDynamik::zeitschritt(double tmax) {
    if (tmax <= 0.0) {  // dm:80-83
        std::cout << "Fehler: Negative Zeitschrittweite " << tmax << " gefordert, breche ab!" << std::endl;
        return 1.0e10;
    }
    push_if_changed();
    if (cfg->p.noise_mode == BDK_NOISE_UPLOAD && !cfg->p.no_wca) berechne_zufallskraefte();
    // the parameter's travel cap (zeitschritt_debug may have set another)
    check(h, bdk_set_max_travel(h, cfg->p.max_travel), "set_max_travel");
    double dt = 0.0;
    check(h, bdk_step(h, tmax, &dt), "step");
    // the reference's arrays always hold the current positions (dm:116-157
    // updates them in place); a single step keeps the aliases current
    sync_positions();
    return dt;
}

void RunZustand::RunDynamik::zeitschritt_debug(double mrd, TimestepInfo &info, bool print) {
    push_if_changed();
    if (cfg->p.noise_mode == BDK_NOISE_UPLOAD && !cfg->p.no_wca) berechne_zufallskraefte();
    check(h, bdk_set_max_travel(h, mrd), "set_max_travel");
    bdk_step_info_t s;
    check(h, bdk_step_info(h, 1.0, &s), "step_info");
    sync_positions();
    info.num_WCA_berechnet = (int)s.num_wca;
    info.maxWCAx = s.max_wca[0];
    info.maxWCAy = s.max_wca[1];
    info.maxWCAtype = s.max_wca_type;
    info.maxWCAindex = s.max_wca_index;
    info.maxWCA_numWW1 = s.max_wca_nlist[0];
    info.maxWCA_numWW2 = s.max_wca_nlist[1];
    info.maxKapx = s.max_kap[0];
    info.maxKapy = s.max_kap[1];
    info.maxKaptype = s.max_kap_type;
    info.maxKapindex = s.max_kap_index;
    info.maxRNDx = s.max_rnd[0];
    info.maxRNDy = s.max_rnd[1];
    info.maxRNDtype = s.max_rnd_type;
    info.maxRNDindex = s.max_rnd_index;
    info.dt = s.dt;
    info.num_Zellwechsel1 = s.cell_changes[0];
    info.num_Zellwechsel2 = s.cell_changes[1];
    if (print) info.ausgabe();
}

void RunZustand::RunDynamik::TimestepInfo::ausgabe() const {
    std::cout << "WCA pairs evaluated: " << num_WCA_berechnet << "; largest WCA force (" << maxWCAx << ","
              << maxWCAy << ") on particle " << maxWCAindex << " of species " << maxWCAtype << ", neighbour list "
              << maxWCA_numWW1 << "+" << maxWCA_numWW2 << "\n"
              << "cell changes: " << num_Zellwechsel1 << "+" << num_Zellwechsel2 << "\n"
              << "largest capillary force (" << maxKapx << "," << maxKapy << ") on particle " << maxKapindex
              << " of species " << maxKaptype << "\n"
              << "largest random force (" << maxRNDx << "," << maxRNDy << ") on particle " << maxRNDindex
              << " of species " << maxRNDtype << "\n"
              << "dt = " << dt << std::endl;
}

RunZustand::RunZustand() : cfg_(config_from_parameter_h()) {}

RunZustand::~RunZustand() { obs.close(); }

void RunZustand::ausgabe_jeansgroessen() const {
    // mean density, Jeans time 1/(rho v) and the Jeans wave number with the
    // hard-disc correction of am:324-348 (close packing rho_c = 2/sqrt 3)
    const double rho = (cfg_.p.n1 + cfg_.p.n2) / (cfg_.p.L * cfg_.p.L);
    const double rho_c = 2.0 / std::sqrt(3.0);
    const double v = cfg_.p.kap;
    const double tJ = 1.0 / (rho * v);
    const double corr = 1.0 - 3.0 * rho * rho / (rho_c * rho_c) - 8.0 * rho * rho * rho / (rho_c * rho_c * rho_c);
    const double kJ = std::sqrt(v / cfg_.p.T * rho / corr);
    std::cout << "Jeans-Zeit: \t\t" << tJ << std::endl;
    std::cout << "Jeans-Wellenzahl:\t" << kJ << std::endl;
}

void RunZustand::init_zufallspos() {
    bd_params p = cfg_.p;
    std::vector<int32_t> g1(2 * (size_t)p.n1), g2(2 * (size_t)p.n2);
    std::vector<double> r1(2 * (size_t)p.n1), r2(2 * (size_t)p.n2);
    bdk_mt_init_positions(dyn.mt, &p, g1.data(), r1.data(), g2.data(), r2.data());
    for (int i = 0; i < p.n1; ++i) {
        r1_git[i][0] = g1[2 * i];
        r1_git[i][1] = g1[2 * i + 1];
        r1_rel[i][0] = r1[2 * i];
        r1_rel[i][1] = r1[2 * i + 1];
    }
    for (int i = 0; i < p.n2; ++i) {
        r2_git[i][0] = g2[2 * i];
        r2_git[i][1] = g2[2 * i + 1];
        r2_rel[i][0] = r2[2 * i];
        r2_rel[i][1] = r2[2 * i + 1];
    }
}

void RunZustand::init(int nr_) {
    nr = nr_;
    obs_nr = 0;
    const int n1 = cfg_.p.n1, n2 = cfg_.p.n2;
    git_store1_.assign(2 * (size_t)n1, 0);
    git_store2_.assign(2 * (size_t)n2, 0);
    rel_store1_.assign(2 * (size_t)n1, 0.0);
    rel_store2_.assign(2 * (size_t)n2, 0.0);
    git_ptr1_.resize(n1);
    git_ptr2_.resize(n2);
    rel_ptr1_.resize(n1);
    rel_ptr2_.resize(n2);
    for (int i = 0; i < n1; ++i) {
        git_ptr1_[i] = &git_store1_[2 * (size_t)i];
        rel_ptr1_[i] = &rel_store1_[2 * (size_t)i];
    }
    for (int i = 0; i < n2; ++i) {
        git_ptr2_[i] = &git_store2_[2 * (size_t)i];
        rel_ptr2_[i] = &rel_store2_[2 * (size_t)i];
    }
    r1_git = git_ptr1_.data();
    r2_git = git_ptr2_.data();
    r1_rel = rel_ptr1_.data();
    r2_rel = rel_ptr2_.data();
    dyn.cfg = &cfg_;
    if (!dyn.mt) bdk_mt_create(cfg_.seed + (uint64_t)nr, &dyn.mt);
    if (cfg_.startpos_methode != 1) {
        std::cout << "Fehler: startpos_methode=" << cfg_.startpos_methode
                  << ", this library implements 1 (random start, dm:1647-1678)" << std::endl;
        std::exit(1);
    }
    init_zufallspos();
    dyn.r1_git = r1_git;  // am:305-308
    dyn.r2_git = r2_git;
    dyn.r1_rel = r1_rel;
    dyn.r2_rel = r2_rel;
    obs.r1_git = r1_git;
    obs.r2_git = r2_git;
    obs.r1_rel = r1_rel;
    obs.r2_rel = r2_rel;
    dyn.init();
    obs.obs_init(nr, cfg_);
}

void RunZustand::zeitschritte_bis_obs(time_t last_write) {
    (void)last_write;
    dyn.push_if_changed();
    if (cfg_.p.noise_mode != BDK_NOISE_UPLOAD) {
        // the library runs the loop of dm:444-455 on the device (fixed dt: a
        // closed step count; adaptive dt: logged dt's summed in long double)
        // and returns the clock exactly as hi + lo
        int64_t k = 0;
        double hi = 0.0, lo = 0.0;
        check(dyn.h, bdk_run_interval_clock(dyn.h, cfg_.obs_dt, &k, &hi, &lo), "run_interval");
        schritte_seit_obs += (int)k;
        t = (long double)hi + (long double)lo;
        return;
    }
    while (t < cfg_.obs_dt) {
        t += (long double)dyn.zeitschritt((double)((long double)cfg_.obs_dt - t));
        schritte_seit_obs++;
    }
}

void RunZustand::obs_point(int nr_obspoint) {
    dyn.sync_positions();
    if (cfg_.korrfunk) obs.record_korrelationsfunktion(dyn.h, nr_obspoint);
    if (cfg_.rhovonk) obs.record_ftrho(dyn.h, nr_obspoint);
    if (cfg_.rhofftw) obs.record_rhoFFTW(dyn.h, nr_obspoint);
    if (cfg_.auswerten_animation) obs.schreibe_pos(nr_obspoint);
    if (cfg_.abstand) obs.schreibe_abstaende(nr_obspoint);
    schritte_seit_obs = 0;
    t = 0.0L;
    obs_nr++;
}

void RunZustand::RunObs::obs_init(int run, const RunConfig &c) {
    cfg = &c;
    double d[6];
    bdk_params_derived(&c.p, d);
    B = d[2];
    if (c.abstand) {  // au:71-74, a new file every run
        if (ab2) std::fclose(ab2);
        ab2 = std::fopen("abstaende2.txt", "w");
        if (ab2) std::fprintf(ab2, "# Format: t TAB mean distance of the species-2 particles from the box centre \n\n");
    }
    if (!c.auswerten_animation) return;
    path1 = "pos1_" + std::to_string(run + 1) + ".txt";  // au:58-59
    path2 = c.p.n2 ? "pos2_" + std::to_string(run + 1) + ".txt" : "";
    FILE *f = std::fopen(path1.c_str(), "w");
    if (f) {
        std::fprintf(f, "# Format: t TAB x TAB y \n\n");
        std::fclose(f);
    }
    if (!path2.empty()) {
        f = std::fopen(path2.c_str(), "w");
        if (f) {
            std::fprintf(f, "# Format: t TAB x TAB y \n\n");
            std::fclose(f);
        }
    }
}

void RunZustand::RunObs::schreibe_pos(int nr) {
    for (int sp = 1; sp <= 2; ++sp) {
        const int n = sp == 1 ? cfg->p.n1 : cfg->p.n2;
        if (sp == 2 && n == 0) return;
        FILE *f = std::fopen((sp == 1 ? path1 : path2).c_str(), "a");
        if (!f) return;
        int **git = sp == 1 ? r1_git : r2_git;
        double **rel = sp == 1 ? r1_rel : r2_rel;
        for (int i = 0; i < n; ++i) {
            double x = rel[i][0] + B * git[i][0];
            double y = rel[i][1] + B * git[i][1];
            std::fprintf(f, "%g \t %g \t %g \n", nr * cfg->obs_dt, x, y);
        }
        std::fprintf(f, "\n");
        std::fclose(f);
    }
}

void RunZustand::RunObs::record_korrelationsfunktion(bdk_t *h, int nr) {
    const int nb = cfg->korr_bins;
    if ((int)g11.size() < cfg->obs_anzahl) {
        g11.assign(cfg->obs_anzahl, std::vector<double>(nb, 0.0));
        g12.assign(cfg->obs_anzahl, std::vector<double>(nb, 0.0));
        g22.assign(cfg->obs_anzahl, std::vector<double>(nb, 0.0));
    }
    if (nr < 0 || nr >= cfg->obs_anzahl) return;
    check(h, bdk_pair_correlation(h, nb, cfg->korr_dr, g11[nr].data(), g12[nr].data(), g22[nr].data()),
          "pair_correlation");
}

void RunZustand::RunObs::schreibe_korrfunk(int run) const {
    const std::string name = "g_run" + std::to_string(run) + ".txt";
    FILE *f = std::fopen(name.c_str(), "w");
    if (!f) return;
    const bool mixed = cfg->korrfunk_mixed;
    std::fprintf(f, mixed ? "# Format: t TAB r TAB g11 TAB g12 TAB g22\n\n" : "# Format: t TAB r TAB g11 TAB g22\n\n");
    for (int o = 0; o < (int)g11.size(); ++o) {
        const double t = o * cfg->obs_dt;
        for (int b = 0; b < cfg->korr_bins; ++b) {
            const double x = b * cfg->korr_dr;
            if (mixed)
                std::fprintf(f, "%g \t %g \t %g \t %g \t %g \n", t, x, g11[o][b], g12[o][b], g22[o][b]);
            else
                std::fprintf(f, "%g \t %g \t %g \t %g \n", t, x, g11[o][b], g22[o][b]);
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

void RunZustand::RunObs::record_rhoFFTW(bdk_t *h, int nr) {
    const int nb = cfg->rhofftw_bins;
    if ((int)rho1_re.size() < cfg->obs_anzahl) {
        for (auto *v : {&rho1_re, &rho1_im, &rho2_re, &rho2_im}) v->assign(cfg->obs_anzahl, std::vector<double>(nb, 0.0));
    }
    if (nr < 0 || nr >= cfg->obs_anzahl || nb < 1) return;
    check(h, bdk_rho_k(h, 1, cfg->dq_dqft, nb, rho1_re[nr].data(), rho1_im[nr].data()), "rho_k");
    if (cfg->p.n2) check(h, bdk_rho_k(h, 2, cfg->dq_dqft, nb, rho2_re[nr].data(), rho2_im[nr].data()), "rho_k");
}

void RunZustand::RunObs::schreibe_rhoFFT(int run) const {
    const std::string name = "rhoFFT_run" + std::to_string(run) + ".txt";
    FILE *f = std::fopen(name.c_str(), "w");
    if (!f) return;
    const bool one = cfg->p.n2 == 0;
    std::fprintf(f, one ? "# Format: t TAB q TAB rho_re TAB rho_im \n\n"
                        : "# Format: t TAB q TAB rho1_re TAB rho1_im TAB rho2_re TAB rho2_im \n\n");
    for (int o = 0; o < (int)rho1_re.size(); ++o) {
        const double t = o * cfg->obs_dt;
        for (int i = 0; i < cfg->rhofftw_bins; ++i) {
            const double q = cfg->dq_rhofftw * i;
            if (one)
                std::fprintf(f, "%g \t %g \t %g \t %g \n", t, q, rho1_re[o][i], rho1_im[o][i]);
            else
                std::fprintf(f, "%g \t %g \t %g \t %g \t %g \t %g \n", t, q, rho1_re[o][i], rho1_im[o][i],
                             rho2_re[o][i], rho2_im[o][i]);
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

void RunZustand::RunObs::record_ftrho(bdk_t *h, int nr) {
    const int qb = cfg->ftrho_qbins;
    if (ftrho_order.empty()) {  // init_ftrho: shells in order of first appearance
        for (int i = 0; i < qb; ++i)
            for (int j = 0; j < qb; ++j) {
                const int ind2 = i * i + j * j;
                if (ind2 > qb * qb) break;
                int b = -1;
                for (int k = 0; k < (int)ftrho_qabs2.size(); ++k)
                    if (ftrho_qabs2[k] == ind2) {
                        b = k;
                        break;
                    }
                if (b < 0) {
                    ftrho_qabs2.push_back(ind2);
                    b = (int)ftrho_qabs2.size() - 1;
                }
                ftrho_order.push_back(b);
            }
        const std::vector<double> z(ftrho_qabs2.size(), 0.0);
        for (auto *v : {&ft1_re, &ft1_im, &ft2_re, &ft2_im}) v->assign(cfg->obs_anzahl, z);
    }
    if (nr < 0 || nr >= cfg->obs_anzahl) return;
    std::vector<double> re((size_t)qb * qb), im((size_t)qb * qb);
    for (int sp = 1; sp <= (cfg->p.n2 ? 2 : 1); ++sp) {
        int32_t np = 0;
        check(h, bdk_ftrho_sums(h, sp, qb, cfg->ftrho_dq, re.data(), im.data(), &np), "ftrho_sums");
        std::vector<double> &R = sp == 1 ? ft1_re[nr] : ft2_re[nr], &I = sp == 1 ? ft1_im[nr] : ft2_im[nr];
        std::fill(R.begin(), R.end(), 0.0);
        std::fill(I.begin(), I.end(), 0.0);
        for (int k = 0; k < np && k < (int)ftrho_order.size(); ++k) {  // au:341-344
            R[ftrho_order[k]] += re[k];
            I[ftrho_order[k]] += im[k];
        }
    }
}

void RunZustand::RunObs::schreibe_ftrho(int run) const {
    const std::string name = "ftrho_run" + std::to_string(run) + ".txt";
    FILE *f = std::fopen(name.c_str(), "w");
    if (!f) return;
    const bool one = cfg->p.n2 == 0;
    std::fprintf(f, one ? "# Format: t TAB q TAB rho_re TAB rho_im \n\n"
                        : "# Format: t TAB q TAB rho1_re TAB rho1_im TAB rho2_re TAB rho2_im \n\n");
    for (int o = 0; o < (int)ft1_re.size(); ++o) {
        const double t = o * cfg->obs_dt;
        for (int i = 0; i < (int)ft1_re[o].size(); ++i) {
            const double q = cfg->dq * cfg->ftrho_dq * std::sqrt((double)ftrho_qabs2[i]);  // au:983
            if (one)
                std::fprintf(f, "%g \t %g \t %g \t %g \n", t, q, ft1_re[o][i], ft1_im[o][i]);
            else
                std::fprintf(f, "%g \t %g \t %g \t %g \t %g \t %g \n", t, q, ft1_re[o][i], ft1_im[o][i],
                             ft2_re[o][i], ft2_im[o][i]);
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

void RunZustand::schreibe_obs() {
    if (cfg_.korrfunk) obs.schreibe_korrfunk(nr);
    if (cfg_.rhovonk) obs.schreibe_ftrho(nr);
    if (cfg_.rhofftw) obs.schreibe_rhoFFT(nr);
}

void RunZustand::RunObs::schreibe_abstaende(int nr) {
    if (!ab2) return;
    const double mitte = 0.5 * cfg->p.L;
    double mean = 0.0;
    for (int i = 0; i < cfg->p.n2; ++i) {
        const double dx = r2_rel[i][0] + B * r2_git[i][0] - mitte;
        const double dy = r2_rel[i][1] + B * r2_git[i][1] - mitte;
        mean += std::sqrt(dx * dx + dy * dy);
    }
    mean /= cfg->p.n2;
    std::fprintf(ab2, "%g \t %g \n", nr * cfg->obs_dt, mean);
    std::fflush(ab2);
}

void RunZustand::RunObs::close() {
    if (ab2) std::fclose(ab2);
    ab2 = nullptr;
}
