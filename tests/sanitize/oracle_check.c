/* Sanitizer run of the FP64 oracle (test infrastructure): the reference's
 * random start (init_zufallspos, its overlapping pairs and rel up to 2B),
 * steps of every scheme and noise mode, adaptive dt, the parameter.h
 * switches, the checkpoint loop and the pos writer, so that AddressSanitizer
 * and UndefinedBehaviorSanitizer see every path of bd_oracle.c. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bd_oracle.h"

static bdo_params base(int scheme, int noise) {
    bdo_params p;
    memset(&p, 0, sizeof p);
    p.n1 = 120;
    p.n2 = 80;
    p.scheme = scheme;
    p.Z = 32;
    p.L = 16.0;
    p.lambda = 4.0;
    p.kap = 1.0;
    p.T = 1.0;
    p.dt_max = 1e-4;
    p.mrd = INFINITY;
    p.s11_22 = p.s11_12 = p.e22_11 = p.e12_11 = 1.0;
    p.f2_f1 = 0.5;
    p.max_pair_wca = 1e5;
    p.noise_mode = noise;
    p.noise_var = 1.0;
    p.seed = 99;
    return p;
}

int main(void) {
    for (int scheme = 0; scheme < 3; ++scheme)
        for (int noise = 0; noise < 3; ++noise) {
            bdo_params p = base(scheme, noise);
            if (noise == 2) p.mrd = 0.05;           /* adaptive dt */
            if (scheme == 1) p.s11_22 = 1.2, p.s11_12 = 1.1, p.e22_11 = 0.8, p.e12_11 = 0.9;
            if (scheme == 2 && noise == 1) p.restrict_radial = 1;
            bdo_state *s = bdo_create(&p);
            bdo_init_zufallspos(s);
            for (int k = 0; k < 20; ++k) bdo_zeitschritt(s, 1.0);
            if (bdo_zeitschritte_bis_obs(s, 3e-4) < 1) return 1;
            bdo_write_pos_header("/tmp/bdk_asan_pos.txt");
            bdo_append_pos(s, 1, "/tmp/bdk_asan_pos.txt", 1, 0.1);
            bdo_append_pos(s, 2, "/tmp/bdk_asan_pos.txt", 1, 0.1);
            double *f = malloc(sizeof(double) * 2 * (p.n1 + p.n2));
            for (int which = 0; which < 3; ++which)
                for (int sp = 1; sp <= 2; ++sp) bdo_get_forces(s, which, sp, f);
            double *m = malloc(sizeof(double) * p.Z * p.Z);
            for (int which = 0; which < 3; ++which) bdo_get_mesh(s, which, m);
            int32_t lst[512];
            bdo_get_list(s, 1, 0, 0, lst, 512);
            bdo_get_list(s, 2, 3, 4, lst, 512);
            free(f);
            free(m);
            bdo_destroy(s);
        }
    /* switches: noWCA, idgas_only, uploaded noise */
    for (int k = 0; k < 3; ++k) {
        bdo_params p = base(2, k == 2 ? 3 : 1);
        if (k == 0) p.no_wca = 1;
        if (k == 1) p.idgas_only = 1;
        bdo_state *s = bdo_create(&p);
        bdo_init_zufallspos(s);
        double *n1 = calloc(2 * p.n1, sizeof(double)), *n2 = calloc(2 * p.n2, sizeof(double));
        for (int j = 0; j < 5; ++j) {
            if (k == 2) bdo_set_noise(s, n1, n2);
            bdo_zeitschritt(s, 1.0);
        }
        free(n1);
        free(n2);
        bdo_destroy(s);
    }
    printf("oracle_check: OK\n");
    return 0;
}
