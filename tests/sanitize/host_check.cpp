// Sanitizer run of the product library's host side (test infrastructure):
// every C-ABI path that needs no GPU -- parameter defaults, derived
// constants and validation, footprints, the checkpoint step count, the MT
// stream and the reference's random start, disc noise, the initial-state
// generators, the startpos reader, and the job server's admission failure
// (a device ordinal that does not exist), whose ownership of the failed
// jobs AddressSanitizer checks on destroy.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bdk.h"

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::fprintf(stderr, "host_check: failed %s (line %d)\n", #c, __LINE__); \
            return 1;                                                     \
        }                                                                 \
    } while (0)

int main() {
    bd_params p;
    CHECK(bdk_params_default(&p) == BDK_OK);
    p.n1 = 300;
    p.n2 = 200;
    p.L = 30.0;
    p.Z = 32;
    p.lambda = 5.0;
    double d[6];
    CHECK(bdk_params_derived(&p, d) == BDK_OK && d[1] == 26.0);
    uint64_t fp = 0;
    CHECK(bdk_footprint(&p, &fp) == BDK_OK && fp > 0);
    int64_t k = 0;
    CHECK(bdk_steps_for_interval(0.1, 1e-4, &k) == BDK_OK && k == 1001);
    // validation failures (before any device call)
    bd_params bad = p;
    bad.Z = 44;  // a factor 11
    bdk_t *h = nullptr;
    CHECK(bdk_create(&bad, &h) == BDK_ERR_ARG && h == nullptr && std::strlen(bdk_last_error(nullptr)) > 0);
    bad = p;
    bad.L = 2.0;
    CHECK(bdk_create(&bad, &h) == BDK_ERR_ARG);
    // MT stream, reference start, disc noise
    bdk_mt_t *mt = nullptr;
    CHECK(bdk_mt_create(5489, &mt) == BDK_OK);
    for (int i = 0; i < 1000; ++i) {
        const double u = bdk_mt_uniform(mt, -1.0, 1.0);
        CHECK(u >= -1.0 && u < 1.0);
    }
    std::vector<int32_t> g1(2 * 300), g2(2 * 200), g(2 * 500);
    std::vector<double> r1(2 * 300), r2(2 * 200), r(2 * 500);
    CHECK(bdk_mt_init_positions(mt, &p, g1.data(), r1.data(), g2.data(), r2.data()) == BDK_OK);
    CHECK(bdk_mt_disc_noise(mt, 300, 200, r1.data(), r2.data()) == BDK_OK);
    bdk_mt_destroy(mt);
    // initial states
    CHECK(bdk_init_uniform(&p, 7, 0.9, g.data(), r.data()) == BDK_OK);
    double clipped = -1.0;
    CHECK(bdk_init_blob(&p, 7, 5.0, 0.1, 0.95, g.data(), r.data(), &clipped) == BDK_OK && clipped >= 0.0);
    // startpos reader: the reference's format
    const char *path = "/tmp/bdk_asan_startpos.txt";
    FILE *f = std::fopen(path, "w");
    CHECK(f != nullptr);
    std::fprintf(f, "# header\n\n");
    for (int i = 0; i < 500; ++i)
        std::fprintf(f, "%d %d %g %g\n", i < 300 ? i : i - 300, i < 300 ? 1 : 2, 0.05 * i, 29.0 - 0.05 * i);
    std::fclose(f);
    CHECK(bdk_read_startpos(&p, path, g.data(), r.data()) == BDK_OK);
    CHECK(bdk_read_startpos(&p, "/nonexistent/startpos.txt", g.data(), r.data()) != BDK_OK);
    // job server on a device that does not exist: every admission fails
    bdk_server_t *srv = nullptr;
    CHECK(bdk_serve_create(99, 1ull << 32, 2, 10, &srv) == BDK_OK);
    int64_t ids[4];
    CHECK(bdk_init_uniform(&p, 9, 0.9, g.data(), r.data()) == BDK_OK);
    for (int j = 0; j < 4; ++j) CHECK(bdk_serve_submit(srv, &p, g.data(), r.data(), 20 + j, &ids[j]) == BDK_OK);
    int32_t a = -1, q = -1, dn = -1;
    CHECK(bdk_serve_step(srv, &a, &q, &dn) == BDK_OK && a == 0 && q == 0 && dn == 4);
    for (int j = 0; j < 4; ++j) {
        double ta = -1.0, td = -1.0;
        CHECK(bdk_serve_result(srv, ids[j], g.data(), r.data(), &ta, &td) != BDK_OK);
        CHECK(ta >= 0.0 && td >= ta);
    }
    int64_t ps = -1;
    double busy = -1.0;
    uint64_t used = 1;
    CHECK(bdk_serve_stats(srv, &ps, &busy, &used) == BDK_OK && ps == 0 && used == 0);
    bdk_serve_destroy(srv);
    bdk_batch_t *b = nullptr;
    CHECK(bdk_batch_create(nullptr, 0, &b) == BDK_ERR_ARG);
    std::printf("host_check: OK\n");
    return 0;
}
