// Emits known-answer data of the C++ standard library's mt19937_64 and
// uniform_real_distribution<double>(0,1), drawn through a std::bind copy of
// the engine, which is how zufall.cpp:18-24 draws its uniforms.  Used to pin
// the oracle's MT restatement.  Output: JSON on stdout.
#include <cstdio>
#include <functional>
#include <random>

int main() {
    std::mt19937_64 def;  // default seed 5489
    unsigned long long x = 0;
    for (int i = 0; i < 10000; ++i) x = def();
    std::printf("{\n  \"default_seed_10000th\": \"%llu\",\n", x);

    const unsigned long long seeds[3] = {12345ULL, 1ULL, 18446744073709551557ULL};
    std::printf("  \"streams\": [\n");
    for (int s = 0; s < 3; ++s) {
        std::mt19937_64 g(seeds[s]);
        std::printf("    {\"seed\": \"%llu\", \"raw\": [", seeds[s]);
        for (int i = 0; i < 700; ++i) std::printf("%s\"%llu\"", i ? ", " : "", (unsigned long long)g());
        std::printf("],\n");
        std::mt19937_64 g2(seeds[s]);
        std::uniform_real_distribution<double> dist(0.0, 1.0);
        auto u = std::bind(dist, g2);
        std::printf("     \"canonical\": [");
        for (int i = 0; i < 2000; ++i) std::printf("%s\"%a\"", i ? ", " : "", u());
        std::printf("]}%s\n", s < 2 ? "," : "");
    }
    std::printf("  ]\n}\n");
    return 0;
}
