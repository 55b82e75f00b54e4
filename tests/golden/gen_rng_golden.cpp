// Writes libstdc++ <random> known-answer vectors for std::default_random_engine
// (minstd_rand0): raw draws, generate_canonical<double,53> and
// bernoulli_distribution decisions.  These come from the C++ standard library,
// which is what the reference links (src/random.cpp:14), not from the reference.
// Build: g++ -O2 gen_rng_golden.cpp        (no FMA contraction possible)
//        g++ -O2 -mfma -ffp-contract=fast  (canonical sum fused, as -march=native may do)
#include <cstdio>
#include <random>
int main()
{
    const unsigned long seeds[] = {0UL, 1UL, 12345UL, 20240520UL, 2147483647UL, 2147483648UL, 4294967295UL};
    std::default_random_engine kat;
    kat.discard(9999);
    std::printf("{\n  \"kat_10000th\": %lu,\n  \"seeds\": [\n", (unsigned long)kat());
    const double ps[] = {0.0, 0.3, 0.5, 0.999, 1.0};
    for (unsigned si = 0; si < sizeof(seeds) / sizeof(seeds[0]); ++si) {
        std::default_random_engine e; e.seed(seeds[si]);
        std::printf("    {\"seed\": %lu, \"raw\": [", seeds[si]);
        for (int i = 0; i < 8; ++i) std::printf("%s%lu", i ? ", " : "", (unsigned long)e());
        e.seed(seeds[si]);
        std::printf("], \"canonical_hex\": [");
        for (int i = 0; i < 64; ++i) std::printf("%s\"%a\"", i ? ", " : "", std::generate_canonical<double, 53>(e));
        e.seed(seeds[si]);
        std::printf("], \"bernoulli\": [");
        int first = 1;
        for (int i = 0; i < 40; ++i) {
            const double p = ps[i % 5];
            std::printf("%s[%.17g, %d]", first ? "" : ", ", p, std::bernoulli_distribution(p)(e) ? 1 : 0);
            first = 0;
        }
        std::printf("], \"state_after_bernoulli\": %lu}%s\n", (unsigned long)e(), si + 1 < sizeof(seeds) / sizeof(seeds[0]) ? "," : "");
    }
    std::printf("  ]\n}\n");
    return 0;
}
