// Writes tests/golden/lindblad_rng.json: draws of the C++ standard library that the trajectory
// solver's random streams are built from, for the oracle's restatement to be checked against.
//   g++ -O2 -std=c++17 tests/golden/make_lindblad_rng.cpp -o /tmp/mk && /tmp/mk > tests/golden/lindblad_rng.json
// Streams: std::mt19937_64 seeded with splitmix64(seed ^ splitmix64(i + 1)) (the per-trajectory
// seeding rule of the reference's solver); 40 uniform_real_distribution<double>(0, 1) draws and
// 40 normal_distribution<double>(0, 0.1) draws (fresh generators) for seeds 1 and 2025,
// trajectories 0..3; and splitmix64 of a few inputs.
#include <cstdint>
#include <cstdio>
#include <random>

static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

int main() {
    std::printf("{\n  \"splitmix64\": [");
    const uint64_t xs[] = {0, 1, 1234567, 0xFFFFFFFFFFFFFFFFULL};
    for (int i = 0; i < 4; ++i) std::printf("%s[\"%llu\", \"%llu\"]", i ? ", " : "", (unsigned long long)xs[i], (unsigned long long)splitmix64(xs[i]));
    std::printf("],\n  \"streams\": [\n");
    const uint64_t seeds[] = {1, 2025};
    bool first = true;
    for (uint64_t seed : seeds)
        for (uint64_t i = 0; i < 4; ++i) {
            std::mt19937_64 a(splitmix64(seed ^ splitmix64(i + 1))), b(splitmix64(seed ^ splitmix64(i + 1)));
            std::uniform_real_distribution<double> uni(0.0, 1.0);
            std::normal_distribution<double> gauss(0.0, 0.1);
            std::printf("%s    {\"seed\": %llu, \"index\": %llu, \"uniform\": [", first ? "" : ",\n", (unsigned long long)seed, (unsigned long long)i);
            first = false;
            for (int k = 0; k < 40; ++k) std::printf("%s%.17g", k ? ", " : "", uni(a));
            std::printf("], \"normal\": [");
            for (int k = 0; k < 40; ++k) std::printf("%s%.17g", k ? ", " : "", gauss(b));
            std::printf("]}");
        }
    std::printf("\n  ]\n}\n");
    return 0;
}
