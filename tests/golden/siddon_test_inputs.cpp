// Writes the 10000 segment pairs that the reference's test/SiddonTest.cpp
// draws (std::default_random_engine seeded with 0, uniform_real_distribution
// over [-500, 500), p1.x p1.y p1.z p2.x p2.y p2.z per segment; SiddonTest.cpp:8-10,27-32)
// to stdout as raw little-endian doubles, 6 per segment.
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char** argv) {
    int n = argc > 1 ? std::atoi(argv[1]) : 10000;
    std::default_random_engine gen{0};
    std::uniform_real_distribution<double> dist(-500.0, 500.0);
    for (int i = 0; i < n; ++i) {
        double v[6];
        for (int k = 0; k < 6; ++k) v[k] = dist(gen);
        std::fwrite(v, sizeof(double), 6, stdout);
    }
    return 0;
}
