// Prints libstdc++'s mt19937_64, uniform_int_distribution<int>(1, INT_MAX)
// and normal_distribution<float/double> draws for a seed, one per line.
// Used by tests/test_oracle_pins.py to pin the oracle's restatement.
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char **argv)
{
    unsigned long long seed = std::strtoull(argv[1], nullptr, 10);
    int n = std::atoi(argv[2]);
    {
        std::mt19937_64 g(seed);
        for (int i = 0; i < n; ++i) std::printf("%llu\n", (unsigned long long)g());
    }
    {
        std::mt19937_64 g(seed);
        std::uniform_int_distribution<int> u(1, INT_MAX);
        for (int i = 0; i < n; ++i) std::printf("%d\n", u(g));
    }
    {
        std::mt19937_64 g(seed);
        std::normal_distribution<float> d(0.f, 1.f);
        for (int i = 0; i < n; ++i) std::printf("%.9e\n", (double)d(g));
    }
    {
        std::mt19937_64 g(seed);
        std::normal_distribution<double> d(0.0, 1.0);
        for (int i = 0; i < n; ++i) std::printf("%.17e\n", d(g));
    }
    return 0;
}
