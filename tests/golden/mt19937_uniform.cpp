// Golden vectors for the Monte-Carlo time draws of MultiQW.cpp:387,417-421:
// std::mt19937 seeded with 29552825458725 and std::uniform_real_distribution<float>(s, 2s),
// compiled as the reference's README builds it (g++ -O3 -march=native).
// Output: binary float32 draws for each s in argv, 100 per s in one stream,
// preceded by the first 1000 raw mt19937 outputs (uint32) of the same seed.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
int main(int argc, char** argv)
{
    std::mt19937 raw(29552825458725);
    std::vector<unsigned> r(1000);
    for (auto& x : r) x = raw();
    fwrite(r.data(), 4, r.size(), stdout);
    std::mt19937 gen(29552825458725);
    for (int a = 1; a < argc; a++) {
        float s = strtof(argv[a], nullptr);
        std::uniform_real_distribution<float> d(s, 2 * s);
        for (int j = 0; j < 100; j++) {
            float v = d(gen);
            fwrite(&v, 4, 1, stdout);
        }
    }
    return 0;
}
