// Round trip through the drop-in C++ interface, written the way the
// reference's examples/example.cpp:13-85 uses PFSOFTlib: IDSOFT of random
// coefficients into a fresh grid, DSOFT back, and its pass criterion -- every
// |re| and |im| difference <= 1e-11.  Inputs come from a seeded generator
// instead of the reference's time-seeded rand() (random.hpp:89-96).
#include "pfsoft_amd.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

using namespace pfsoft;

static double next_uniform(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 2.0 * ((double)(z >> 11) / 9007199254740992.0) - 1.0;
}

int main(int argc, char** argv)
{
    int bandwidth = argc > 1 ? atoi(argv[1]) : 16;
    grid3D<complex<double> > sample(2 * bandwidth);
    DSOFTFourierCoefficients coef(bandwidth);
    DSOFTFourierCoefficients rec_coef(bandwidth);

    uint64_t seed = 42;
    for (int l = 0; l < bandwidth; ++l)
        for (int M = -l; M <= l; ++M)
            for (int Mp = -l; Mp <= l; ++Mp) {
                coef(l, M, Mp).re = next_uniform(&seed);
                coef(l, M, Mp).im = next_uniform(&seed);
            }

    FourierTransforms::IDSOFT(coef, sample);
    FourierTransforms::DSOFT(sample, rec_coef);

    bool equal = true;
    double epsilon = 1e-11, maxerr = 0;
    int cnt_fc = 0;
    for (int m = 0; m < bandwidth; ++m)
        for (int n = -m; n <= m; ++n)
            for (int k = -m; k <= m; ++k) {
                double dr = fabs(coef(m, n, k).re - rec_coef(m, n, k).re);
                double di = fabs(coef(m, n, k).im - rec_coef(m, n, k).im);
                if (dr > epsilon || di > epsilon) equal = false;
                maxerr = fmax(maxerr, fmax(dr, di));
                cnt_fc++;
            }
    printf("#coefficients:   %d\n", cnt_fc);
    printf("Bandbreite:      %d\n", bandwidth);
    printf("max abs error:   %.3e\n", maxerr);
    printf("Correct result:  %s\n", (equal ? "Yes" : "No"));

    // the reference's warn-and-return path on a shape mismatch leaves fc alone
    grid3D<complex<double> > odd(2 * bandwidth + 1);
    FourierTransforms::DSOFT(odd, rec_coef);
    return equal ? 0 : 1;
}
