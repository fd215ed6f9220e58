// lobes.cpp -- Gaussian derivative lobes on the host.
//
// Compiled without floating-point contraction: the reference evaluates these
// float expressions on the host, one rounding per operation.
#include <cmath>
#include <cstddef>


namespace tucu {

constexpr size_t LobeSizeMax = 17;      // FIRFilter.h:26

// detail::lobeSize (FIRGaussianConvolver.cu:52-98): truncate where |lobe|
// falls to 1 % of its peak, then round up to a supported size class.
size_t lobe_size(const float lobe[], bool even)
{
    const float epsilon = 0.01f;
    float th = 0;
    for (size_t i = LobeSizeMax; i-- > 0; )
        if (std::fabs(lobe[i]) >= th) th = std::fabs(lobe[i]);
    th *= epsilon;

    size_t siz;
    for (siz = LobeSizeMax; siz-- > 0; )
        if (std::fabs(lobe[siz]) > th) { ++siz; break; }

    if (even) return siz <= 2 ? 3 : siz <= 5 ? 5 : siz <= 9 ? 9 : 17;
    return siz <= 1 ? 2 : siz <= 4 ? 4 : siz <= 8 ? 8 : 16;
}

// Samples of the Gaussian and its first two derivatives at integer offsets
// (FIRGaussianConvolver.h:130-139).
static void gauss_samples(float sigma, float lb0[], float lb1[], float lb2[])
{
    for (size_t i = 0; i < LobeSizeMax; ++i) {
        const float dx = float(i) / sigma, dxdx = dx * dx;
        lb0[i] = std::exp(-0.5f * dxdx);
        lb1[i] = -dx * lb0[i];
        lb2[i] = (dxdx - 1.0f) * lb0[i];
    }
}

// Normalisation of FIRGaussianConvolver.h:141-163 for lobe lengths s0, s1, s2:
//   lobe0[i] = lb0[s0-1-i] / |lb0[0] + 2 sum_{1<=i<s0} lb0[i]|
//   lobe1[i] = lb1[s1-i]   / |sum_{0<=i<s1} 2 i lb1[i]|
//   lobe2[i] = lb2[s2-1-i] / |sum_{1<=i<s2} i^2 lb2[i]|
static void normalise(const float lb0[], const float lb1[], const float lb2[],
                      size_t s0, size_t s1, size_t s2,
                      float* lobe0, float* lobe1, float* lobe2)
{
    float sum = lb0[0];
    for (size_t i = 1; i < s0; ++i) sum += (2.0f * lb0[i]);
    for (size_t i = 0; i < s0; ++i) lobe0[i] = lb0[s0 - 1 - i] / std::fabs(sum);

    sum = 0.0f;
    for (size_t i = 0; i < s1; ++i) sum += (2.0f * float(i) * lb1[i]);
    for (size_t i = 0; i < s1; ++i) lobe1[i] = lb1[s1 - i] / std::fabs(sum);

    sum = 0.0f;
    for (size_t i = 1; i < s2; ++i) sum += (float(i * i) * lb2[i]);
    for (size_t i = 0; i < s2; ++i) lobe2[i] = lb2[s2 - 1 - i] / std::fabs(sum);
}

// FIRGaussianConvolver2::initialize(sigma, lobe0, lobe1, lobe2)
// (FIRGaussianConvolver.h:123-170).
void gauss_lobes_reference(float sigma, float* lobe0, int* n0, float* lobe1, int* n1,
                           float* lobe2, int* n2)
{
    float lb0[LobeSizeMax], lb1[LobeSizeMax], lb2[LobeSizeMax];
    gauss_samples(sigma, lb0, lb1, lb2);
    const size_t s0 = lobe_size(lb0, true);
    const size_t s1 = lobe_size(lb1, false);
    const size_t s2 = lobe_size(lb2, true);
    normalise(lb0, lb1, lb2, s0, s1, s2, lobe0, lobe1, lobe2);
    *n0 = int(s0); *n1 = int(s1); *n2 = int(s2);
}

// Extension rule: radius r = ceil(4 sigma), clamped to [1, 16]; symmetric
// lobes of r+1 entries, the antisymmetric one of r, normalised by the same
// formulas as the reference rule.
void gauss_lobes_ceil4(float sigma, float* lobe0, int* n0, float* lobe1, int* n1,
                       float* lobe2, int* n2)
{
    float lb0[LobeSizeMax], lb1[LobeSizeMax], lb2[LobeSizeMax];
    gauss_samples(sigma, lb0, lb1, lb2);
    int r = int(std::ceil(4.0f * sigma));
    r = r < 1 ? 1 : r > 16 ? 16 : r;
    normalise(lb0, lb1, lb2, size_t(r) + 1, size_t(r), size_t(r) + 1, lobe0, lobe1, lobe2);
    *n0 = r + 1; *n1 = r; *n2 = r + 1;
}

}   // namespace tucu
