// C entry point over mqw::compress_first_stage (multistageqw_amd/csrc/mqw_basis.cpp)
// for the CPU tests: built with g++ by tests/test_basis_host.py, no device code.
#include <string.h>

#include <vector>

#include "mqw_internal.h"

extern "C" int harness_compress(int kc, const double* gram, int S, const double* coefs, double tol, int* R,
                                double* B, double* u)
{
    std::vector<const double*> cp((size_t)S);
    for (int j = 0; j < S; j++) cp[(size_t)j] = coefs + (size_t)j * kc;
    mqw::FirstStageCompression fc;
    if (!mqw::compress_first_stage(kc, gram, cp, tol, fc)) return 0;
    *R = fc.R;
    memcpy(B, fc.B.data(), sizeof(double) * fc.B.size());
    memcpy(u, fc.u.data(), sizeof(double) * fc.u.size());
    return 1;
}
