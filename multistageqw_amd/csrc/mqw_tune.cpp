// Autotuner of the sweep kernels.  For one n it times every candidate
// configuration (kernel kind and tile bits, loads-in-flight variant, XCD chunk
// bits) on synthetic term jobs shaped like a batched solve, keeps the fastest
// in the context and, given a cache path, records it as a line
//   <gfx arch> <n> <tile bits> <variant> <chunk bits> <GB/s>
// so later processes reuse the choice without measuring.  MQW_TUNE_VERBOSE=1
// prints every candidate's rate to stderr.
#include <hip/hip_runtime_api.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "mqw_internal.h"

using namespace mqw;

namespace {

struct Cand {
    int tb, variant, lc;
};

std::vector<Cand> candidates(int n)
{
    std::vector<Cand> v;
    for (int tb : {11, 12}) {
        if (tb > n) continue;
        for (int var = 0; var < term_variants(); var++)
            for (int lc : {0, 3, 6, 8})
                if (lc <= n - tb) v.push_back({tb, var, lc});
    }
    for (int hb : {13, 14}) {
        if (hb > n) continue;
        for (int var = 0; var < half_variants(hb); var++)
            for (int lc : {3, 4, 5, 6, 7})
                if (lc <= n - hb + 1) v.push_back({hb, var, lc});
    }
    if (v.empty()) v.push_back({std::min(n, 12), 0, 0});
    return v;
}

bool cache_lookup(const char* path, const std::string& arch, int n, Cand& out)
{
    if (!path) return false;
    FILE* f = fopen(path, "r");
    if (!f) return false;
    char a[128];
    int nn, tb, var, lc;
    double gbs;
    bool found = false;
    while (fscanf(f, "%127s %d %d %d %d %lf", a, &nn, &tb, &var, &lc, &gbs) == 6)
        if (arch == a && nn == n) {
            out = {tb, var, lc};
            found = true;
        }
    fclose(f);
    return found;
}

std::string device_arch(int device)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
    std::string a = p.gcnArchName;
    const size_t colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
}

} // namespace

extern "C" mqw_status mqw_tune(mqw_ctx* c, int n, const char* cache_path, int* tile_bits, int* chunk_bits,
                               int* variant)
{
    if (!c) return MQW_EINVAL;
    if (n < MQW_MIN_N || n > MQW_MAX_N) return fail(c, MQW_EINVAL, "n=%d outside [%d, %d]", n, MQW_MIN_N, MQW_MAX_N);
    hipSetDevice(c->device);
    const std::string arch = device_arch(c->device);
    Cand best{0, 0, 0};
    const char* retune = getenv("MQW_RETUNE");
    bool have = !(retune && *retune == '1') && cache_lookup(cache_path, arch, n, best);
    double best_gbs = 0.0;
    if (!have) {
        const size_t N = (size_t)1 << n;
        // about 2^27 amplitudes per launch, like a batch of trajectories
        const size_t T = std::max<size_t>(1, ((size_t)1 << 27) >> n);
        float* st = (float*)dev_get(c, "tune_state", sizeof(float) * 6 * N * T);
        // a diagonal per instance, dozens of trajectories (samples) per instance
        const size_t NE = std::max<size_t>(1, std::min<size_t>(T / 32, 64));
        float* E = (float*)dev_get(c, "tune_E", sizeof(float) * N * NE);
        double* part = (double*)dev_get(c, "tune_part", sizeof(double) * 2 * N);
        TermJob<float>* jd = (TermJob<float>*)dev_get(c, "tune_jobs", sizeof(TermJob<float>) * 2 * T);
        if (!st || !E || !part || !jd) return fail(c, MQW_ENOMEM, "device allocation failed while tuning n=%d", n);
        HIPCHK(c, hipMemsetAsync(st, 0, sizeof(float) * 6 * N * T, c->stream));
        HIPCHK(c, hipMemsetAsync(E, 0, sizeof(float) * N * NE, c->stream));
        // two job tables: b1 = A, b2 = out = B, psi = C, then A and B swapped
        std::vector<TermJob<float>> hj(2 * T);
        for (int ph = 0; ph < 2; ph++)
            for (size_t t = 0; t < T; t++) {
                float* A = st + (t * 3 + 0) * 2 * N;
                float* B = st + (t * 3 + 1) * 2 * N;
                TermJob<float>& j = hj[ph * T + t];
                j = TermJob<float>{};
                j.b1 = ph ? B : A;
                j.b2 = j.out = ph ? A : B;
                j.psi = st + (t * 3 + 2) * 2 * N;
                j.E = E + (t % NE) * N;
                j.norm_part = part;
                j.alpha = 0.1f;
                j.beta = 0.01f;
                j.w1r = 1.0f;
                j.w2r = 1.0f;
                j.cr = 0.5f;
                j.flags = SRC1_ARRAY | SRC2_ARRAY;
            }
        HIPCHK(c, hipMemcpyAsync(jd, hj.data(), sizeof(TermJob<float>) * hj.size(), hipMemcpyHostToDevice, c->stream));
        hipEvent_t e0, e1;
        HIPCHK(c, hipEventCreate(&e0));
        HIPCHK(c, hipEventCreate(&e1));
        const int reps = 8;
        const char* vb = getenv("MQW_TUNE_VERBOSE");
        const bool verbose = vb && *vb == '1';
        for (const Cand& cd : candidates(n)) {
            const bool half = cd.tb >= 13;
            auto launch = [&](int q) {
                const void* jobs = jd + (size_t)(q & 1) * T;
                return half ? launch_half_term(cd.tb, cd.variant, jobs, (uint32_t)T, n, cd.lc, 0, c->stream)
                            : launch_term(MQW_FP32, cd.tb, cd.variant, jobs, (uint32_t)T, n, cd.lc, c->stream);
            };
            if (launch(0) || launch(1)) continue;
            HIPCHK(c, hipEventRecord(e0, c->stream));
            int bad = 0;
            for (int q = 0; q < reps; q++) bad |= launch(q);
            HIPCHK(c, hipEventRecord(e1, c->stream));
            HIPCHK(c, hipEventSynchronize(e1));
            if (bad) continue;
            float ms = 0.0f;
            hipEventElapsedTime(&ms, e0, e1);
            const double gbs = 36.0 * (double)N * (double)T * reps / (ms * 1e-3) / 1e9;
            if (verbose) fprintf(stderr, "mqw_tune n=%d tb=%d variant=%d lc=%d %.1f GB/s\n", n, cd.tb, cd.variant, cd.lc, gbs);
            if (gbs > best_gbs) {
                best_gbs = gbs;
                best = cd;
            }
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        if (best.tb == 0) return fail(c, MQW_EDEVICE, "no sweep configuration ran for n=%d", n);
        if (cache_path) {
            FILE* f = fopen(cache_path, "a");
            if (f) {
                fprintf(f, "%s %d %d %d %d %.1f\n", arch.c_str(), n, best.tb, best.variant, best.lc, best_gbs);
                fclose(f);
            }
        }
    }
    c->tuned_tb[n] = best.tb;
    c->tuned_var[n] = best.variant;
    c->tuned_lc[n] = best.lc;
    if (tile_bits) *tile_bits = best.tb;
    if (chunk_bits) *chunk_bits = best.lc;
    if (variant) *variant = best.variant;
    return MQW_OK;
}
