// Autotuner of the sweep kernels, in two stages for one n:
//   1. every candidate configuration (kernel kind and tile bits, variant,
//      XCD chunk bits) times a few launches on synthetic term jobs shaped
//      like a batched solve;
//   2. the fastest few of those, the measured default configuration and,
//      from n = 22, the high-bit pass (LDS and register pencils), each run a real solve of
//      a generated SK instance (m = 3, enough samples for about 2^26
//      amplitudes in flight, every stage as term sweeps); the fastest by the
//      sweeps' device time wins.
// The choice is kept in the context and, given a cache path, recorded as a line
//   <gfx arch> <n> <tile bits> <variant> <chunk bits> <high pass> <GB/s>
// so later processes reuse it without measuring.  MQW_RETUNE=1 measures again;
// MQW_TUNE_VERBOSE=1 prints every candidate's rate to stderr.
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
    int hp; // high_pass option: -1 by n, 0 off, 1..3 on
};

// From n = 14 only the real-half kernel: the shared bases (first and second
// stage, adjoint readouts, direct m = 1 chains) run on it, so a complex-tile
// choice would turn them off for every m > 1 instance of that n.
std::vector<Cand> candidates(int n)
{
    std::vector<Cand> v;
    for (int tb : {11, 12}) {
        if (tb > n || n >= 14) continue;
        for (int var = 0; var < term_variants(); var++)
            for (int lc : {0, 3, 6, 8})
                if (lc <= n - tb) v.push_back({tb, var, lc, 0});
    }
    for (int hb : {13, 14}) {
        if (hb > n) continue;
        for (int var = 0; var < half_variants(hb); var++)
            for (int lc : {3, 4, 5, 6, 7})
                if (lc <= n - hb + 1) v.push_back({hb, var, lc, 0});
    }
    if (v.empty()) v.push_back({std::min(n, 12), 0, 0, 0});
    return v;
}

bool cache_lookup(const char* path, const std::string& arch, int n, Cand& out)
{
    if (!path) return false;
    FILE* f = fopen(path, "r");
    if (!f) return false;
    char line[256];
    bool found = false;
    while (fgets(line, sizeof line, f)) {
        char a[128];
        int nn, tb, var, lc, hp;
        double gbs;
        if (sscanf(line, "%127s %d %d %d %d %d %lf", a, &nn, &tb, &var, &lc, &hp, &gbs) == 7 && arch == a &&
            nn == n) {
            out = {tb, var, lc, hp};
            found = true;
        }
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

void forget(mqw_ctx* c, int n)
{
    c->tuned_tb[n] = 0;
    c->tuned_lc[n] = -1;
    c->tuned_var[n] = 0;
    c->tuned_hp[n] = -1;
}

// Stage 1: GB/s of every candidate on synthetic jobs (0 when it did not run).
mqw_status synthetic_rates(mqw_ctx* c, int n, const std::vector<Cand>& cands, std::vector<double>& gbs, bool verbose)
{
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
    EventPair ev;
    HIPCHK(c, ev.create());
    const int reps = 8;
    gbs.assign(cands.size(), 0.0);
    for (size_t i = 0; i < cands.size(); i++) {
        const Cand& cd = cands[i];
        const bool half = cd.tb >= 13;
        auto launch = [&](int q) {
            const void* jobs = jd + (size_t)(q & 1) * T;
            return half ? launch_half_term(cd.tb, cd.variant, jobs, (uint32_t)T, n, cd.lc, 0, c->stream)
                        : launch_term(MQW_FP32, cd.tb, cd.variant, jobs, (uint32_t)T, n, cd.lc, c->stream);
        };
        if (launch(0) || launch(1)) continue;
        HIPCHK(c, hipEventRecord(ev.a, c->stream));
        int bad = 0;
        for (int q = 0; q < reps; q++) bad |= launch(q);
        HIPCHK(c, hipEventRecord(ev.b, c->stream));
        HIPCHK(c, hipEventSynchronize(ev.b));
        if (bad) continue;
        float ms = 0.0f;
        hipEventElapsedTime(&ms, ev.a, ev.b);
        gbs[i] = 36.0 * (double)N * (double)T * reps / (ms * 1e-3) / 1e9;
        if (verbose)
            fprintf(stderr, "mqw_tune n=%d synthetic tb=%d variant=%d lc=%d %.1f GB/s\n", n, cd.tb, cd.variant, cd.lc,
                    gbs[i]);
    }
    return MQW_OK;
}

// Stage 2: algorithmic GB/s of a real solve with the configuration (0 on failure).
double solve_rate(mqw_ctx* c, int n, const std::vector<double>& rec, int samples, const Cand& cd)
{
    mqw_opts o;
    mqw_opts_default(&o);
    o.samples = samples;
    o.small_n_max = 0;
    o.tile_bits = cd.tb;
    o.chunk_bits = cd.lc;
    o.variant = cd.variant;
    o.high_pass = cd.hp;
    // every stage as term sweeps: the configuration under test is the sweep's
    o.shared_first_stage = 0;
    o.shared_second_stage = 0;
    o.final_readout = 0;
    float p = 0.0f, tt = 0.0f;
    if (mqw_solve(c, n, 3, rec.data(), 1, &o, &p, &tt) != MQW_OK) return 0.0;
    if (mqw_solve(c, n, 3, rec.data(), 1, &o, &p, &tt) != MQW_OK) return 0.0;
    return c->stats.sweep_ms > 0 ? c->stats.algorithmic_bytes / (c->stats.sweep_ms * 1e-3) / 1e9 : 0.0;
}

} // namespace

extern "C" mqw_status mqw_tune(mqw_ctx* c, int n, const char* cache_path, int* tile_bits, int* chunk_bits,
                               int* variant, int* high_pass)
{
    if (!c) return MQW_EINVAL;
    if (n < MQW_MIN_N || n > MQW_MAX_N) return fail(c, MQW_EINVAL, "n=%d outside [%d, %d]", n, MQW_MIN_N, MQW_MAX_N);
    hipSetDevice(c->device);
    const std::string arch = device_arch(c->device);
    Cand best{0, 0, 0, 0};
    const char* retune = getenv("MQW_RETUNE");
    const char* vb = getenv("MQW_TUNE_VERBOSE");
    const bool verbose = vb && *vb == '1';
    bool have = !(retune && *retune == '1') && cache_lookup(cache_path, arch, n, best);
    double best_gbs = 0.0;
    if (!have) {
        forget(c, n);
        const std::vector<Cand> cands = candidates(n);
        std::vector<double> gbs;
        mqw_status st = synthetic_rates(c, n, cands, gbs, verbose);
        if (st) return st;
        std::vector<size_t> order(cands.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return gbs[a] > gbs[b]; });
        // stage 2 candidates: the 5 fastest synthetic ones, the default, and
        // from n = 22 the high-bit pass (LDS and register pencils) with 2^14
        // tiles and the LDS pass with the 2 fastest real-half candidates
        std::vector<Cand> real;
        for (size_t i = 0; i < order.size() && real.size() < 5; i++)
            if (gbs[order[i]] > 0) real.push_back(cands[order[i]]);
        const SweepCfg d = choose_sweep(c, MQW_FP32, n, 0, -1, -1, -1);
        real.push_back({d.tb, d.variant, d.lc, d.hi > 0 ? (d.hwb < 0 ? 4 : d.hwb + 1) : 0});
        if (n >= 22 && d.half) {
            real.push_back({14, 4, 5, 2});
            real.push_back({14, 4, 5, 4});
            int added = 0;
            for (size_t i = 0; i < order.size() && added < 2; i++)
                if (cands[order[i]].tb >= 13 && gbs[order[i]] > 0) {
                    Cand h = cands[order[i]];
                    h.hp = 2;
                    real.push_back(h);
                    added++;
                }
        }
        // a generated instance, samples for about 2^26 amplitudes in flight
        std::vector<double> rec((size_t)n * (n + 1) / 2);
        generate_sk(n, 1, 7, rec.data());
        const int samples = (int)std::max<size_t>(2, std::min<size_t>(256, ((size_t)1 << 26) >> n));
        for (const Cand& cd : real) {
            const double r = solve_rate(c, n, rec, samples, cd);
            if (verbose)
                fprintf(stderr, "mqw_tune n=%d solve tb=%d variant=%d lc=%d high_pass=%d %.1f GB/s\n", n, cd.tb,
                        cd.variant, cd.lc, cd.hp, r);
            if (r > best_gbs) {
                best_gbs = r;
                best = cd;
            }
        }
        c->stats = mqw_stats{};
        if (best.tb == 0) return fail(c, MQW_EDEVICE, "no sweep configuration ran for n=%d", n);
        if (cache_path) {
            FILE* f = fopen(cache_path, "a");
            if (f) {
                fprintf(f, "%s %d %d %d %d %d %.1f\n", arch.c_str(), n, best.tb, best.variant, best.lc, best.hp,
                        best_gbs);
                fclose(f);
            }
        }
    }
    c->tuned_tb[n] = best.tb;
    c->tuned_var[n] = best.variant;
    c->tuned_lc[n] = best.lc;
    c->tuned_hp[n] = best.hp;
    if (tile_bits) *tile_bits = best.tb;
    if (chunk_bits) *chunk_bits = best.lc;
    if (variant) *variant = best.variant;
    if (high_pass) *high_pass = best.hp;
    return MQW_OK;
}
