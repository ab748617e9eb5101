// mqw: drop-in command line of MultiQW (MultiQW.cpp:241-458).
//
//   mqw n m filename [start problems]
//
// Reads `problems` instances of n(n+1)/2 float64 from `filename` starting at
// record `start` (both honoured only when both are given, MultiQW.cpp:259),
// runs the m-stage walk of each on the GPU, prints "<total_t> <prob>" per
// problem and appends the float32 probabilities to output_<n>_<m>.
// Deviations from the reference are errors instead of undefined behaviour:
// n outside [4, 30], a missing or short file, problems > the file holds.
//
// Environment (optional): MQW_DEVICE (0), MQW_SAMPLES (100), MQW_PRECISION
// (fp32|fp64), MQW_ENERGY (compat|exact), MQW_BATCH (problems per GPU call),
// MQW_FLUSH (1 = append each batch's results as soon as it finishes, so an
// interrupted run keeps them and can be resumed with `start`; the default
// writes one block at exit like MultiQW.cpp:457).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/mqw.h"

static int env_int(const char* name, int dflt)
{
    const char* v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

int main(int argc, char** argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s n m filename [start problems]\n", argv[0]);
        return -1;
    }
    const int n = atoi(argv[1]);
    const int m = atoi(argv[2]);
    const char* filename = argv[3];
    long start = 0, problems = MQW_DEFAULT_PROBLEMS;
    if (argc == 6) {
        start = atol(argv[4]);
        problems = atol(argv[5]);
    }
    if (n < MQW_MIN_N || n > MQW_MAX_N || m < 1 || m > MQW_MAX_STAGES || start < 0 || problems < 0) {
        fprintf(stderr, "mqw: need %d <= n <= %d, 1 <= m <= %d, start >= 0, problems >= 0\n", MQW_MIN_N, MQW_MAX_N,
                MQW_MAX_STAGES);
        return 1;
    }
    long have = 0;
    if (mqw_count_instances(filename, n, &have) != MQW_OK) {
        fprintf(stderr, "mqw: cannot read %s\n", filename);
        return 1;
    }
    if (start + problems > have) {
        fprintf(stderr, "mqw: %s holds %ld records of n=%d; asked for %ld from %ld\n", filename, have, n, problems,
                start);
        return 1;
    }
    mqw_opts o;
    mqw_opts_default(&o);
    o.samples = env_int("MQW_SAMPLES", MQW_DEFAULT_SAMPLES);
    const char* prec = getenv("MQW_PRECISION");
    if (prec && strcmp(prec, "fp64") == 0) o.precision = MQW_FP64;
    const char* en = getenv("MQW_ENERGY");
    if (en && strcmp(en, "exact") == 0) o.energy_mode = MQW_ENERGY_EXACT;
    long batch = env_int("MQW_BATCH", n <= 14 ? 2000 : (n <= 20 ? 32 : 4));
    if (batch < 1) batch = 1;
    const bool flush = env_int("MQW_FLUSH", 0) != 0;

    mqw_ctx* ctx = nullptr;
    mqw_status st = mqw_create(env_int("MQW_DEVICE", 0), &ctx);
    if (st != MQW_OK) {
        fprintf(stderr, "mqw: no usable GPU (%s)\n", mqw_status_string(st));
        return 1;
    }
    const long P = n * (n + 1) / 2;
    std::vector<float> results((size_t)problems);
    std::vector<double> recs;
    for (long b0 = 0; b0 < problems; b0 += batch) {
        const long cnt = problems - b0 < batch ? problems - b0 : batch;
        recs.resize((size_t)(cnt * P));
        if (mqw_read_instances(filename, n, start + b0, cnt, recs.data()) != MQW_OK) {
            fprintf(stderr, "mqw: short read from %s\n", filename);
            mqw_destroy(ctx);
            return 1;
        }
        std::vector<float> tt((size_t)cnt);
        st = mqw_solve(ctx, n, m, recs.data(), (int)cnt, &o, results.data() + b0, tt.data());
        if (st != MQW_OK) {
            fprintf(stderr, "mqw: %s: %s\n", mqw_status_string(st), mqw_last_error(ctx));
            mqw_destroy(ctx);
            return 1;
        }
        for (long k = 0; k < cnt; k++) printf("%g %g\n", (double)tt[k], (double)results[b0 + k]);
        fflush(stdout);
        if (flush && mqw_append_output(nullptr, n, m, results.data() + b0, cnt) != MQW_OK) {
            fprintf(stderr, "mqw: cannot append to output_%d_%d\n", n, m);
            mqw_destroy(ctx);
            return 1;
        }
    }
    mqw_destroy(ctx);
    if (!flush && mqw_append_output(nullptr, n, m, results.data(), problems) != MQW_OK) {
        fprintf(stderr, "mqw: cannot append to output_%d_%d\n", n, m);
        return 1;
    }
    return 0;
}
