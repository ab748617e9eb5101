// Host engine of libmqw: context, device workspace, energy build, the plan of
// Clenshaw sweeps for a batch of trajectories, hipGraph capture, and the C ABI
// declared in include/mqw.h.
//
// Reference: Asa-Hopkins/MultiStageQW MultiQW.cpp.  A "trajectory" is one
// chain of stages that starts from |+>: for m > 1 each Monte-Carlo sample is a
// trajectory of m stages (the state is reset per sample, :447-451); for m == 1
// the samples are chained (sorted, differenced times, :426-430) into one
// trajectory of `samples` stages with a readout after each.  Each stage is a
// Chebyshev series of K terms evaluated by the Clenshaw recurrence; its K-1
// sweeps (the reference's injection-only top term is folded into the first
// sweep) run as term-kernel launches in lockstep across the trajectories of a
// batch, aligned so every trajectory ends its stage on the same launch.
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "mqw_internal.h"

using namespace mqw;

mqw_status mqw::fail(mqw_ctx* c, mqw_status s, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return s;
}

namespace {
// Host phase times on stderr with MQW_TRACE=1: wall clock since the current
// solve_group began (per host thread).
thread_local std::chrono::steady_clock::time_point g_trace_t0;
bool trace_on()
{
    static const bool on = [] {
        const char* v = getenv("MQW_TRACE");
        return v && *v == '1';
    }();
    return on;
}
void trace(const char* fmt, ...)
{
    if (!trace_on()) return;
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - g_trace_t0).count();
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[mqw %9.3f ms] ", ms);
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    va_end(ap);
}
} // namespace

// Grow-only named device buffer.  Growing synchronises both streams first.
// Beyond 8 GiB a buffer grows in steps of 8 GiB: allocating tens of GiB takes
// about a second, and nearby sizes (the bases of different instances) then
// share one allocation.
void* mqw::dev_get(mqw_ctx* c, const std::string& name, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    DevBuf& b = c->bufs[name];
    if (b.cap >= bytes) return b.ptr;
    hipStreamSynchronize(c->stream);
    if (c->stream2) hipStreamSynchronize(c->stream2);
    if (b.ptr) hipFree(b.ptr);
    b.ptr = nullptr;
    b.cap = 0;
    const size_t step = bytes > ((size_t)8 << 30) ? ((size_t)8 << 30) : 4096;
    const size_t want = (bytes + step - 1) / step * step;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
        hipGetLastError();
        b.ptr = nullptr;
        trace("device buffer %s: %zu bytes failed", name.c_str(), want);
        return nullptr;
    }
    b.cap = want;
    trace("device buffer %s: %zu bytes", name.c_str(), want);
    return b.ptr;
}

void mqw::release_bufs(mqw_ctx* c)
{
    hipStreamSynchronize(c->stream);
    if (c->stream2) hipStreamSynchronize(c->stream2);
    for (auto& kv : c->bufs)
        if (kv.second.ptr) hipFree(kv.second.ptr);
    c->bufs.clear();
}

// Smallest n at which the high-bit pass is on by default (at n = 24 the
// register variant over 5 bits runs 4.8 % more applies/s than partner loads
// alone), and the row width of the LDS variant (128 B rows: measured best of
// 64, 128 and 256 B at n = 23..28).
static constexpr int kHighPassMinN = 24;
static constexpr int kHighPassWb = 1;

static int default_tile_bits(int precision, int n)
{
    const int tb = precision == MQW_FP32 ? 12 : 11;
    return std::min(n, tb);
}

static int default_chunk_bits(int n, int tb)
{
    // runs of 64 tiles (a sub-cube of the outer bits) per XCD measured best
    return std::min(n - tb, 6);
}

// Which sweep kernel runs a group: the real-half kernel (FP32, n >= its tile
// bits) or the complex-tile kernel.  `tb` is the tile bits of that kernel and
// `parts` the number of partial norm sums one output leaves.
SweepCfg mqw::choose_sweep(const mqw_ctx* c, int precision, int n, int tile_bits, int chunk_bits, int variant,
                           int high_pass)
{
    SweepCfg k;
    if (high_pass < 0 && c->tuned_hp[n] >= 0) high_pass = c->tuned_hp[n];
    k.variant = variant >= 0 ? variant : c->tuned_var[n];
    int tb = tile_bits > 0 ? tile_bits : c->tuned_tb[n];
    int dflt_lc = -1;
    // Untuned FP32 default from n = 14, measured on MI355X over batches of
    // 2^31/2^n trajectories (at most 1600): real-half tiles, non-temporal
    // output stores; 2^14-float tiles with one partner load in flight while
    // the outer tile bits fit a 2^5..2^6-tile XCD chunk (n <= 21), 2^13
    // tiles with two loads in flight above, and 2^14 tiles again once the
    // high-bit pass takes the top bits (n >= kHighPassMinN).
    if (tb <= 0 && precision == MQW_FP32 && n >= 14) {
        int dv;
        if (n == 14) {
            tb = 13, dv = 4, dflt_lc = 3;
        } else if (n <= 19) {
            tb = 14, dv = 3, dflt_lc = 5;
        } else if (n <= 21) {
            tb = 14, dv = 3, dflt_lc = 6;
        } else if (n < kHighPassMinN || high_pass == 0) {
            tb = 13, dv = 4, dflt_lc = 5;
        } else {
            tb = 14, dv = 4, dflt_lc = 5;
        }
        if (variant < 0) k.variant = dv;
    }
    if (precision == MQW_FP32 && n >= 13 && tb > 0 && half_tile_bits_supported(tb)) {
        if (tb > n) tb = 13;
        k.half = true;
        k.tb = tb;
        k.parts = 2u << (n - tb);
        k.lc = chunk_bits >= 0 ? chunk_bits
                               : (c->tuned_lc[n] >= 0 ? c->tuned_lc[n]
                                                      : (dflt_lc >= 0 ? dflt_lc : std::min(n - tb + 1, 6)));
    } else {
        if (tb <= 0 || tb > n || !term_tile_bits_supported(precision, tb)) tb = default_tile_bits(precision, n);
        k.tb = tb;
        k.parts = 1u << (n - tb);
        k.lc = chunk_bits >= 0 ? chunk_bits : (c->tuned_lc[n] >= 0 ? c->tuned_lc[n] : default_chunk_bits(n, tb));
    }
    k.lc = std::max(0, std::min(k.lc, 24));
    // High-bit pass: from n = kHighPassMinN by default (measured), or as asked
    // (high_pass = 1, 2, 3: LDS pencil rows of 64, 128, 256 B, up to 10 - wb
    // of the outer bits; 4: register pencils for the 4..7 bits beyond the XCD
    // chunk).  The sweep keeps partner loads for the rest.
    const bool pass_on = high_pass > 0 || (high_pass < 0 && n >= kHighPassMinN);
    if (k.half && pass_on) {
        int wb, hi;
        // by n: the register variant while the bits beyond the chunk are at
        // most 7 (measured 3 % faster at n = 25, 26), else 128 B LDS pencils
        if (high_pass < 0 && n - k.tb - k.lc <= 7) high_pass = 4;
        if (high_pass == 4) {
            // register variant: just the outer bits beyond the XCD chunk, 4..7 of them
            wb = -1;
            hi = std::max(4, std::min(7, n - k.tb - k.lc));
        } else {
            wb = high_pass > 0 ? std::min(high_pass, 3) - 1 : kHighPassWb;
            hi = std::min(10 - wb, n - k.tb);
        }
        hi = std::min(hi, n - k.tb);
        if (high_pass_supported(n, hi, wb)) {
            k.hi = hi;
            k.hwb = wb;
        }
    }
    return k;
}

namespace {

bool finite_rec(int n, const double* rec)
{
    const int P = n * (n + 1) / 2;
    for (int k = 0; k < P; k++)
        if (!isfinite(rec[k])) return false;
    return true;
}

// ---------------------------------------------------------------------------
// Per-problem host state.
// ---------------------------------------------------------------------------
struct Problem {
    int n = 0, m = 0;
    const double* rec = nullptr;
    float J[kMaxN * kMaxN];
    float h[kMaxN];
    mqw_schedule sch{};
    std::vector<float> times; // m * samples
    mqw_energy_info info{};
    int out = 0;              // index into the caller's outputs
    size_t prob_base = 0;     // first readout slot
};

struct Segment {
    float gamma, onenorm;
    double s = 0.0;        // onenorm * t: the argument of the coefficients
    std::vector<double> c; // FP64 coefficients, K = c.size()
    int64_t prob_slot;     // -1 = no readout
};

struct Traj {
    int prob;                     // index into the group's problem list
    std::vector<Segment> segs;
    const void* psi_in = nullptr; // input of the first stage; nullptr = |+>
    bool first_done = false;      // the first stage comes from the shared basis (a combine op)
    bool second_done = false;     // the first two stages come from the shared bases (a combine2 op)
    bool adjoint = false;         // the last stage is read out by the adjoint dots (no terms)
    bool adjoint2 = false;        // the last two stages are read out by the two-stage adjoint (no terms)
    std::vector<float> c1q;       // with a compressed first-stage basis: the sample's coordinates in it
    std::vector<float> cxq;       // with a compressed last-stage adjoint basis: the sample's coordinates in it
    bool direct_done = false;     // an m == 1 chain evaluated directly (direct_m1): no terms run
    bool mid_done = false;        // the first three stages from shared bases (the third from the
                                  // compressed basis of the second-stage states, a combine2 op)
    std::vector<float> cmq;       // the sample's coordinates in that basis
    int seg0 = 0;                 // with no combine op: the first segment to run, psi_in its input
};

// Chebyshev coefficients of every segment.  Segments are independent (the
// Miller scratch of cheb_coeffs_vec is per thread), so large batches spread
// over up to 16 host threads; the values do not depend on the split.
void fill_coeffs(std::vector<Traj>& trajs, double tol, bool to_float)
{
    std::vector<Segment*> segs;
    for (Traj& tr : trajs)
        for (Segment& sg : tr.segs) segs.push_back(&sg);
    auto work = [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) {
            Segment& sg = *segs[i];
            cheb_coeffs_vec(sg.s, tol, sg.c);
            if (to_float)
                for (double& v : sg.c) v = (double)(float)v; // the reference casts to float (:440)
        }
    };
    const size_t total = segs.size();
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nt = std::min<size_t>({(size_t)(hw ? hw : 1), 16, total / 1024 + 1});
    if (nt <= 1) {
        work(0, total);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; t++) th.emplace_back(work, total * t / nt, total * (t + 1) / nt);
    work(0, total / nt);
    for (std::thread& x : th) x.join();
}

// ---------------------------------------------------------------------------
// Energies for a group of problems of one n: E (float) or E64 (double) per
// problem on the device, infos copied back to the host.
// ---------------------------------------------------------------------------
mqw_status group_energies(mqw_ctx* c, int n, std::vector<Problem*>& ps, int mode, int precision, void* E_out)
{
    const size_t N = (size_t)1 << n;
    const size_t cnt = ps.size();
    const size_t nb = energy_blocks(n);
    // FP32: every problem in one set of launches (block tables per problem);
    // FP64: one problem after another through a shared float buffer
    const bool batched = precision == MQW_FP32;
    const size_t nt = batched ? cnt : 1; // problems whose block tables coexist
    float* Jh = (float*)dev_get(c, "energy_Jh", sizeof(float) * (size_t)(n * n + n) * cnt);
    double* scr = (double*)dev_get(c, "energy_scratch", sizeof(double) * 4 * nb * nt);
    double* e_inits = (double*)dev_get(c, "energy_e_init", sizeof(double) * cnt);
    EnergyDev* info = (EnergyDev*)dev_get(c, "energy_info", sizeof(EnergyDev) * cnt);
    float* hp_tmp = nullptr;
    if (precision == MQW_FP64) hp_tmp = (float*)dev_get(c, "energy_hp_tmp", sizeof(float) * N);
    if (!Jh || !scr || !e_inits || !info || (precision == MQW_FP64 && !hp_tmp))
        return fail(c, MQW_ENOMEM, "device allocation failed in energy build (n=%d, %zu problems)", n, cnt);
    std::vector<float> hJh((size_t)(n * n + n) * cnt);
    std::vector<double> he(cnt);
    for (size_t p = 0; p < cnt; p++) {
        memcpy(&hJh[p * (n * n + n)], ps[p]->J, sizeof(float) * n * n);
        memcpy(&hJh[p * (n * n + n) + n * n], ps[p]->h, sizeof(float) * n);
        he[p] = mode == MQW_ENERGY_EXACT ? e_init_exact(n, ps[p]->J, ps[p]->h) : e_init_compat(n, ps[p]->J, ps[p]->h);
    }
    // (the copies read pageable host memory: the stream drains before it goes)
    struct Drain {
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{c->stream};
    HIPCHK(c, hipMemcpyAsync(Jh, hJh.data(), sizeof(float) * hJh.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(e_inits, he.data(), sizeof(double) * cnt, hipMemcpyHostToDevice, c->stream));
    EnergyScratch es{scr, scr + nb * nt, (uint32_t*)(scr + 2 * nb * nt), scr + 3 * nb * nt};
    if (batched && launch_energies(n, (uint32_t)cnt, Jh, e_inits, mode == MQW_ENERGY_EXACT, (float*)E_out, nullptr,
                                   1, es, info, c->stream))
        return fail(c, MQW_EDEVICE, "energy kernels: %s", last_kernel_error());
    for (size_t p = 0; !batched && p < cnt; p++) {
        const float* Jd = Jh + p * (n * n + n);
        const float* hd = Jd + n * n;
        float* hp = precision == MQW_FP32 ? (float*)E_out + p * N : hp_tmp;
        double* e64 = precision == MQW_FP64 ? (double*)E_out + p * N : nullptr;
        if (launch_energy(n, Jd, hd, he[p], mode == MQW_ENERGY_EXACT, hp, e64, 1, es, info + p, c->stream))
            return fail(c, MQW_EDEVICE, "energy kernels: %s", last_kernel_error());
    }
    std::vector<EnergyDev> hinfo(cnt);
    HIPCHK(c, hipMemcpyAsync(hinfo.data(), info, sizeof(EnergyDev) * cnt, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t p = 0; p < cnt; p++) {
        mqw_energy_info& o = ps[p]->info;
        o.e0 = hinfo[p].e0;
        o.emax = hinfo[p].emax;
        o.eloc = hinfo[p].eloc;
        o.shift = hinfo[p].shift;
        o.e_abs = hinfo[p].e_abs;
        o.e_init = he[p];
    }
    return MQW_OK;
}

// ---------------------------------------------------------------------------
// Sweep plans (global-memory path).
//
// Each trajectory becomes a linear sequence of term jobs across all of its
// stages.  Trajectories run in groups whose buffers fit the device's memory-
// side cache (Infinity Cache): a group's launch q holds job q of every member
// still running, so members progress independently through their stages and
// nobody idles at stage boundaries.  Groups run one after another on the
// stream and reuse the same buffer slots.  The stage norm is reduced inside
// the next stage's first term (F_NORM_IN) and a readout folds into it
// (F_READOUT); a reduce launch per group reads out the last stage.
// ---------------------------------------------------------------------------
struct Op {
    int kind; // 0 = term launch, 1 = reduce launch, 2 = high-bit pass over the jobs of the next term launch,
              // 3 = combine launch (shared first stage), 4 = real-half term launch (basis recurrences),
              // 5 = adjoint readout dots, 6 = adjoint readout finalize, 7 = combine2 launch (second stage),
              // 8 = pair dots of the two-stage adjoint, 9 = on-chip recurrences (the plan's chains)
    size_t off;
    uint32_t count;
    uint32_t aux = 0;  // kinds 3, 5: largest basis size of its jobs / tiles; kind 7: largest tile
    uint32_t aux2 = 0; // kinds 5, 6, 8: amplitude chunks of the dots; kind 7: 1 = split-FP16 kernel
    uint32_t aux3 = 0; // kind 5: 1 = dots on the matrix cores (split FP16)
};

template <typename real>
struct Plan {
    std::vector<TermJob<real>> jobs;
    std::vector<ReduceJob> rjobs;
    std::vector<CombineJob> cjobs;
    std::vector<CombineItem> citems;
    std::vector<float> ccoef;
    std::vector<ReadTile> rtiles;
    std::vector<ReadItem> ritems;
    std::vector<Combine2Tile> c2tiles;
    std::vector<Combine2Item> c2items;
    size_t dots_doubles = 0;  // largest readout-dot scratch of one op
    double aux_bytes = 0.0;   // algorithmic bytes of the combine and readout launches
    double gemm_flops = 0.0;        // matrix-core flops the combine2 launches execute (padded rows)
    double gemm_useful_flops = 0.0; // of which on real samples
    size_t c2a_halves = 0;          // FP16 coefficient matrices of the split-FP16 combine2 launches
    double gemm_bytes = 0.0;        // their algorithmic bytes (also in aux_bytes)
    uint64_t aux_launches = 0;
    std::vector<Op> ops;
    std::vector<ChainJob> chains;        // on-chip recurrences (op kind 9): the chains,
    std::vector<uint32_t> corder, cgoff; // their order by group and each group's first
    uint64_t amp_terms = 0;
    uint64_t psi_const_amp_terms = 0;
    uint64_t terms = 0;
    uint64_t sweeps = 0;
};

// Buffers of one trajectory slot: three planar states and two partial-sum arrays.
template <typename real>
struct Slot {
    real* st[3];
    double* part[2];
    real* hs = nullptr; // high-bit sums of the current term's s1 (planar), when the pass is on
};

// The input of an adjoint trajectory's read-out stages: its state and partial
// norms, and the slot's other state and partial-sum array, free once the
// trajectory's terms have run (the two-stage adjoint writes phi' there).
struct AdjointIn {
    const void* psi = nullptr;
    const double* norm_part = nullptr;
    void* free_st = nullptr;
    double* free_part = nullptr;
};

// Event pairs and graphs of launched op runs, with each run's timing class
// (0 sweep, 1 basis recurrences, 2 GEMM, 3 other auxiliary launches); kept
// until the work has finished and its times are read.
struct Timed {
    std::deque<EventPair> evs;
    std::deque<GraphGuard> graphs;
    std::vector<int> cls;
};

// Elapsed times of finished runs, by class.
void collect(Timed& tm, double* ms_sweep, double* ms_aux, double* ms_basis, double* ms_gemm)
{
    for (size_t q = 0; q < tm.cls.size() && q < tm.evs.size(); q++) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, tm.evs[q].a, tm.evs[q].b) != hipSuccess) continue;
        *(tm.cls[q] ? ms_aux : ms_sweep) += t;
        if (tm.cls[q] == 1 && ms_basis) *ms_basis += t;
        if (tm.cls[q] == 2 && ms_gemm) *ms_gemm += t;
    }
}

template <typename real>
void traj_jobs(int n, uint32_t parts, const Traj& tr, const Slot<real>& sl, const real* E, uint32_t eloc,
               double* probs, std::vector<TermJob<real>>& out, ReduceJob& fin, bool& has_fin, int hi = 0,
               AdjointIn* adj = nullptr)
{
    const size_t N = (size_t)1 << n;
    const real psi0 = (real)(float)(1 / sqrt((double)N));
    int p = -1, a = 0, b = 1, par = 0;
    int64_t prev_slot = -1;
    size_t s0 = 0;
    if (tr.first_done || tr.second_done || tr.mid_done) {
        // a combine op left the first (combine2: second, the middle combine2:
        // third) stage's output in st[0] and its partial norms in part[0]
        p = 0;
        a = 1;
        b = 2;
        par = 1;
        s0 = tr.mid_done ? 3 : tr.second_done ? 2 : 1;
        prev_slot = tr.segs[s0 - 1].prob_slot;
    } else {
        s0 = (size_t)tr.seg0;
    }
    auto rc = [](double v) { return (real)v; };
    // an adjoint trajectory stops before its last stage, a two-stage adjoint
    // one before its last two
    const size_t s1 = tr.segs.size() - (tr.adjoint2 ? 2 : tr.adjoint ? 1 : 0);
    for (size_t s = s0; s < s1; s++) {
        const Segment& sg = tr.segs[s];
        const std::vector<double>& cc = sg.c;
        const int K = (int)cc.size();
        const int S = std::max(K - 1, 1);
        const bool pconst = p < 0 && !tr.psi_in;
        const real* psi = p >= 0 ? sl.st[p] : (const real*)tr.psi_in;
        for (int d = 0; d < S; d++) {
            TermJob<real> j{};
            j.psi = psi;
            j.psi0 = psi0;
            j.E = E;
            j.eloc = eloc;
            uint32_t flags = pconst ? F_PSI_CONST : 0u;
            if (d == 0 && p >= 0) {
                flags |= F_NORM_IN;
                j.norm_in = sl.part[par ^ 1];
                if (prev_slot >= 0) {
                    flags |= F_READOUT;
                    j.prob = probs + prev_slot;
                }
            }
            if (K == 1) {
                flags |= SRC1_ZERO | SRC2_ZERO | F_FINAL;
                j.out = sl.st[a];
                j.cr = rc(cc[0]);
            } else {
                const int r = K - 2 - d;
                j.out = sl.st[(d & 1) ? b : a];
                const double top = cc[K - 1];
                const bool top_odd = ((K - 1) & 1) != 0;
                if (d == 0) {
                    flags |= SRC1_PSI;
                    j.w1r = top_odd ? real(0) : rc(top);
                    j.w1i = top_odd ? rc(top) : real(0);
                } else {
                    flags |= SRC1_ARRAY;
                    j.b1 = sl.st[((d - 1) & 1) ? b : a];
                    j.w1r = real(1);
                }
                if (d == 0) {
                    flags |= SRC2_ZERO;
                } else if (d == 1) {
                    flags |= SRC2_PSI;
                    j.w2r = top_odd ? real(0) : rc(top);
                    j.w2i = top_odd ? rc(top) : real(0);
                } else {
                    flags |= SRC2_ARRAY;
                    j.b2 = j.out;
                    j.w2r = real(1);
                }
                const bool odd = (r & 1) != 0;
                j.cr = odd ? real(0) : rc(cc[r]);
                j.ci = odd ? rc(cc[r]) : real(0);
                const double sc = (double)sg.onenorm * (r == 0 ? 1.0 : 0.5);
                j.alpha = (real)(1.0 / sc);
                j.beta = (real)((double)sg.gamma / sc);
                if (r == 0) flags |= F_FINAL;
                // s1 is an array (not |+>): its high-bit sums come from the pass
                if (hi > 0 && !(d == 0 && pconst)) {
                    flags |= F_HSUM;
                    j.hsum = sl.hs;
                }
            }
            if (flags & F_FINAL) j.norm_part = sl.part[par];
            j.flags = flags;
            out.push_back(j);
        }
        const int fout = (K == 1) ? a : (((S - 1) & 1) ? b : a);
        int others[2], k = 0;
        for (int q = 0; q < 3; q++)
            if (q != fout) others[k++] = q;
        p = fout;
        a = others[0];
        b = others[1];
        par ^= 1;
        prev_slot = sg.prob_slot;
    }
    if (tr.adjoint || tr.adjoint2) {
        has_fin = false;
        if (adj) {
            adj->psi = sl.st[p];
            adj->norm_part = sl.part[par ^ 1];
            adj->free_st = sl.st[a];
            adj->free_part = sl.part[par];
        }
        return;
    }
    has_fin = prev_slot >= 0;
    if (has_fin) {
        fin = ReduceJob{};
        fin.norm_part = sl.part[par ^ 1];
        fin.out = sl.st[p];
        fin.N = N;
        fin.ntiles = parts;
        fin.eloc = eloc;
        fin.prob = probs + prev_slot;
    }
}

// Per-problem shared bases (FP32): the first stage's V (nullptr = none, kc
// vectors with V_0 = |+> implicit) and the last stage's adjoint X (kx vectors
// from X_0 = e_eloc).
struct Bases {
    std::vector<const float*> V;
    std::vector<int> kc;
    std::vector<const float*> X;
    std::vector<int> kx;
    // shared second stage: W (kc x kc2 vectors, l >= 1 stored), its row table and |+> as a vector
    std::vector<const float*> W;
    std::vector<int> kc2;
    std::vector<const uint32_t*> qtab;
    std::vector<uint32_t> q0, q1;
    const float* plus = nullptr;
    bool split = true; // the GEMM on FP16 matrix ops with split operands (else FP32 matrix ops)
    // per instance: 1 = V holds a compressed first-stage basis Q_1.. (kc = 2R
    // vectors, Q_0 at plusv), computed before the batch; no V recurrence then
    std::vector<char> cmp;
    std::vector<const float*> plusv;
    // two-stage adjoint (m >= 4): the compressed last-stage adjoint basis Z
    // (kz = 2R vectors, Z_0 .. Z_{kz-1} N floats apart, computed before the
    // batch), Y_lr = T_l(Ht_{m-1}) Z_r (l >= 1 stored like W) and its GEMM
    // row table; kz = 0 = none
    std::vector<const float*> Z;
    std::vector<int> kz;
    std::vector<const float*> Y;
    std::vector<int> ky;
    std::vector<const uint32_t*> ytab;
    std::vector<uint32_t> y0, y1;
    // the compressed middle stage (stage 3): U (ku = 2R directions of the
    // samples' second-stage states, N floats apart, made in the batch), W3_lr =
    // T_l(Ht_3) U_r (l >= 1 stored like W) and its GEMM row table; ku = 0 = none
    std::vector<const float*> U;
    std::vector<int> ku;
    std::vector<const float*> W3;
    std::vector<int> k3;
    std::vector<const uint32_t*> mtab;
    std::vector<uint32_t> m0, m1;
};

// Rows (l, k) of the second-stage GEMM in two classes, each padded to a
// multiple of kC2RowPad: real ((l&1) == (k&1), minus when both are odd), then
// imaginary.
void second_stage_rows(int kc1, int kc2, std::vector<uint32_t>& rows, uint32_t& q0, uint32_t& q1)
{
    rows.clear();
    for (int cls = 0; cls < 2; cls++) {
        const size_t b = rows.size();
        for (int l = 0; l < kc2; l++)
            for (int k = 0; k < kc1; k++) {
                if (((l & 1) != (k & 1)) != (cls == 1)) continue;
                const uint32_t sg = (cls == 0 && (l & 1)) ? 2u : 1u;
                rows.push_back((uint32_t)l | ((uint32_t)k << 8) | (sg << 16));
            }
        while ((rows.size() - b) % kC2RowPad) rows.push_back(0u); // l = k = 0 with no weight
        (cls ? q1 : q0) = (uint32_t)(rows.size() - b);
    }
}

// Readout-dot tiles for the adjoint members of one group: per problem, tiles
// of up to kReadTile samples; the dots are split over amplitude chunks so a
// launch has about 2048 workgroups.
template <typename real>
void readout_ops(int n, uint32_t parts, const std::vector<Traj>& trajs, const std::vector<size_t>& grp,
                 const std::vector<AdjointIn>& adj, const Bases& B, double* probs, Plan<real>& P)
{
    const size_t N = (size_t)1 << n;
    std::map<int, std::vector<size_t>> byprob;
    for (size_t i = 0; i < grp.size(); i++)
        if (trajs[grp[i]].adjoint) byprob[trajs[grp[i]].prob].push_back(i);
    if (byprob.empty()) return;
    const size_t t0 = P.rtiles.size();
    uint32_t kx_max = 1;
    for (const auto& kv : byprob) {
        const int pr = kv.first, kx = B.kx[pr];
        const std::vector<size_t>& mem = kv.second;
        for (size_t i0 = 0; i0 < mem.size(); i0 += kReadTile) {
            const size_t ns = std::min<size_t>(kReadTile, mem.size() - i0);
            ReadTile rt{};
            rt.X = B.X[pr];
            rt.kx = (uint32_t)kx;
            rt.ns = (uint32_t)ns;
            rt.item_off = (uint32_t)P.ritems.size();
            float amax = 0.0f; // largest coefficient magnitude of the tile
            for (size_t q = 0; q < ns; q++) {
                const size_t i = mem[i0 + q];
                const Traj& tr = trajs[grp[i]];
                const std::vector<double>& cc = tr.segs.back().c;
                ReadItem it{};
                it.psi = (const float*)adj[i].psi;
                it.norm_part = adj[i].norm_part;
                it.prob = probs + tr.segs.back().prob_slot;
                it.coef_off = (uint32_t)P.ccoef.size();
                for (int k = 0; k < kx; k++) {
                    P.ccoef.push_back(k < (int)cc.size() ? (float)cc[k] : 0.0f);
                    amax = std::max(amax, fabsf(P.ccoef.back()));
                }
                P.ritems.push_back(it);
            }
            // FP16 coefficients: a power of two taking the largest to [2^13, 2^14]
            const int e = amax > 0.0f ? std::max(-30, std::min(60, (int)floor(log2(16384.0 / amax)))) : 0;
            rt.ascale = (float)ldexp(1.0, e);
            rt.oscale = (float)ldexp(1.0, -e - 12); // and the basis's 2^12
            P.rtiles.push_back(rt);
            kx_max = std::max<uint32_t>(kx_max, (uint32_t)kx);
            // reads the kx basis vectors and the ns states once
            P.aux_bytes += (double)N * (4.0 * kx + 8.0 * (double)ns);
        }
    }
    const uint32_t ntiles = (uint32_t)(P.rtiles.size() - t0);
    // chunks of at least 4096 amplitudes; about 2048 workgroups per launch
    uint32_t nchunk = 1;
    while ((uint64_t)nchunk * ntiles < 2048 && (N / (nchunk * 2)) >= 4096) nchunk *= 2;
    P.dots_doubles = std::max<size_t>(P.dots_doubles, (size_t)ntiles * nchunk * kReadTile * 2);
    // the matrix-core dots up to n = kReadMfmaMaxN, the FP32 vector dots beyond
    Op d{5, t0, ntiles, kx_max, nchunk, B.split && n <= kReadMfmaMaxN ? 1u : 0u};
    Op f{6, t0, ntiles, kx_max, nchunk};
    P.ops.push_back(d);
    P.ops.push_back(f);
    P.aux_launches += 2;
    (void)parts;
}

// Pair-dot tiles of the two-stage adjoint members of one group, after their
// phi' GEMM: per problem, tiles of up to kReadTile samples whose items pair
// phi' (the slot's free state) with the input psi of stage m - 1; the dots are
// split over amplitude chunks like the readout's, and the readout's finalize
// divides |amplitude|^2 by psi's norm.
template <typename real>
void pair_dot_ops(int n, const std::vector<Traj>& trajs, const std::vector<size_t>& grp,
                  const std::vector<AdjointIn>& adj, double* probs, Plan<real>& P)
{
    const size_t N = (size_t)1 << n;
    std::map<int, std::vector<size_t>> byprob;
    for (size_t i = 0; i < grp.size(); i++)
        if (trajs[grp[i]].adjoint2) byprob[trajs[grp[i]].prob].push_back(i);
    if (byprob.empty()) return;
    const size_t t0 = P.rtiles.size();
    for (const auto& kv : byprob) {
        const std::vector<size_t>& mem = kv.second;
        for (size_t i0 = 0; i0 < mem.size(); i0 += kReadTile) {
            const size_t ns = std::min<size_t>(kReadTile, mem.size() - i0);
            ReadTile rt{};
            rt.ns = (uint32_t)ns;
            rt.item_off = (uint32_t)P.ritems.size();
            for (size_t q = 0; q < ns; q++) {
                const size_t i = mem[i0 + q];
                ReadItem it{};
                it.psi = (const float*)adj[i].psi;
                it.phi = (const float*)adj[i].free_st;
                it.norm_part = adj[i].norm_part;
                it.prob = probs + trajs[grp[i]].segs.back().prob_slot;
                P.ritems.push_back(it);
            }
            P.rtiles.push_back(rt);
            // reads both planar states of every item once
            P.aux_bytes += 16.0 * (double)N * (double)ns;
        }
    }
    const uint32_t ntiles = (uint32_t)(P.rtiles.size() - t0);
    uint32_t nchunk = 1;
    while ((uint64_t)nchunk * ntiles * kReadTile < 4096 && (N / (nchunk * 2)) >= 4096) nchunk *= 2;
    P.dots_doubles = std::max<size_t>(P.dots_doubles, (size_t)ntiles * nchunk * kReadTile * 2);
    P.ops.push_back(Op{8, t0, ntiles, 0, nchunk});
    P.ops.push_back(Op{6, t0, ntiles, 0, nchunk});
    P.aux_launches += 2;
}

// Combine jobs for the first_done members of one group: per problem, batches
// of up to kCombineBatch samples, each writing its slot's st[0] and part[0].
template <typename real>
void combine_ops(int n, const std::vector<Traj>& trajs, const std::vector<size_t>& grp, const Bases& B,
                 const std::vector<Slot<real>>& slots, Plan<real>& P)
{
    if constexpr (std::is_same<real, float>::value) {
        const size_t N = (size_t)1 << n;
        std::map<int, std::vector<size_t>> byprob; // problem -> member indices
        for (size_t i = 0; i < grp.size(); i++)
            if (trajs[grp[i]].first_done) byprob[trajs[grp[i]].prob].push_back(i);
        if (byprob.empty()) return;
        Op op{3, P.cjobs.size(), 0, 0};
        for (const auto& kv : byprob) {
            const int pr = kv.first, kc = B.kc[pr];
            const std::vector<size_t>& mem = kv.second;
            for (size_t i0 = 0; i0 < mem.size(); i0 += kCombineBatch) {
                const size_t ns = std::min<size_t>(kCombineBatch, mem.size() - i0);
                CombineJob cj{};
                cj.V = B.V[pr];
                cj.kc = (uint32_t)kc;
                cj.ns = (uint32_t)ns;
                cj.item_off = (uint32_t)P.citems.size();
                cj.psi0 = (float)(1 / sqrt((double)N));
                for (size_t q = 0; q < ns; q++) {
                    const size_t i = mem[i0 + q];
                    const std::vector<double>& cc = trajs[grp[i]].segs[0].c;
                    CombineItem it{};
                    it.out = slots[i].st[0];
                    it.norm_part = slots[i].part[0];
                    it.coef_off = (uint32_t)P.ccoef.size();
                    for (int k = 0; k < kc; k++) P.ccoef.push_back(k < (int)cc.size() ? (float)cc[k] : 0.0f);
                    P.citems.push_back(it);
                }
                P.cjobs.push_back(cj);
                op.count++;
                op.aux = std::max<uint32_t>(op.aux, (uint32_t)kc);
                // reads the kc - 1 stored vectors once, writes both halves of every sample
                P.aux_bytes += (double)N * (4.0 * (kc - 1) + 8.0 * (double)ns);
            }
        }
        P.ops.push_back(op);
        P.aux_launches++;
    } else {
        (void)n, (void)trajs, (void)grp, (void)B, (void)slots, (void)P;
    }
}

// Combine2 tiles of one group, per problem in tiles of up to kC2Tile
// (kC2TileH) samples: the second_done members' second stage, written to their
// slot's st[0] and part[0]; or, with adj, the two-stage adjoint members'
// phi'_j = P_{m-1}^T phi_j (basis Y from Z, coefficients: the sample's
// coordinates in Z and its stage m - 1 coefficients), written to the free
// state of their slot.
template <typename real>
void combine2_ops(int n, const std::vector<Traj>& trajs, const std::vector<size_t>& grp, const Bases& B,
                  const std::vector<Slot<real>>& slots, Plan<real>& P, const std::vector<AdjointIn>* adj = nullptr,
                  bool mid = false)
{
    if constexpr (std::is_same<real, float>::value) {
        const size_t N = (size_t)1 << n;
        std::map<int, std::vector<size_t>> byprob;
        for (size_t i = 0; i < grp.size(); i++)
            if (adj ? trajs[grp[i]].adjoint2 : mid ? trajs[grp[i]].mid_done : trajs[grp[i]].second_done)
                byprob[trajs[grp[i]].prob].push_back(i);
        if (byprob.empty()) return;
        Op op{7, P.c2tiles.size(), 0, 0, B.split ? 1u : 0u};
        const size_t tsz = B.split ? kC2TileH : kC2Tile;
        for (const auto& kv : byprob) {
            const int pr = kv.first;
            const int kc1 = adj ? B.kz[pr] : mid ? B.ku[pr] : B.kc[pr];
            const int kc2 = adj ? B.ky[pr] : mid ? B.k3[pr] : B.kc2[pr];
            const std::vector<size_t>& mem = kv.second;
            for (size_t i0 = 0; i0 < mem.size(); i0 += tsz) {
                const size_t ns = std::min<size_t>(tsz, mem.size() - i0);
                Combine2Tile ct{};
                const bool cq = adj || mid || (!B.cmp.empty() && B.cmp[pr]);
                if (adj) {
                    ct.plus = B.Z[pr];
                    ct.V = B.Z[pr] + N;
                    ct.W = B.Y[pr];
                    ct.qtab = B.ytab[pr];
                    ct.q0 = B.y0[pr];
                    ct.q1 = B.y1[pr];
                } else if (mid) {
                    ct.plus = B.U[pr];
                    ct.V = B.U[pr] + N;
                    ct.W = B.W3[pr];
                    ct.qtab = B.mtab[pr];
                    ct.q0 = B.m0[pr];
                    ct.q1 = B.m1[pr];
                } else {
                    ct.plus = cq ? B.plusv[pr] : B.plus;
                    ct.V = B.V[pr];
                    ct.W = B.W[pr];
                    ct.qtab = B.qtab[pr];
                    ct.q0 = B.q0[pr];
                    ct.q1 = B.q1[pr];
                }
                ct.kc1 = (uint32_t)kc1;
                ct.kc2 = (uint32_t)kc2;
                ct.item_off = (uint32_t)P.c2items.size();
                ct.ns = (uint32_t)ns;
                float m1 = 0.0f, m2 = 0.0f; // largest coefficient magnitudes of the tile
                for (size_t q = 0; q < ns; q++) {
                    const size_t i = mem[i0 + q];
                    const Traj& tr = trajs[grp[i]];
                    Combine2Item it{};
                    it.out = adj ? (float*)(*adj)[i].free_st : slots[i].st[0];
                    it.norm_part = adj ? (*adj)[i].free_part : slots[i].part[0];
                    const std::vector<float>& u = adj ? tr.cxq : mid ? tr.cmq : tr.c1q;
                    const std::vector<double>& c2 =
                        adj ? tr.segs[tr.segs.size() - 2].c : mid ? tr.segs[2].c : tr.segs[1].c;
                    it.c1_off = (uint32_t)P.ccoef.size();
                    for (int k = 0; k < kc1; k++) {
                        if (cq)
                            P.ccoef.push_back(k < (int)u.size() ? u[k] : 0.0f);
                        else
                            P.ccoef.push_back(k < (int)tr.segs[0].c.size() ? (float)tr.segs[0].c[k] : 0.0f);
                        m1 = std::max(m1, fabsf(P.ccoef.back()));
                    }
                    it.c2_off = (uint32_t)P.ccoef.size();
                    for (int l = 0; l < kc2; l++) {
                        P.ccoef.push_back(l < (int)c2.size() ? (float)c2[l] : 0.0f);
                        m2 = std::max(m2, fabsf(P.ccoef.back()));
                    }
                    P.c2items.push_back(it);
                }
                // FP16 coefficient matrix: a power of two taking the largest
                // product |c1| |c2| to [2^13, 2^14] (far below FP16's 65504)
                const double amax = (double)m1 * (double)m2;
                const int e = amax > 0.0 ? std::max(-30, std::min(60, (int)floor(log2(16384.0 / amax)))) : 0;
                ct.ascale = (float)ldexp(1.0, e);
                ct.oscale = (float)ldexp(1.0, -e - 12); // and the basis's 2^12
                P.c2tiles.push_back(ct);
                op.count++;
                op.aux = std::max<uint32_t>(op.aux, (uint32_t)ns);
            }
        }
        // per tile: rows x padded samples, both FP16 parts
        const uint32_t nsp = combine2h_nsp(op.aux);
        for (size_t q = op.off; q < op.off + op.count; q++) {
            Combine2Tile& ct = P.c2tiles[q];
            const double rows = (double)(ct.q0 + ct.q1);
            const uint32_t ns = ct.ns;
            if (B.split) {
                ct.a_off = (uint64_t)P.c2a_halves;
                P.c2a_halves += 2 * (size_t)(ct.q0 + ct.q1) * nsp;
                P.gemm_flops += 3.0 * 2.0 * rows * (double)nsp * (double)N;
            } else {
                P.gemm_flops += 2.0 * rows * 32.0 * (double)((ns + 31) / 32) * (double)N;
            }
            // reads every basis row once, writes both halves of every sample
            P.aux_bytes += (double)N * (4.0 * rows + 8.0 * (double)ns);
            P.gemm_useful_flops += 2.0 * rows * (double)ns * (double)N;
            P.gemm_bytes += (double)N * (4.0 * rows + 8.0 * (double)ns);
        }
        P.ops.push_back(op);
        P.aux_launches++;
    } else {
        (void)n, (void)trajs, (void)grp, (void)B, (void)slots, (void)P;
    }
}

// Groups of trajectories (indices into trajs), each run to completion before the next.
template <typename real>
void build_plan(int n, uint32_t parts, const std::vector<Traj>& trajs, const std::vector<std::vector<size_t>>& groups,
                const std::vector<Problem*>& ps, const real* E, const std::vector<Slot<real>>& slots, double* probs,
                Plan<real>& P, int hi = 0, const Bases* B = nullptr)
{
    const size_t N = (size_t)1 << n;
    for (const auto& grp : groups) {
        if (B) combine_ops<real>(n, trajs, grp, *B, slots, P);
        if (B && !B->W.empty()) combine2_ops<real>(n, trajs, grp, *B, slots, P);
        if (B && !B->U.empty()) combine2_ops<real>(n, trajs, grp, *B, slots, P, nullptr, true);
        std::vector<std::vector<TermJob<real>>> seq(grp.size());
        std::vector<ReduceJob> fins;
        std::vector<AdjointIn> adj(grp.size());
        size_t L = 0;
        for (size_t i = 0; i < grp.size(); i++) {
            const Traj& tr = trajs[grp[i]];
            ReduceJob fin;
            bool has = false;
            traj_jobs<real>(n, parts, tr, slots[i], E + (size_t)tr.prob * N, ps[tr.prob]->info.eloc, probs, seq[i],
                            fin, has, hi, &adj[i]);
            if (has) fins.push_back(fin);
            L = std::max(L, seq[i].size());
            P.terms += seq[i].size();
        }
        for (size_t q = 0; q < L; q++) {
            Op op{0, P.jobs.size(), 0};
            bool any_hsum = false;
            for (size_t i = 0; i < grp.size(); i++)
                if (q < seq[i].size()) {
                    P.jobs.push_back(seq[i][q]);
                    op.count++;
                    P.amp_terms += N;
                    if (seq[i][q].flags & F_PSI_CONST) P.psi_const_amp_terms += N;
                    any_hsum |= (seq[i][q].flags & F_HSUM) != 0;
                }
            if (any_hsum) P.ops.push_back(Op{2, op.off, op.count});
            P.ops.push_back(op);
            P.sweeps++;
        }
        if (B && !B->Z.empty()) {
            combine2_ops<real>(n, trajs, grp, *B, slots, P, &adj);
            pair_dot_ops<real>(n, trajs, grp, adj, probs, P);
        }
        if (B && !B->X.empty()) readout_ops<real>(n, parts, trajs, grp, adj, *B, probs, P);
        if (!fins.empty()) {
            P.ops.push_back(Op{1, P.rjobs.size(), (uint32_t)fins.size()});
            P.rjobs.insert(P.rjobs.end(), fins.begin(), fins.end());
        }
    }
}

template <typename real>
// Runs the plan's ops in order.  Consecutive ops of one class (sweep: term,
// reduce and high-bit pass launches; aux: basis, combine and readout
// launches) form one hipGraph (or direct launches) between two events, so
// the sweep's device time is measured apart from the rest.
mqw_status run_plan(mqw_ctx* c, int precision, int n, const SweepCfg& k, Plan<real>& P, bool use_graph,
                    double* ms_sweep, double* ms_aux, double* ms_basis = nullptr, double* ms_gemm = nullptr,
                    Timed* pending = nullptr, hipStream_t stream = nullptr, const std::string& tag = "plan")
{
    // pending: launch only; the caller synchronises and collects the timings.
    // stream / tag: the stream and the name prefix of the device job buffers
    const hipStream_t stm = stream ? stream : c->stream;
    // the combine, second-stage and readout tables index their coefficients
    // and items with 32-bit offsets
    const size_t lim = 0xffffffffull;
    if (P.ccoef.size() > lim || P.citems.size() > lim || P.c2items.size() > lim || P.ritems.size() > lim)
        return fail(c, MQW_EINVAL, "plan tables exceed 2^32 entries (%zu coefficients); solve fewer problems per call",
                    P.ccoef.size());
    TermJob<real>* jd =
        (TermJob<real>*)dev_get(c, tag + "_jobs", sizeof(TermJob<real>) * std::max<size_t>(1, P.jobs.size()));
    ReduceJob* rd = (ReduceJob*)dev_get(c, tag + "_rjobs", sizeof(ReduceJob) * std::max<size_t>(1, P.rjobs.size()));
    if (!jd || !rd) return fail(c, MQW_ENOMEM, "device allocation failed for %zu sweep jobs", P.jobs.size());
    CombineJob* cd = nullptr;
    CombineItem* ci = nullptr;
    float* cf = nullptr;
    ReadTile* rtd = nullptr;
    ReadItem* rid = nullptr;
    double* dots = nullptr;
    if (!P.ccoef.empty()) {
        cf = (float*)dev_get(c, tag + "_ccoef", sizeof(float) * P.ccoef.size());
        if (!cf) return fail(c, MQW_ENOMEM, "device allocation failed for coefficient rows");
        HIPCHK(c, hipMemcpyAsync(cf, P.ccoef.data(), sizeof(float) * P.ccoef.size(), hipMemcpyHostToDevice,
                                 stm));
    }
    if (!P.cjobs.empty()) {
        cd = (CombineJob*)dev_get(c, tag + "_cjobs", sizeof(CombineJob) * P.cjobs.size());
        ci = (CombineItem*)dev_get(c, tag + "_citems", sizeof(CombineItem) * P.citems.size());
        if (!cd || !ci) return fail(c, MQW_ENOMEM, "device allocation failed for combine jobs");
        HIPCHK(c, hipMemcpyAsync(cd, P.cjobs.data(), sizeof(CombineJob) * P.cjobs.size(), hipMemcpyHostToDevice,
                                 stm));
        HIPCHK(c, hipMemcpyAsync(ci, P.citems.data(), sizeof(CombineItem) * P.citems.size(), hipMemcpyHostToDevice,
                                 stm));
    }
    Combine2Tile* c2t = nullptr;
    Combine2Item* c2i = nullptr;
    uint16_t* c2a = nullptr;
    if (P.c2a_halves) {
        c2a = (uint16_t*)dev_get(c, tag + "_c2a", sizeof(uint16_t) * P.c2a_halves);
        if (!c2a) return fail(c, MQW_ENOMEM, "device allocation failed for FP16 coefficient matrices");
    }
    if (!P.c2tiles.empty()) {
        c2t = (Combine2Tile*)dev_get(c, tag + "_c2tiles", sizeof(Combine2Tile) * P.c2tiles.size());
        c2i = (Combine2Item*)dev_get(c, tag + "_c2items", sizeof(Combine2Item) * P.c2items.size());
        if (!c2t || !c2i) return fail(c, MQW_ENOMEM, "device allocation failed for second-stage tiles");
        HIPCHK(c, hipMemcpyAsync(c2t, P.c2tiles.data(), sizeof(Combine2Tile) * P.c2tiles.size(),
                                 hipMemcpyHostToDevice, stm));
        HIPCHK(c, hipMemcpyAsync(c2i, P.c2items.data(), sizeof(Combine2Item) * P.c2items.size(),
                                 hipMemcpyHostToDevice, stm));
    }
    if (!P.rtiles.empty()) {
        rtd = (ReadTile*)dev_get(c, tag + "_rtiles", sizeof(ReadTile) * P.rtiles.size());
        rid = (ReadItem*)dev_get(c, tag + "_ritems", sizeof(ReadItem) * P.ritems.size());
        dots = (double*)dev_get(c, tag + "_dots", sizeof(double) * P.dots_doubles);
        if (!rtd || !rid || !dots) return fail(c, MQW_ENOMEM, "device allocation failed for readout tiles");
        HIPCHK(c, hipMemcpyAsync(rtd, P.rtiles.data(), sizeof(ReadTile) * P.rtiles.size(), hipMemcpyHostToDevice,
                                 stm));
        HIPCHK(c, hipMemcpyAsync(rid, P.ritems.data(), sizeof(ReadItem) * P.ritems.size(), hipMemcpyHostToDevice,
                                 stm));
    }
    if (!P.jobs.empty())
        HIPCHK(c, hipMemcpyAsync(jd, P.jobs.data(), sizeof(TermJob<real>) * P.jobs.size(), hipMemcpyHostToDevice,
                                 stm));
    ChainJob* chd = nullptr;
    uint32_t *cord = nullptr, *cgo = nullptr, *csync = nullptr;
    if (!P.chains.empty()) {
        chd = (ChainJob*)dev_get(c, tag + "_chains", sizeof(ChainJob) * P.chains.size());
        cord = (uint32_t*)dev_get(c, tag + "_corder", sizeof(uint32_t) * P.corder.size());
        cgo = (uint32_t*)dev_get(c, tag + "_cgoff", sizeof(uint32_t) * P.cgoff.size());
        csync = (uint32_t*)dev_get(c, tag + "_csync", sizeof(uint32_t) * chain_sync_words());
        if (!chd || !cord || !cgo || !csync || !c->chain_err)
            return fail(c, MQW_ENOMEM, "device allocation failed for %zu chains", P.chains.size());
        HIPCHK(c, hipMemcpyAsync(chd, P.chains.data(), sizeof(ChainJob) * P.chains.size(), hipMemcpyHostToDevice,
                                 stm));
        HIPCHK(c, hipMemcpyAsync(cord, P.corder.data(), sizeof(uint32_t) * P.corder.size(), hipMemcpyHostToDevice,
                                 stm));
        HIPCHK(c, hipMemcpyAsync(cgo, P.cgoff.data(), sizeof(uint32_t) * P.cgoff.size(), hipMemcpyHostToDevice,
                                 stm));
    }
    if (!P.rjobs.empty())
        HIPCHK(c, hipMemcpyAsync(rd, P.rjobs.data(), sizeof(ReduceJob) * P.rjobs.size(), hipMemcpyHostToDevice,
                                 stm));
    auto issue = [&](size_t b, size_t e) -> int {
        for (size_t q = b; q < e; q++) {
            const Op& op = P.ops[q];
            int rc;
            if (op.kind == 1)
                rc = launch_reduce(precision, rd + op.off, op.count, stm);
            else if (op.kind == 3)
                rc = launch_combine(cd + op.off, op.count, ci, cf, n, k.parts, op.aux, stm);
            else if (op.kind == 4)
                rc = launch_half_real(k.tb, k.variant, jd + op.off, op.count, n, k.lc, stm);
            else if (op.kind == 9)
                rc = launch_chains(chd, cord, cgo, csync, c->chain_err, n, k.chain_polls, stm);
            else if (op.kind == 5 && op.aux3)
                rc = launch_readout_dots_mfma(rtd + op.off, op.count, rid, cf, n, op.aux2, dots, stm);
            else if (op.kind == 5)
                rc = launch_readout_dots(rtd + op.off, op.count, rid, cf, n, op.aux2, op.aux, dots, stm);
            else if (op.kind == 6)
                rc = launch_readout_finalize(rtd + op.off, op.count, rid, op.aux2, k.parts, dots, stm);
            else if (op.kind == 8)
                rc = launch_pair_dots(rtd + op.off, op.count, rid, n, op.aux2, dots, stm);
            else if (op.kind == 7 && op.aux2)
                rc = launch_combine2h(c2t + op.off, op.count, c2i, cf, c2a, n, k.parts, op.aux, stm);
            else if (op.kind == 7)
                rc = launch_combine2(c2t + op.off, op.count, c2i, cf, n, k.parts, op.aux, stm);
            else if (op.kind == 2)
                rc = launch_high_pass(k.hi, k.hwb, jd + op.off, op.count, n, stm);
            else if (k.nmask > 0)
                rc = launch_mask_term(precision, jd + op.off, op.count, n, k.masks, k.nmask, k.xy, stm);
            else if (k.half)
                rc = launch_half_term(k.tb, k.variant, jd + op.off, op.count, n, k.lc, k.hi, stm);
            else
                rc = launch_term(precision, k.tb, k.variant, jd + op.off, op.count, n, k.lc, stm);
            if (rc) return rc;
        }
        return 0;
    };
    // timing classes: 0 sweep (terms, reduce, high-bit pass), 1 basis recurrences,
    // 2 second-stage GEMM, 3 other auxiliary launches (combine, readout)
    auto op_class = [](int kind) { return kind <= 2 ? 0 : kind == 4 || kind == 9 ? 1 : kind == 7 ? 2 : 3; };
    struct Run {
        size_t b, e;
        int cls;
    };
    std::vector<Run> runs;
    for (size_t q = 0; q < P.ops.size(); q++) {
        const int cl = op_class(P.ops[q].kind);
        if (runs.empty() || runs.back().cls != cl) runs.push_back(Run{q, q, cl});
        runs.back().e = q + 1;
    }
    Timed local;
    Timed& tm = pending ? *pending : local;
    std::deque<EventPair>& evs = tm.evs;
    std::deque<GraphGuard>& graphs = tm.graphs;
    mqw_status st = MQW_OK;
    for (const Run& r : runs) {
        tm.cls.push_back(r.cls);
        evs.emplace_back();
        EventPair& ev = evs.back();
        HIPCHK(c, ev.create());
        HIPCHK(c, hipEventRecord(ev.a, stm));
        if (use_graph) {
            graphs.emplace_back();
            GraphGuard& gr = graphs.back();
            HIPCHK(c, hipStreamBeginCapture(stm, hipStreamCaptureModeThreadLocal));
            const int rc = issue(r.b, r.e);
            hipError_t ce = hipStreamEndCapture(stm, &gr.g);
            if (rc || ce != hipSuccess)
                return fail(c, MQW_EDEVICE, "graph capture failed: %s / %s", last_kernel_error(),
                            hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&gr.ge, gr.g, nullptr, nullptr, 0);
            if (ie != hipSuccess) return fail(c, MQW_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ie));
            hipError_t le = hipGraphLaunch(gr.ge, stm);
            if (le != hipSuccess) {
                st = fail(c, MQW_EDEVICE, "hipGraphLaunch: %s", hipGetErrorString(le));
                break;
            }
        } else if (issue(r.b, r.e)) {
            st = fail(c, MQW_EDEVICE, "launch: %s", last_kernel_error());
            break;
        }
        HIPCHK(c, hipEventRecord(ev.b, stm));
    }
    if (pending) return st; // the caller synchronises and collects
    hipError_t se = hipStreamSynchronize(stm);
    if (se != hipSuccess) return fail(c, MQW_EDEVICE, "stream sync: %s", hipGetErrorString(se));
    if (st) return st;
    collect(tm, ms_sweep, ms_aux, ms_basis, ms_gemm);
    return st;
}

// ---------------------------------------------------------------------------
// Small n: every trajectory runs inside one workgroup (mqw_lds.hip).
// ---------------------------------------------------------------------------
mqw_status run_lds(mqw_ctx* c, int n, const std::vector<Traj>& trajs, const std::vector<Problem*>& ps,
                   const float* E, double* probs_d, bool use_graph)
{
    const size_t N = (size_t)1 << n;
    std::vector<LdsTraj> ht(trajs.size());
    std::vector<LdsSeg> hs;
    std::vector<float> hc;
    uint64_t amp_terms = 0;
    for (size_t t = 0; t < trajs.size(); t++) {
        const Traj& tr = trajs[t];
        ht[t].E = E + (size_t)tr.prob * N;
        ht[t].seg_off = (uint32_t)hs.size();
        ht[t].nseg = (uint32_t)tr.segs.size();
        ht[t].eloc = ps[tr.prob]->info.eloc;
        ht[t].pad = 0;
        for (const Segment& sg : tr.segs) {
            LdsSeg g{};
            g.alpha = (float)(2.0 / (double)sg.onenorm);
            g.beta = (float)(2.0 * (double)sg.gamma / (double)sg.onenorm);
            g.coef_off = (uint32_t)hc.size();
            g.K = (uint32_t)sg.c.size();
            g.prob_slot = (int32_t)sg.prob_slot;
            for (double v : sg.c) hc.push_back((float)v);
            hs.push_back(g);
            amp_terms += (uint64_t)std::max<size_t>(sg.c.size() - 1, 1) * N;
        }
    }
    if (hc.size() > 0xffffffffull) return fail(c, MQW_EINVAL, "coefficient table too large");
    LdsTraj* dt = (LdsTraj*)dev_get(c, "lds_trajs", sizeof(LdsTraj) * ht.size());
    LdsSeg* ds = (LdsSeg*)dev_get(c, "lds_segs", sizeof(LdsSeg) * hs.size());
    float* dc = (float*)dev_get(c, "lds_coefs", sizeof(float) * hc.size());
    if (!dt || !ds || !dc) return fail(c, MQW_ENOMEM, "device allocation failed for LDS trajectory tables");
    HIPCHK(c, hipMemcpyAsync(dt, ht.data(), sizeof(LdsTraj) * ht.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(ds, hs.data(), sizeof(LdsSeg) * hs.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dc, hc.data(), sizeof(float) * hc.size(), hipMemcpyHostToDevice, c->stream));
    const int grid = lds_grid(n, c->device);
    EventPair ev;
    HIPCHK(c, ev.create());
    HIPCHK(c, hipEventRecord(ev.a, c->stream));
    if (use_graph) {
        GraphGuard gr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = launch_lds(n, dt, (uint32_t)ht.size(), ds, dc, probs_d, nullptr, grid, c->stream);
        hipError_t ce = hipStreamEndCapture(c->stream, &gr.g);
        if (rc || ce != hipSuccess)
            return fail(c, MQW_EDEVICE, "graph capture of the LDS kernel failed (%d): %s", rc,
                        hipGetErrorString(ce));
        HIPCHK(c, hipGraphInstantiate(&gr.ge, gr.g, nullptr, nullptr, 0));
        hipError_t le = hipGraphLaunch(gr.ge, c->stream);
        HIPCHK(c, hipEventRecord(ev.b, c->stream));
        hipError_t se = hipStreamSynchronize(c->stream);
        if (le != hipSuccess) return fail(c, MQW_EDEVICE, "hipGraphLaunch: %s", hipGetErrorString(le));
        if (se != hipSuccess) return fail(c, MQW_EDEVICE, "stream sync: %s", hipGetErrorString(se));
    } else {
        if (launch_lds(n, dt, (uint32_t)ht.size(), ds, dc, probs_d, nullptr, grid, c->stream))
            return fail(c, MQW_EDEVICE, "LDS kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
        HIPCHK(c, hipEventRecord(ev.b, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ev.a, ev.b);
    c->stats.gpu_ms += ms;
    c->stats.lds_trajectories += trajs.size();
    c->stats.amp_terms += amp_terms;
    return MQW_OK;
}

// ---------------------------------------------------------------------------
// Shared bases, real halves only (launch_half_real), for every problem with
// bases.kc > 1 or bases.kx > 1; launch q holds step q of every recurrence
// still running:
//   first stage  V_{q+1} = 2 Ht_1 V_q - V_{q-1},  V_1 = Ht_1 |+>
//   last stage   X_{q+1} = 2 Ht_m X_q - X_{q-1},  X_1 = Ht_m e_eloc
// with Ht_i = (H_P - gamma_i H_G) / onenorm_i, the stage's gamma and bound as
// its Clenshaw terms use.
// ---------------------------------------------------------------------------
// With pending, the launches go to `stream` and are not waited for: the
// plans (whose host tables the copies read) move into keep, the events into
// pending, and the caller synchronises and collects the times.
// tag names the device buffers of the plan's tables: work on the main stream
// while a pending plan runs on the second stream must use another.
// A launch of on-chip recurrences left running (run_bases with a ChainDefer),
// so the host can prepare the next plan meanwhile: its plan (whose host tables
// are the sources of its copies) stays here until finish_chains has
// synchronised the stream, which the destructor also does on an early return.
struct ChainDefer {
    bool active = false;
    hipStream_t stm = nullptr;
    Timed tm;
    std::deque<Plan<float>> keep;
    double bytes = 0.0; // algorithmic bytes of the launch
    uint64_t csteps = 0;
    ~ChainDefer()
    {
        if (active) (void)hipStreamSynchronize(stm);
    }
};

// Waits for a deferred launch of on-chip recurrences.  *ok: its outputs stand;
// false: it aborted (a group's workgroups on several XCDs, or a wait that ran
// out), this context runs its recurrences one launch per step from now on,
// and the caller reruns these (they rewrite every row).
mqw_status finish_chains(mqw_ctx* c, ChainDefer& d, bool* ok)
{
    *ok = true;
    if (!d.active) return MQW_OK;
    d.active = false;
    HIPCHK(c, hipStreamSynchronize(d.stm));
    double ms_sweep = 0.0, ms_aux = 0.0, ms_basis = 0.0;
    collect(d.tm, &ms_sweep, &ms_aux, &ms_basis, nullptr);
    d.tm = Timed{};
    d.keep.clear();
    uint32_t ev = 0;
    HIPCHK(c, hipMemcpy(&ev, c->chain_err, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (ev == 0) {
        c->stats.aux_bytes += d.bytes;
        c->stats.basis_bytes += d.bytes;
        c->stats.basis_launches += 1;
        c->stats.aux_launches += 1;
        c->stats.chain_launches += 1;
        c->stats.chain_steps += d.csteps;
        c->stats.gpu_ms += ms_sweep + ms_aux;
        c->stats.aux_ms += ms_sweep + ms_aux;
        c->stats.basis_ms += ms_basis;
        return MQW_OK;
    }
    (void)hipMemset(c->chain_err, 0, sizeof(uint32_t));
    c->chain_capable = false;
    trace("on-chip recurrences aborted (%s): one launch per step from now on",
          (ev & 2u) ? "workgroups of a group on several XCDs" : "a wait ran out");
    *ok = false;
    return MQW_OK;
}

// defer: on-chip recurrences are left running (see ChainDefer); otherwise the
// call returns when its recurrences are done.
mqw_status run_bases(mqw_ctx* c, int n, const SweepCfg& k, const std::vector<Problem*>& ps,
                     const std::vector<Traj>& trajs, const float* E, const Bases& B, bool use_graph,
                     Timed* pending = nullptr, hipStream_t stream = nullptr, std::deque<Plan<float>>* keep = nullptr,
                     const std::string& tag = "bases", ChainDefer* defer = nullptr)
{
    const hipStream_t stm = stream ? stream : c->stream;
    const size_t N = (size_t)1 << n;
    const float psi0 = (float)(1 / sqrt((double)N));
    std::vector<const Segment*> seg0(ps.size(), nullptr), segl(ps.size(), nullptr);
    for (const Traj& tr : trajs)
        if (!seg0[tr.prob]) {
            seg0[tr.prob] = &tr.segs.front();
            segl[tr.prob] = &tr.segs.back();
        }
    double bytes = 0.0;
    // X_0 = e_eloc, ahead of the recurrence on the same stream
    for (size_t p = 0; p < ps.size(); p++)
        if (B.kx[p] > 0) {
            if (launch_fill(const_cast<float*>(B.X[p]), N, 0.0f, ps[p]->info.eloc, 1.0f, stm))
                return fail(c, MQW_EDEVICE, "unit vector: %s", last_kernel_error());
            bytes += 4.0 * N;
        }
    // One plan of recurrence steps.  Step t holds V_{t+1} and X_{t+1} of every
    // instance and, as a wavefront, the second stage's W_{l+1,k} with
    // l + k = t: W_1k needs V_k (made at step k - 1), W_{l+1,k} needs W_lk and
    // W_{l-1,k} (steps t - 1, t - 2), so every launch is wide from the first
    // steps on instead of a narrow V/X phase followed by the W launches.
    //   V_{k+1} = 2 Ht_1 V_k - V_{k-1},  X_{k+1} = 2 Ht_m X_k - X_{k-1}
    //   W_{l+1,k} = 2 Ht_2 W_lk - W_{l-1,k},  W_1k = Ht_2 V_k,  W_0k = V_k (V_0 = |+>)
    // kc2 holds zeros for the instances (or the whole batch) without a shared second stage
    // a compressed instance's first-stage basis Q is ready before the batch:
    // no V recurrence, and every W chain starts at step 0 from its Q_k
    auto cmp = [&](size_t p) { return !B.cmp.empty() && B.cmp[p] != 0; };
    bool has_w = false, need_plus = false;
    for (size_t p = 0; p < B.kc2.size(); p++) {
        has_w |= B.kc2[p] > 0;
        need_plus |= B.kc2[p] > 0 && !cmp(p);
    }
    if (need_plus && !B.plus) return fail(c, MQW_EINVAL, "second-stage bases without the |+> row");
    std::vector<const Segment*> seg1(ps.size(), nullptr), segy(ps.size(), nullptr);
    if (has_w)
        for (const Traj& tr : trajs)
            if (!seg1[tr.prob] && tr.segs.size() > 1) seg1[tr.prob] = &tr.segs[1];
    // two-stage adjoint chains: the operator of stage m - 1
    bool has_y = false;
    for (size_t p = 0; p < B.kz.size(); p++) has_y |= B.kz[p] > 0 && B.ky[p] > 0;
    if (has_y)
        for (const Traj& tr : trajs)
            if (!segy[tr.prob] && tr.segs.size() > 2) segy[tr.prob] = &tr.segs[tr.segs.size() - 2];
    // middle-stage chains: the operator of stage 3
    bool has_m = false;
    std::vector<const Segment*> segm(ps.size(), nullptr);
    for (size_t p = 0; p < B.ku.size(); p++) has_m |= B.ku[p] > 0 && B.k3[p] > 0;
    if (has_m)
        for (const Traj& tr : trajs)
            if (!segm[tr.prob] && tr.segs.size() > 3) segm[tr.prob] = &tr.segs[2];
    int steps = 0;
    for (size_t p = 0; p < ps.size(); p++) {
        if (has_m && B.ku[p] > 0 && B.k3[p] > 0) {
            if (!segm[p]) return fail(c, MQW_EINVAL, "middle-stage basis of an instance with m < 4");
            steps = std::max(steps, B.k3[p] - 1);
        }
        steps = std::max({steps, cmp(p) ? 0 : B.kc[p] - 1, B.kx[p] - 1});
        if (has_w && B.kc2[p] > 0) steps = std::max(steps, (cmp(p) ? 0 : B.kc[p] - 1) + B.kc2[p] - 1);
        if (has_y && B.kz[p] > 0 && B.ky[p] > 0) {
            if (!segy[p]) return fail(c, MQW_EINVAL, "two-stage adjoint basis of an instance with m < 3");
            steps = std::max(steps, B.ky[p] - 1);
        }
    }
    if (need_plus) {
        // |+> as a vector: the GEMM's row (l, k) = (0, 0)
        if (launch_fill(const_cast<float*>(B.plus), N, psi0, N, 0.0f, stm))
            return fail(c, MQW_EDEVICE, "fill: %s", last_kernel_error());
    }
    Plan<float> P;
    // one launch per step where the on-chip recurrences apply sums in their order
    const uint32_t late = n >= kChainMinN && n <= kChainMaxN ? (uint32_t)F_LATE : 0u;
    auto job = [&](size_t p, const Segment& sg, int q) {
        TermJob<float> j{};
        j.E = E + p * N;
        j.psi0 = psi0;
        j.w1r = 1.0f;
        const double sc = (double)sg.onenorm * (q == 0 ? 1.0 : 0.5);
        j.alpha = (float)(1.0 / sc);
        j.beta = (float)((double)sg.gamma / sc);
        return j;
    };
    // On-chip recurrences (launch_chains): every chain runs all its steps in
    // one launch, so every chain must start from a vector ready before it
    // (no uncompressed W chain, which waits for its V_k).  The same rows, steps
    // and scalars as the per-step plan below, and bit-identical outputs.
    // (c->chain_capable: not after a launch of this context aborted)
    bool use_chains = k.chains && c->chain_capable && !pending;
    for (size_t p = 0; use_chains && p < ps.size(); p++)
        if (has_w && B.kc2[p] > 0 && !cmp(p)) use_chains = false;
    if (use_chains) {
        double cbytes = 0.0;
        uint64_t csteps = 0;
        auto add = [&](size_t p, const Segment& sg, const float* seed, float* out, int nsteps) {
            if (nsteps <= 0) return;
            const TermJob<float> j0 = job(p, sg, 0), j1 = job(p, sg, 1);
            // the kernel takes alpha0 E as (alpha E) / 2
            if (j1.alpha != 2.0f * j0.alpha) use_chains = false;
            ChainJob cj{};
            cj.E = E + p * N;
            cj.seed = seed;
            cj.out = out;
            cj.alpha0 = j0.alpha;
            cj.beta0 = j0.beta;
            cj.alpha = j1.alpha;
            cj.beta = j1.beta;
            cj.a1 = 1.0f;
            cj.a2 = 1.0f;
            cj.psi0 = psi0;
            cj.steps = (uint32_t)nsteps;
            P.chains.push_back(cj);
            // HBM: the diagonal and the seed read once, each output written
            // (s_{k-1} comes back from LDS, the partners from the XCD's L2)
            cbytes += 4.0 * N * (1.0 + (seed ? 1.0 : 0.0) + nsteps);
            csteps += (uint64_t)nsteps;
        };
        for (size_t p = 0; p < ps.size(); p++) {
            if (!cmp(p) && B.kc[p] > 1) add(p, *seg0[p], nullptr, const_cast<float*>(B.V[p]), B.kc[p] - 1);
            if (B.kx[p] > 1)
                add(p, *segl[p], B.X[p], const_cast<float*>(B.X[p]) + N, B.kx[p] - 1);
            const int kc2 = has_w ? B.kc2[p] : 0;
            if (kc2 > 0) // compressed: Q_0 at plusv, Q_kk = V_kk
                for (int kk = 0; kk < B.kc[p]; kk++)
                    add(p, *seg1[p], kk > 0 ? B.V[p] + (size_t)(kk - 1) * N : B.plusv[p],
                        const_cast<float*>(B.W[p]) + (size_t)kk * (kc2 - 1) * N, kc2 - 1);
            if (has_y && B.kz[p] > 0 && B.ky[p] > 0)
                for (int kk = 0; kk < B.kz[p]; kk++)
                    add(p, *segy[p], B.Z[p] + (size_t)kk * N,
                        const_cast<float*>(B.Y[p]) + (size_t)kk * (B.ky[p] - 1) * N, B.ky[p] - 1);
            if (has_m && B.ku[p] > 0 && B.k3[p] > 0)
                for (int kk = 0; kk < B.ku[p]; kk++)
                    add(p, *segm[p], B.U[p] + (size_t)kk * N,
                        const_cast<float*>(B.W3[p]) + (size_t)kk * (B.k3[p] - 1) * N, B.k3[p] - 1);
        }
        if (use_chains && !P.chains.empty()) {
            // longest chains first, each to the group with the fewest steps so far
            const uint32_t ngroups = (uint32_t)kChainWorkgroups >> (n - kChainMinN);
            std::vector<uint32_t> idx(P.chains.size());
            for (uint32_t i = 0; i < idx.size(); i++) idx[i] = i;
            std::stable_sort(idx.begin(), idx.end(),
                             [&](uint32_t a, uint32_t b) { return P.chains[a].steps > P.chains[b].steps; });
            std::vector<uint64_t> load(ngroups, 0);
            std::vector<std::vector<uint32_t>> lists(ngroups);
            for (uint32_t i : idx) {
                const uint32_t g = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
                load[g] += P.chains[i].steps;
                lists[g].push_back(i);
            }
            P.cgoff.assign(1, 0);
            for (uint32_t g = 0; g < ngroups; g++) {
                P.corder.insert(P.corder.end(), lists[g].begin(), lists[g].end());
                P.cgoff.push_back((uint32_t)P.corder.size());
            }
            P.ops.push_back(Op{9, 0, (uint32_t)P.chains.size()});
            double ms_sweep = 0.0, ms_aux = 0.0, ms_basis = 0.0;
            if (defer) {
                defer->keep.push_back(std::move(P));
                defer->active = true; // (from here the destructor waits for the stream)
                defer->stm = stm;
                defer->bytes = bytes + cbytes;
                defer->csteps = csteps;
                return run_plan<float>(c, MQW_FP32, n, k, defer->keep.back(), use_graph, &ms_sweep, &ms_aux, &ms_basis,
                                       nullptr, &defer->tm, stm, tag);
            }
            mqw_status st = run_plan<float>(c, MQW_FP32, n, k, P, use_graph, &ms_sweep, &ms_aux, &ms_basis, nullptr,
                                            nullptr, stm, tag);
            if (st) return st;
            uint32_t ev = 0;
            HIPCHK(c, hipMemcpy(&ev, c->chain_err, sizeof(uint32_t), hipMemcpyDeviceToHost));
            if (ev == 0) {
                c->stats.aux_bytes += bytes + cbytes;
                c->stats.basis_bytes += bytes + cbytes;
                c->stats.basis_launches += 1;
                c->stats.aux_launches += 1;
                c->stats.chain_launches += 1;
                c->stats.chain_steps += csteps;
                c->stats.gpu_ms += ms_sweep + ms_aux;
                c->stats.aux_ms += ms_sweep + ms_aux;
                c->stats.basis_ms += ms_basis;
                return MQW_OK;
            }
            // a group's workgroups on several XCDs, or a wait that ran out: this
            // context runs its recurrences one launch per step from now on,
            // starting with these (they rewrite every row)
            (void)hipMemset(c->chain_err, 0, sizeof(uint32_t));
            c->chain_capable = false;
            trace("on-chip recurrences aborted (%s): one launch per step from now on",
                  (ev & 2u) ? "workgroups of a group on several XCDs" : "a wait ran out");
            SweepCfg k1 = k;
            k1.chains = false;
            return run_bases(c, n, k1, ps, trajs, E, B, use_graph, pending, stream, keep, tag);
        }
        P.chains.clear();
    }
    // Chains from seeds: row (l, kk) of rows at W + (kk (kc2 - 1) + l - 1) N for
    // l >= 1, row (0, kk) = seed(kk) (nullptr: the constant psi0); step q holds
    // l = q - kk, or l = q when every seed is ready before the batch (cq).
    Op op{4, 0, 0};
    auto chains = [&](size_t p, float* W, int kc1, int kc2, const Segment& sg, bool cq, int q, auto seed) {
        auto row = [&](int ll, int kk) { return W + ((size_t)kk * (kc2 - 1) + (size_t)(ll - 1)) * N; };
        for (int kk = 0; kk < kc1 && (cq || kk <= q); kk++) {
            const int l = cq ? q : q - kk;
            if (l >= kc2 - 1) continue;
            TermJob<float> j = job(p, sg, l);
            j.out = row(l + 1, kk);
            uint32_t flags = 0;
            double b = 8.0; // E read, output write
            if (l == 0) {
                if (!seed(kk)) {
                    flags |= SRC1_PSI | F_PSI_CONST;
                } else {
                    flags |= SRC1_ARRAY;
                    j.b1 = seed(kk);
                    b += 4.0;
                }
                flags |= SRC2_ZERO;
            } else {
                flags |= SRC1_ARRAY;
                j.b1 = row(l, kk);
                j.w2r = 1.0f;
                b += 4.0;
                if (l == 1) {
                    if (!seed(kk)) {
                        flags |= SRC2_PSI | F_PSI_CONST;
                    } else {
                        flags |= SRC2_ARRAY;
                        j.b2 = seed(kk);
                        b += 4.0;
                    }
                } else {
                    flags |= SRC2_ARRAY;
                    j.b2 = row(l - 1, kk);
                    b += 4.0;
                }
            }
            j.flags = flags | late;
            P.jobs.push_back(j);
            op.count++;
            bytes += b * N;
        }
    };
    for (int q = 0; q < steps; q++) {
        op = Op{4, P.jobs.size(), 0};
        for (size_t p = 0; p < ps.size(); p++) {
            if (!cmp(p) && q < B.kc[p] - 1) {
                float* V = const_cast<float*>(B.V[p]); // V_1 .. V_{kc-1}
                TermJob<float> j = job(p, *seg0[p], q);
                j.out = V + (size_t)q * N;
                uint32_t flags = F_PSI_CONST;
                if (q == 0) {
                    flags |= SRC1_PSI | SRC2_ZERO;
                    bytes += 8.0 * N;
                } else {
                    flags |= SRC1_ARRAY;
                    j.b1 = V + (size_t)(q - 1) * N;
                    j.w2r = 1.0f;
                    if (q == 1) {
                        flags |= SRC2_PSI;
                        bytes += 12.0 * N;
                    } else {
                        flags |= SRC2_ARRAY;
                        j.b2 = V + (size_t)(q - 2) * N;
                        bytes += 16.0 * N;
                    }
                }
                j.flags = flags | late;
                P.jobs.push_back(j);
                op.count++;
            }
            if (q < B.kx[p] - 1) {
                float* X = const_cast<float*>(B.X[p]); // X_0 .. X_{kx-1}
                TermJob<float> j = job(p, *segl[p], q);
                j.out = X + (size_t)(q + 1) * N;
                j.b1 = X + (size_t)q * N;
                uint32_t flags = SRC1_ARRAY;
                if (q == 0) {
                    flags |= SRC2_ZERO;
                    bytes += 12.0 * N;
                } else {
                    flags |= SRC2_ARRAY;
                    j.b2 = X + (size_t)(q - 1) * N;
                    j.w2r = 1.0f;
                    bytes += 16.0 * N;
                }
                j.flags = flags | late;
                P.jobs.push_back(j);
                op.count++;
            }
            const int kc2 = has_w ? B.kc2[p] : 0;
            if (kc2 > 0) {
                const float* V = B.V[p];
                const bool cq = cmp(p);
                // W_0kk: |+> (nullptr: the constant psi0) for kk = 0, else V_kk;
                // compressed: Q_0 at plusv, Q_kk = V_kk
                chains(p, const_cast<float*>(B.W[p]), B.kc[p], kc2, *seg1[p], cq, q, [&](int kk) -> const float* {
                    return kk > 0 ? V + (size_t)(kk - 1) * N : cq ? B.plusv[p] : nullptr;
                });
            }
            // Y_{l+1,r} = 2 Ht_{m-1} Y_lr - Y_{l-1,r} from Y_0r = Z_r (ready before the batch)
            if (has_y && B.kz[p] > 0 && B.ky[p] > 0) {
                const float* Z = B.Z[p];
                chains(p, const_cast<float*>(B.Y[p]), B.kz[p], B.ky[p], *segy[p], true, q,
                       [&](int kk) -> const float* { return Z + (size_t)kk * N; });
            }
            // W3_{l+1,r} = 2 Ht_3 W3_lr - W3_{l-1,r} from W3_0r = U_r
            if (has_m && B.ku[p] > 0 && B.k3[p] > 0) {
                const float* U = B.U[p];
                chains(p, const_cast<float*>(B.W3[p]), B.ku[p], B.k3[p], *segm[p], true, q,
                       [&](int kk) -> const float* { return U + (size_t)kk * N; });
            }
        }
        P.ops.push_back(op);
    }
    c->stats.aux_bytes += bytes;
    c->stats.basis_bytes += bytes;
    if (P.jobs.empty()) return MQW_OK;
    c->stats.basis_launches += P.ops.size();
    c->stats.aux_launches += P.ops.size();
    if (pending) {
        mqw_status st = run_plan<float>(c, MQW_FP32, n, k, P, use_graph, nullptr, nullptr, nullptr, nullptr, pending,
                                        stm, tag);
        keep->push_back(std::move(P)); // the copies read its host tables
        return st;
    }
    double ms_sweep = 0.0, ms_aux = 0.0, ms_basis = 0.0; // every op here is a basis launch
    mqw_status st = run_plan<float>(c, MQW_FP32, n, k, P, use_graph, &ms_sweep, &ms_aux, &ms_basis, nullptr, nullptr,
                                    stm, tag);
    if (st) return st;
    c->stats.gpu_ms += ms_sweep + ms_aux;
    c->stats.aux_ms += ms_sweep + ms_aux;
    c->stats.basis_ms += ms_basis;
    return MQW_OK;
}

// ---------------------------------------------------------------------------
// m == 1 instances evaluated directly.  The reference chains its samples: it
// sorts the m == 1 stage times, differences them, and propagates each sample's
// state from the previous sample's by the increment (MultiQW.cpp:426-430,447).
// Sample j's state is therefore exp(iH T_j)|+> with T_j the sum of the first
// j + 1 increments, up to each step's truncation (its tail, <= tol) and
// renormalisation.  Here it is evaluated as one series,
//     psi_j = sum_k a_jk V_k,  V_k = T_k(Ht)|+>,  a_jk = c_k(onenorm T_j) i^(k&1),
// from the instance's first-stage basis V: the amplitude is
// sum_k a_jk V_k[E_loc] and the squared norm c_e^T G_e c_e + c_o^T G_o c_o with
// G the Gram matrix per parity class (gram_kernel, FP64).  No term sweep runs.
// In chunks of instances whose V fit in vcap bytes; the probabilities go to
// probs (host) and probs_d.
// ---------------------------------------------------------------------------
mqw_status direct_m1(mqw_ctx* c, int n, const SweepCfg& k, const std::vector<Problem*>& ps, std::vector<Traj>& trajs,
                     const float* E, double tol, bool use_graph, size_t vcap, std::vector<double>& probs,
                     double* probs_d)
{
    const size_t N = (size_t)1 << n;
    const float psi0 = (float)(1 / sqrt((double)N));
    struct Chain {
        size_t t;                           // its trajectory
        int kc = 0;
        std::vector<std::vector<double>> cf; // per sample: coefficients at onenorm * T_j (float values)
    };
    std::vector<Chain> ch;
    for (size_t t = 0; t < trajs.size(); t++) {
        const Traj& tr = trajs[t];
        if (ps[tr.prob]->m != 1 || tr.segs.empty()) continue;
        Chain q;
        q.t = t;
        q.cf.resize(tr.segs.size());
        double T = 0.0;
        const double onenorm = (double)tr.segs.front().onenorm;
        for (size_t j = 0; j < tr.segs.size(); j++) {
            T += (double)ps[tr.prob]->times[j];
            cheb_coeffs_vec(onenorm * T, tol, q.cf[j]);
            for (double& v : q.cf[j]) v = (double)(float)v; // the reference casts to float (:440)
            q.kc = std::max(q.kc, (int)q.cf[j].size());
        }
        if (q.kc >= 2 && q.kc <= kGramMaxK) ch.push_back(std::move(q));
    }
    if (ch.empty()) return MQW_OK;
    size_t i0 = 0;
    while (i0 < ch.size()) {
        size_t i1 = i0, vbytes = 0;
        while (i1 < ch.size()) {
            const size_t b = sizeof(float) * N * (size_t)(ch[i1].kc - 1);
            if (i1 > i0 && vbytes + b > vcap) break;
            vbytes += b;
            i1++;
        }
        char* vbuf = (char*)dev_get(c, "cmp_V", vbytes);
        if (!vbuf) return fail(c, MQW_ENOMEM, "device allocation failed for first-stage bases");
        c->stats.first_stage_chunks++;
        Bases vb;
        vb.V.assign(ps.size(), nullptr);
        vb.kc.assign(ps.size(), 0);
        vb.X.assign(ps.size(), nullptr);
        vb.kx.assign(ps.size(), 0);
        vb.kc2.assign(ps.size(), 0);
        size_t off = 0, pairs = 0;
        uint32_t kcmax = 0;
        std::vector<size_t> goff(i1 - i0);
        for (size_t i = i0; i < i1; i++) {
            const size_t p = (size_t)trajs[ch[i].t].prob;
            vb.kc[p] = ch[i].kc;
            vb.V[p] = (const float*)(vbuf + off);
            off += sizeof(float) * N * (size_t)(ch[i].kc - 1);
            goff[i - i0] = pairs;
            pairs += gram_pairs((uint32_t)ch[i].kc);
            kcmax = std::max<uint32_t>(kcmax, (uint32_t)ch[i].kc);
        }
        mqw_status st = run_bases(c, n, k, ps, trajs, E, vb, use_graph);
        if (st) return st;
        const size_t njobs = i1 - i0;
        const size_t pmax = gram_pmax(kcmax);
        const uint32_t nchunk = gram_nchunk(n);
        double* gram_d = (double*)dev_get(c, "cmp_gram", sizeof(double) * pairs);
        double* part_d = (double*)dev_get(c, "cmp_gram_part", sizeof(double) * njobs * nchunk * pmax);
        GramJob* gj_d = (GramJob*)dev_get(c, "cmp_gram_jobs", sizeof(GramJob) * njobs);
        if (!gram_d || !part_d || !gj_d) return fail(c, MQW_ENOMEM, "device allocation failed for Gram matrices");
        std::vector<GramJob> gj(njobs);
        for (size_t i = i0; i < i1; i++) {
            const size_t p = (size_t)trajs[ch[i].t].prob;
            gj[i - i0] = GramJob{vb.V[p], nullptr, gram_d + goff[i - i0], (uint32_t)ch[i].kc, psi0, 0u, 0u};
        }
        EventPair ev;
        HIPCHK(c, ev.create());
        HIPCHK(c, hipEventRecord(ev.a, c->stream));
        HIPCHK(c, hipMemcpyAsync(gj_d, gj.data(), sizeof(GramJob) * njobs, hipMemcpyHostToDevice, c->stream));
        if (launch_gram(gj_d, (uint32_t)njobs, kcmax, n, nchunk, part_d, c->stream))
            return fail(c, MQW_EDEVICE, "gram: %s", last_kernel_error());
        HIPCHK(c, hipEventRecord(ev.b, c->stream));
        std::vector<double> gram(pairs);
        HIPCHK(c, hipMemcpyAsync(gram.data(), gram_d, sizeof(double) * pairs, hipMemcpyDeviceToHost, c->stream));
        // V_k[E_loc], k >= 1: one float per vector, N floats apart
        std::vector<std::vector<float>> ve(njobs);
        for (size_t i = i0; i < i1; i++) {
            const size_t p = (size_t)trajs[ch[i].t].prob;
            std::vector<float>& v = ve[i - i0];
            v.assign((size_t)ch[i].kc, psi0);
            HIPCHK(c, hipMemcpy2DAsync(v.data() + 1, sizeof(float), vb.V[p] + ps[p]->info.eloc, sizeof(float) * N,
                                       sizeof(float), (size_t)(ch[i].kc - 1), hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream)); // cmp_V is reused by the next chunk
        float ms = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&ms, ev.a, ev.b));
        c->stats.gpu_ms += ms;
        c->stats.aux_ms += ms;
        c->stats.aux_launches += 2;
        for (size_t i = i0; i < i1; i++) {
            Chain& q = ch[i];
            Traj& tr = trajs[q.t];
            const Problem& pr = *ps[tr.prob];
            const double* g = &gram[goff[i - i0]];
            const std::vector<float>& v = ve[i - i0];
            c->stats.aux_bytes += 4.0 * (double)N * (double)(q.kc - 1); // the Gram read of V
            for (size_t j = 0; j < q.cf.size(); j++) {
                const std::vector<double>& cf = q.cf[j];
                const int K = (int)cf.size();
                double re = 0.0, im = 0.0, nrm = 0.0;
                for (int k2 = 0; k2 < K; k2++) {
                    ((k2 & 1) ? im : re) += cf[k2] * (double)v[k2];
                    const uint32_t b = gram_pair_start((uint32_t)k2);
                    for (int k1 = k2 & 1; k1 <= k2; k1 += 2)
                        nrm += (k1 == k2 ? 1.0 : 2.0) * cf[k1] * cf[k2] * g[b + (uint32_t)k1 / 2];
                }
                probs[pr.prob_base + j] = (re * re + im * im) / nrm;
            }
            HIPCHK(c, hipMemcpyAsync(probs_d + pr.prob_base, &probs[pr.prob_base], sizeof(double) * q.cf.size(),
                                     hipMemcpyHostToDevice, c->stream));
            tr.direct_done = true;
            c->stats.direct_chains++;
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        i0 = i1;
    }
    return MQW_OK;
}

// ---------------------------------------------------------------------------
// Compression of a shared basis (see GramJob): a basis of kc real vectors
// (V_0 stored at V0, else psi0 * ones or the unit vector at unit1 - 1; V_1 ..
// at V, N floats apart) and the samples' coefficient rows over it.  The Gram
// pair sums run on the device, one launch per size class of the kernel (a
// wide basis would otherwise set the pair count and occupancy of every job);
// one SVD per job and parity class on the host, on up to 16 threads
// (compress_first_stage); then combine launches write the kept directions:
// item r holds the even class's direction 2r in its real half and the odd
// class's 2r + 1 in its imaginary half, so the 2R vectors lie in index order,
// N floats apart.  A basis wider than kCombineMaxKX takes two combine jobs,
// the second adding its vectors 127.. to the first's outputs.
// ---------------------------------------------------------------------------
// per-sample relative error of the kept span: a quarter of the FP32 rounding of
// the basis itself (7.4e-7 relative, measured on bench instances)
constexpr double kCompressTol = 2e-7;

struct Compression {
    int kc = 0;
    const float* V = nullptr;
    const float* V0 = nullptr;
    float psi0 = 0.0f;
    uint32_t unit1 = 0;
    size_t S = 0;
    std::vector<double> rows; // S x kc coefficients
    size_t goff = 0;          // its pair sums in the Gram array
    bool ok = false;
    FirstStageCompression fc;
};

// Gram pair sums and SVDs of the jobs, on c->stream (synchronised on return).
mqw_status compress_svd(mqw_ctx* c, int n, std::vector<Compression>& cj, const std::string& tag)
{
    const size_t N = (size_t)1 << n;
    const size_t njobs = cj.size();
    if (njobs == 0) return MQW_OK;
    const uint32_t nchunk = gram_nchunk(n);
    size_t pairs = 0;
    uint32_t kcmax = 0;
    for (Compression& j : cj) {
        j.goff = pairs;
        pairs += gram_pairs((uint32_t)j.kc);
        kcmax = std::max<uint32_t>(kcmax, (uint32_t)j.kc);
    }
    const size_t pmax = gram_pmax(kcmax);
    double* gram_d = (double*)dev_get(c, tag + "_gram", sizeof(double) * pairs);
    double* part_d = (double*)dev_get(c, tag + "_gram_part", sizeof(double) * njobs * nchunk * pmax);
    GramJob* gj_d = (GramJob*)dev_get(c, tag + "_gram_jobs", sizeof(GramJob) * njobs);
    if (!gram_d || !part_d || !gj_d) return fail(c, MQW_ENOMEM, "device allocation failed for Gram matrices");
    auto gclass = [](int kc) { return kc <= 32 ? 0 : kc <= 64 ? 1 : kc <= 128 ? 2 : 3; };
    std::vector<GramJob> gj;
    std::vector<size_t> cls_end(4, 0);
    std::vector<uint32_t> cls_kc(4, 0);
    for (int g = 0; g < 4; g++) {
        for (const Compression& j : cj)
            if (gclass(j.kc) == g) {
                gj.push_back(GramJob{j.V, j.V0, gram_d + j.goff, (uint32_t)j.kc, j.psi0, j.unit1, 0u});
                cls_kc[g] = std::max<uint32_t>(cls_kc[g], (uint32_t)j.kc);
            }
        cls_end[g] = gj.size();
    }
    EventPair ev;
    HIPCHK(c, ev.create());
    HIPCHK(c, hipEventRecord(ev.a, c->stream));
    HIPCHK(c, hipMemcpyAsync(gj_d, gj.data(), sizeof(GramJob) * njobs, hipMemcpyHostToDevice, c->stream));
    for (int g = 0, a = 0; g < 4; a = (int)cls_end[g], g++)
        if (cls_end[g] > (size_t)a &&
            launch_gram(gj_d + a, (uint32_t)(cls_end[g] - a), cls_kc[g], n, nchunk, part_d, c->stream))
            return fail(c, MQW_EDEVICE, "gram: %s", last_kernel_error());
    HIPCHK(c, hipEventRecord(ev.b, c->stream));
    std::vector<double> gram(pairs);
    HIPCHK(c, hipMemcpyAsync(gram.data(), gram_d, sizeof(double) * pairs, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, ev.a, ev.b));
    c->stats.gpu_ms += ms;
    c->stats.aux_ms += ms;
    c->stats.aux_launches += 2 * (uint64_t)std::count_if(cls_end.begin(), cls_end.end(), [&](size_t e) {
        return e > 0;
    });
    for (const Compression& j : cj) c->stats.aux_bytes += 4.0 * (double)N * (double)(j.kc - (j.V0 ? 0 : 1));
    trace("compression: %zu Gram matrices", njobs);
    auto work = [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) {
            Compression& j = cj[i];
            std::vector<const double*> cp(j.S);
            for (size_t s = 0; s < j.S; s++) cp[s] = &j.rows[s * (size_t)j.kc];
            j.ok = compress_first_stage(j.kc, &gram[j.goff], cp, kCompressTol, j.fc);
        }
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nt = std::min<size_t>({(size_t)(hw ? hw : 1), 16, njobs});
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; t++) th.emplace_back(work, njobs * t / nt, njobs * (t + 1) / nt);
    work(0, njobs / nt);
    for (std::thread& x : th) x.join();
    trace("compression: %zu SVDs", njobs);
    return MQW_OK;
}

// The kept directions of every successful job into out[i] (2R vectors of N
// floats), on c->stream (not synchronised).
mqw_status compress_combine(mqw_ctx* c, int n, const SweepCfg& k, const std::vector<Compression>& cj,
                            const std::vector<float*>& out, const std::string& tag)
{
    const size_t N = (size_t)1 << n;
    size_t nitems = 0;
    uint32_t kcm = 0;
    for (const Compression& j : cj)
        if (j.ok) {
            nitems += (size_t)j.fc.R;
            kcm = std::max<uint32_t>(kcm, (uint32_t)j.kc);
        }
    if (nitems == 0) return MQW_OK;
    double* dummy = (double*)dev_get(c, tag + "_norm", sizeof(double) * k.parts * nitems);
    CombineJob* cj_d = (CombineJob*)dev_get(c, tag + "_cjobs", sizeof(CombineJob) * cj.size());
    CombineJob* cj2_d = (CombineJob*)dev_get(c, tag + "_cjobs2", sizeof(CombineJob) * cj.size());
    CombineItem* ci_d = (CombineItem*)dev_get(c, tag + "_citems", sizeof(CombineItem) * nitems);
    CombineItem* ci2_d = (CombineItem*)dev_get(c, tag + "_citems2", sizeof(CombineItem) * nitems);
    float* cf_d = (float*)dev_get(c, tag + "_coefs", sizeof(float) * nitems * (kcm + 2));
    if (!dummy || !cj_d || !cj2_d || !ci_d || !ci2_d || !cf_d)
        return fail(c, MQW_ENOMEM, "device allocation failed for compressed bases");
    // the second job of a wide basis: its vectors 0 and 1 (basis vectors 126
    // and 127, in the first job) have no weight, so the parity classes stay
    std::vector<CombineJob> jb, jb2;
    std::vector<CombineItem> ci, ci2;
    std::vector<float> cf;
    // These tables are copied to the device from pageable memory, which the
    // runtime may read when it carries the copies out, after this call:
    // the stream is drained on every way out, before they are freed.
    struct Drain {
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{c->stream};
    uint32_t kcb_max = 0;
    for (size_t i = 0; i < cj.size(); i++) {
        const Compression& j = cj[i];
        if (!j.ok) continue;
        const int kc = j.kc, R = j.fc.R;
        const int kca = std::min(kc, kCombineMaxKX), kcb = kc > kCombineMaxKX ? kc - (kCombineMaxKX - 2) : 0;
        jb.push_back(CombineJob{j.V, j.V0, (uint32_t)kca, (uint32_t)R, (uint32_t)ci.size(), j.psi0, j.unit1, 0u});
        if (kcb > 0) {
            jb2.push_back(CombineJob{j.V + (size_t)(kCombineMaxKX - 2) * N, nullptr, (uint32_t)kcb, (uint32_t)R,
                                     (uint32_t)ci2.size(), 0.0f, 0u, 1u});
            kcb_max = std::max<uint32_t>(kcb_max, (uint32_t)kcb);
        }
        const double* B = j.fc.B.data();
        for (int r = 0; r < R; r++) {
            CombineItem it{};
            it.out = out[i] + (size_t)2 * r * N;
            it.norm_part = dummy + ci.size() * k.parts;
            it.coef_off = (uint32_t)cf.size();
            for (int q = 0; q < kca; q++) cf.push_back((float)B[(size_t)(2 * r + (q & 1)) * kc + q]);
            ci.push_back(it);
            if (kcb > 0) {
                it.coef_off = (uint32_t)cf.size();
                for (int q = 0; q < kcb; q++) {
                    const int kk = q + kCombineMaxKX - 2; // basis index of column q
                    cf.push_back(q < 2 ? 0.0f : (float)B[(size_t)(2 * r + (kk & 1)) * kc + kk]);
                }
                ci2.push_back(it);
            }
        }
        c->stats.aux_bytes += (double)N * (4.0 * (kc - (j.V0 ? 0 : 1)) + 8.0 * R);
    }
    HIPCHK(c, hipMemcpyAsync(cj_d, jb.data(), sizeof(CombineJob) * jb.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(ci_d, ci.data(), sizeof(CombineItem) * ci.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(cf_d, cf.data(), sizeof(float) * cf.size(), hipMemcpyHostToDevice, c->stream));
    if (launch_combine(cj_d, (uint32_t)jb.size(), ci_d, cf_d, n, k.parts, std::min<uint32_t>(kcm, kCombineMaxKX),
                       c->stream))
        return fail(c, MQW_EDEVICE, "basis compression: %s", last_kernel_error());
    c->stats.aux_launches++;
    if (!jb2.empty()) {
        HIPCHK(c, hipMemcpyAsync(cj2_d, jb2.data(), sizeof(CombineJob) * jb2.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(ci2_d, ci2.data(), sizeof(CombineItem) * ci2.size(), hipMemcpyHostToDevice,
                                 c->stream));
        if (launch_combine(cj2_d, (uint32_t)jb2.size(), ci2_d, cf_d, n, k.parts, kcb_max, c->stream))
            return fail(c, MQW_EDEVICE, "basis compression: %s", last_kernel_error());
        c->stats.aux_launches++;
    }
    return MQW_OK;
}

// ---------------------------------------------------------------------------
// Compressed bases computed before the batches, in chunks of instances whose
// bases fit in vcap bytes:
// - the first-stage basis V, for the instances whose second stage comes from
//   the shared basis W (kc2 > 0);
// - the last-stage adjoint basis X (X_0 = e_eloc), for the instances whose
//   last two stages are read out by the two-stage adjoint (kxa > 0): the
//   samples' adjoint states phi_j = sum_k a_jk X_k span few directions too.
// The recurrences of every instance of a chunk run in one plan, then the
// compressions above.  kcq[p] = 2R and qbase[p] = Q_0 for V (every
// trajectory gets its coordinates in c1q), kzq[p] = 2R and zbase[p] = Z_0 for
// X (coordinates in cxq); 0 = not compressed.
// ---------------------------------------------------------------------------
mqw_status compress_bases(mqw_ctx* c, int n, const SweepCfg& k, const std::vector<Problem*>& ps,
                          std::vector<Traj>& trajs, const float* E, const std::vector<int>& kc1,
                          const std::vector<int>& kc2, const std::vector<int>& kxa, bool use_graph, size_t vcap,
                          std::vector<int>& kcq, std::vector<const float*>& qbase, std::vector<int>& kzq,
                          std::vector<const float*>& zbase)
{
    const size_t N = (size_t)1 << n;
    const float psi0 = (float)(1 / sqrt((double)N));
    kcq.assign(ps.size(), 0);
    qbase.assign(ps.size(), nullptr);
    kzq.assign(ps.size(), 0);
    zbase.assign(ps.size(), nullptr);
    auto fwd = [&](size_t p) { return kc2[p] > 0 && kc1[p] >= 8; };
    auto adj = [&](size_t p) { return kxa[p] >= 8 && kxa[p] <= kAdjMaxK && kxa[p] <= kGramMaxK; };
    std::vector<size_t> cand;
    for (size_t p = 0; p < ps.size(); p++)
        if (fwd(p) || adj(p)) cand.push_back(p);
    if (cand.empty()) return MQW_OK;
    std::vector<std::vector<size_t>> ptraj(ps.size());
    for (size_t t = 0; t < trajs.size(); t++) ptraj[trajs[t].prob].push_back(t);
    auto pbytes = [&](size_t p) {
        return sizeof(float) * N * (size_t)((fwd(p) ? kc1[p] - 1 : 0) + (adj(p) ? kxa[p] : 0));
    };
    size_t i0 = 0;
    int chunk_id = 0;
    while (i0 < cand.size()) {
        size_t i1 = i0, vbytes = 0;
        while (i1 < cand.size()) {
            const size_t b = pbytes(cand[i1]);
            if (i1 > i0 && vbytes + b > vcap) break;
            vbytes += b;
            i1++;
        }
        char* vbuf = (char*)dev_get(c, "cmp_V", vbytes);
        if (!vbuf) return fail(c, MQW_ENOMEM, "device allocation failed for first-stage bases");
        c->stats.first_stage_chunks++;
        // V and X recurrences of the chunk (no W)
        Bases vb;
        vb.V.assign(ps.size(), nullptr);
        vb.kc.assign(ps.size(), 0);
        vb.X.assign(ps.size(), nullptr);
        vb.kx.assign(ps.size(), 0);
        vb.kc2.assign(ps.size(), 0);
        std::vector<Compression> cj;
        std::vector<std::pair<size_t, int>> who; // (instance, 0 = V / 1 = X)
        size_t off = 0;
        for (size_t i = i0; i < i1; i++) {
            const size_t p = cand[i];
            for (int which = 0; which < 2; which++) {
                if (which ? !adj(p) : !fwd(p)) continue;
                Compression j;
                j.kc = which ? kxa[p] : kc1[p];
                if (which) {
                    vb.kx[p] = kxa[p];
                    vb.X[p] = (const float*)(vbuf + off);
                    off += sizeof(float) * N * (size_t)kxa[p];
                    j.V = vb.X[p] + N;
                    j.unit1 = ps[p]->info.eloc + 1u;
                } else {
                    vb.kc[p] = kc1[p];
                    vb.V[p] = (const float*)(vbuf + off);
                    off += sizeof(float) * N * (size_t)(kc1[p] - 1);
                    j.V = vb.V[p];
                    j.psi0 = psi0;
                }
                // the samples' coefficients of the first stage (V) or of the last (X)
                const std::vector<size_t>& tl = ptraj[p];
                j.S = tl.size();
                j.rows.assign(j.S * (size_t)j.kc, 0.0);
                for (size_t s = 0; s < j.S; s++) {
                    const std::vector<double>& cf = which ? trajs[tl[s]].segs.back().c : trajs[tl[s]].segs[0].c;
                    for (int q = 0; q < j.kc && q < (int)cf.size(); q++) j.rows[s * (size_t)j.kc + q] = cf[q];
                }
                cj.push_back(std::move(j));
                who.emplace_back(p, which);
            }
        }
        trace("compression chunk: %zu instances, %zu bytes", i1 - i0, vbytes);
        mqw_status st = run_bases(c, n, k, ps, trajs, E, vb, use_graph);
        if (st) return st;
        trace("compression chunk: recurrences");
        st = compress_svd(c, n, cj, "cmp");
        if (st) return st;
        // the kept directions, kept for the whole solve
        std::vector<float*> out(cj.size(), nullptr);
        for (int which = 0; which < 2; which++) {
            size_t bytes = 0;
            for (size_t i = 0; i < cj.size(); i++)
                if (cj[i].ok && who[i].second == which) bytes += sizeof(float) * 2 * N * (size_t)cj[i].fc.R;
            if (bytes == 0) continue;
            char* buf = (char*)dev_get(c, (which ? "cmp_Z" : "cmp_Q") + std::to_string(chunk_id), bytes);
            if (!buf) return fail(c, MQW_ENOMEM, "device allocation failed for compressed bases");
            size_t qoff = 0;
            for (size_t i = 0; i < cj.size(); i++)
                if (cj[i].ok && who[i].second == which) {
                    out[i] = (float*)(buf + qoff);
                    qoff += sizeof(float) * 2 * N * (size_t)cj[i].fc.R;
                }
        }
        st = compress_combine(c, n, k, cj, out, "cmp");
        if (st) return st;
        HIPCHK(c, hipStreamSynchronize(c->stream)); // cmp_V is reused by the next chunk
        trace("compression chunk: combines");
        for (size_t i = 0; i < cj.size(); i++) {
            if (!cj[i].ok) continue;
            const size_t p = who[i].first;
            const int which = who[i].second, R = cj[i].fc.R;
            (which ? kzq : kcq)[p] = 2 * R;
            (which ? zbase : qbase)[p] = out[i];
            const std::vector<size_t>& tl = ptraj[p];
            for (size_t s = 0; s < tl.size(); s++) {
                std::vector<float>& u = which ? trajs[tl[s]].cxq : trajs[tl[s]].c1q;
                u.resize((size_t)2 * R);
                for (int q = 0; q < 2 * R; q++) u[q] = (float)cj[i].fc.u[s * 2 * R + q];
            }
        }
        i0 = i1;
        chunk_id++;
    }
    return MQW_OK;
}

// ---------------------------------------------------------------------------
// The compressed third stage, per batch of bases (m >= 4, FP32, split-FP16
// GEMM), for an instance whose third stage is neither read out (the last
// stage, or one of the two-stage adjoint's) nor preceded by an unshared
// second stage.  The samples' second-stage states psi2_j span few
// directions: their real parts and their imaginary parts, each a parity class
// as in the first-stage compression (vector 2j = Re psi2_j, 2j + 1 =
// Im psi2_j; sample j's coefficients are ones at 2j and 2j + 1, so its state
// is sum_k c_jk i^(k&1) V_k).  So:
// 1. the second-stage GEMM writes psi2_j into one contiguous buffer per
//    instance instead of the trajectory slots;
// 2. their Gram pair sums, the SVDs and the combines give U (2R directions)
//    and the samples' coordinates (compress_svd, compress_combine);
// 3. the recurrences W3_lr = T_l(Ht_3) U_r run (run_bases);
// 4. the groups' plans take the third stage as a GEMM over W3 (mid_done),
//    exactly as the second stage over W.
// An instance whose compression fails or does not pay continues from its
// psi2 in the buffer with the third stage as terms (seg0 = 2).  cap bounds
// the W3 bases of the batch.
// ---------------------------------------------------------------------------
mqw_status middle_stage(mqw_ctx* c, int n, const SweepCfg& k, const std::vector<Problem*>& ps,
                        std::vector<Traj>& trajs, const std::vector<std::vector<size_t>>& ptraj,
                        const std::vector<size_t>& cand, const float* E, Bases& B, const std::vector<int>& k3,
                        const std::vector<double>& own3, bool use_graph, size_t cap)
{
    const size_t N = (size_t)1 << n;
    const uint32_t nparts = k.parts;
    size_t ntr = 0;
    for (size_t p : cand) ntr += ptraj[p].size();
    float* buf = (float*)dev_get(c, "mid_psi2", sizeof(float) * 2 * N * ntr);
    double* part = (double*)dev_get(c, "mid_psi2_norm", sizeof(double) * nparts * ntr);
    if (!buf || !part) { // no room: the third stage stays as terms
        trace("third stage: no room for %zu states", ntr);
        return MQW_OK;
    }
    trace("third stage: %zu instances, %zu states", cand.size(), ntr);
    // 1. the second stage's outputs into the buffer, sample j of an instance at
    //    its base + j (every sample of an instance is in this batch)
    std::vector<size_t> grp, base(ps.size(), 0);
    std::vector<Slot<float>> sl;
    for (size_t p : cand) {
        base[p] = grp.size();
        for (size_t t : ptraj[p]) {
            Slot<float> s{};
            s.st[0] = buf + grp.size() * 2 * N;
            s.part[0] = part + grp.size() * nparts;
            sl.push_back(s);
            grp.push_back(t);
        }
    }
    {
        Plan<float> P;
        combine2_ops<float>(n, trajs, grp, B, sl, P);
        double ms_sweep = 0.0, ms_aux = 0.0, ms_basis = 0.0, ms_gemm = 0.0;
        mqw_status st = run_plan<float>(c, MQW_FP32, n, k, P, use_graph, &ms_sweep, &ms_aux, &ms_basis, &ms_gemm,
                                        nullptr, c->stream, "mid");
        if (st) return st;
        c->stats.gpu_ms += ms_sweep + ms_aux;
        c->stats.aux_ms += ms_aux;
        c->stats.gemm_ms += ms_gemm;
        c->stats.gemm_flops += P.gemm_flops;
        c->stats.gemm_useful_flops += P.gemm_useful_flops;
        c->stats.gemm_bytes += P.gemm_bytes;
        c->stats.gemm_launches += 1;
        c->stats.aux_bytes += P.aux_bytes;
        c->stats.aux_launches += P.aux_launches;
    }
    trace("third stage: second-stage states");
    // 2. compression of each instance's states
    std::vector<Compression> cj(cand.size());
    for (size_t i = 0; i < cand.size(); i++) {
        const size_t p = cand[i];
        Compression& j = cj[i];
        j.S = ptraj[p].size();
        j.kc = (int)(2 * j.S);
        j.V0 = buf + base[p] * 2 * N;
        j.V = j.V0 + N;
        j.rows.assign(j.S * (size_t)j.kc, 0.0);
        for (size_t s = 0; s < j.S; s++) j.rows[s * (size_t)j.kc + 2 * s] = j.rows[s * (size_t)j.kc + 2 * s + 1] = 1.0;
    }
    mqw_status st = compress_svd(c, n, cj, "mid");
    if (st) return st;
    // where it pays and fits: the W3 recurrences and GEMM read (20 B per vector
    // and amplitude) against 0.8 x the samples' own third-stage terms
    std::vector<float*> out(cj.size(), nullptr);
    std::vector<size_t> w3off(cj.size(), 0);
    size_t ubytes = 0, wbytes = 0;
    for (size_t i = 0; i < cj.size(); i++) {
        const size_t p = cand[i];
        if (!cj[i].ok) continue;
        const int ku = 2 * cj[i].fc.R;
        const size_t w = sizeof(float) * N * (size_t)ku * (size_t)(k3[p] - 1);
        const bool pays = ku <= kC2MaxK2 && (size_t)ku * (size_t)k3[p] + 2 * kC2RowPad <= (size_t)kC2MaxRows &&
                          20.0 * ku * k3[p] < 0.8 * own3[p] && wbytes + w <= cap;
        trace("third stage: instance %zu, %d of %d directions, k3 %d, W3 %zu bytes%s", p, ku, cj[i].kc, k3[p], w,
              pays ? "" : ": as terms");
        if (!pays) {
            cj[i].ok = false;
            continue;
        }
        w3off[i] = wbytes;
        wbytes += w;
        ubytes += sizeof(float) * N * (size_t)ku;
    }
    char* ub = ubytes ? (char*)dev_get(c, "mid_U", ubytes) : nullptr;
    char* wb = wbytes ? (char*)dev_get(c, "mid_W3", wbytes) : nullptr;
    if (ubytes && (!ub || !wb)) {
        trace("third stage: no room for %zu + %zu bytes of bases", ubytes, wbytes);
        for (Compression& j : cj) j.ok = false; // no room: the third stage stays as terms
    }
    size_t uo = 0;
    for (size_t i = 0; i < cj.size(); i++)
        if (cj[i].ok) {
            out[i] = (float*)(ub + uo);
            uo += sizeof(float) * N * (size_t)(2 * cj[i].fc.R);
        }
    st = compress_combine(c, n, k, cj, out, "mid");
    if (st) return st;
    // 3. the W3 recurrences, and the row tables of the third-stage GEMM
    Bases mb;
    mb.V.assign(ps.size(), nullptr);
    mb.kc.assign(ps.size(), 0);
    mb.X.assign(ps.size(), nullptr);
    mb.kx.assign(ps.size(), 0);
    mb.kc2.assign(ps.size(), 0);
    mb.U.assign(ps.size(), nullptr);
    mb.ku.assign(ps.size(), 0);
    mb.W3.assign(ps.size(), nullptr);
    mb.k3.assign(ps.size(), 0);
    std::vector<uint32_t> rows, all;
    std::vector<size_t> roff(cj.size(), 0);
    B.U.assign(ps.size(), nullptr);
    B.ku.assign(ps.size(), 0);
    B.W3.assign(ps.size(), nullptr);
    B.k3.assign(ps.size(), 0);
    B.mtab.assign(ps.size(), nullptr);
    B.m0.assign(ps.size(), 0);
    B.m1.assign(ps.size(), 0);
    bool any = false;
    for (size_t i = 0; i < cj.size(); i++) {
        if (!cj[i].ok) continue;
        const size_t p = cand[i];
        const int ku = 2 * cj[i].fc.R;
        mb.U[p] = B.U[p] = out[i];
        mb.ku[p] = B.ku[p] = ku;
        mb.W3[p] = B.W3[p] = (const float*)(wb + w3off[i]);
        mb.k3[p] = B.k3[p] = k3[p];
        second_stage_rows(ku, k3[p], rows, B.m0[p], B.m1[p]);
        roff[i] = all.size();
        all.insert(all.end(), rows.begin(), rows.end());
        any = true;
    }
    if (any) {
        uint32_t* rd = (uint32_t*)dev_get(c, "mid_rows", sizeof(uint32_t) * all.size());
        if (!rd) return fail(c, MQW_ENOMEM, "device allocation failed for third-stage rows");
        HIPCHK(c, hipMemcpyAsync(rd, all.data(), sizeof(uint32_t) * all.size(), hipMemcpyHostToDevice, c->stream));
        for (size_t i = 0; i < cj.size(); i++)
            if (cj[i].ok) B.mtab[cand[i]] = rd + roff[i];
        // (the next batch's recurrences may be running on the second stream
        // with the "bases" tables)
        st = run_bases(c, n, k, ps, trajs, E, mb, use_graph, nullptr, nullptr, nullptr, "mid_bases");
        if (st) return st;
    } else {
        B.U.clear();
        B.ku.clear();
    }
    trace("third stage: bases");
    // 4. the trajectories continue from the third stage's GEMM, or from their
    //    second-stage states with the third stage as terms
    for (size_t i = 0; i < cj.size(); i++) {
        const size_t p = cand[i];
        const int R = cj[i].ok ? cj[i].fc.R : 0;
        for (size_t s = 0; s < ptraj[p].size(); s++) {
            Traj& tr = trajs[ptraj[p][s]];
            tr.second_done = false;
            tr.first_done = false;
            if (cj[i].ok) {
                tr.mid_done = true;
                tr.cmq.resize((size_t)2 * R);
                for (int q = 0; q < 2 * R; q++) tr.cmq[q] = (float)cj[i].fc.u[s * 2 * R + q];
            } else {
                tr.seg0 = 2;
                tr.psi_in = buf + (base[p] + s) * 2 * N;
            }
        }
        if (cj[i].ok) {
            c->stats.third_stage_instances++;
            c->stats.third_stage_vectors += (uint64_t)(2 * R);
            c->stats.third_stage_states += (uint64_t)(2 * ptraj[p].size());
        }
    }
    return MQW_OK;
}

// ---------------------------------------------------------------------------
// One group of problems of the same n.
// ---------------------------------------------------------------------------
template <typename real>
mqw_status solve_group(mqw_ctx* c, int n, std::vector<Problem*>& ps, const mqw_opts& o, std::vector<double>& probs)
{
    const size_t N = (size_t)1 << n;
    const int precision = o.precision;
    const int samples = o.samples;
    g_trace_t0 = std::chrono::steady_clock::now();
    trace("solve_group n=%d, %zu problems", n, ps.size());
    SweepCfg k = choose_sweep(c, precision, n, o.tile_bits, o.chunk_bits, o.variant, o.high_pass);
    k.chains = o.chain_recurrences && c->chain_capable && k.half && n >= kChainMinN && n <= kChainMaxN;
    k.chain_polls = o.chain_wait_polls;
    const bool mixer = o.mixer_count > 0;
    if (mixer) {
        // X-string mixer: its own sweep kernel, masks on the device
        for (int q = 0; q < o.mixer_count; q++)
            if ((uint64_t)o.mixer_masks[q] >= (1ull << n))
                return fail(c, MQW_EINVAL, "mixer mask %u has bits above n=%d", o.mixer_masks[q], n);
        uint32_t* md = (uint32_t*)dev_get(c, "mixer_masks", sizeof(uint32_t) * o.mixer_count);
        if (!md) return fail(c, MQW_ENOMEM, "device allocation failed for mixer masks");
        HIPCHK(c, hipMemcpyAsync(md, o.mixer_masks, sizeof(uint32_t) * o.mixer_count, hipMemcpyHostToDevice,
                                 c->stream));
        k = SweepCfg{};
        k.tb = mask_tile_bits(n);
        k.parts = 1u << (n - k.tb);
        k.masks = md;
        k.nmask = o.mixer_count;
        k.xy = o.mixer_kind == MQW_MIXER_XY;
    }
    // spectral bound of the mixer: n for sum_i X_i, the mask count for
    // X-strings, 2 per pair for XY (each X_a X_b + Y_a Y_b has norm 2)
    const float mix_norm = mixer ? (float)o.mixer_count * (o.mixer_kind == MQW_MIXER_XY ? 2.0f : 1.0f)
                                 : (float)(unsigned)n;
    real* E = (real*)dev_get(c, "E_group", sizeof(real) * N * ps.size());
    if (!E) return fail(c, MQW_ENOMEM, "device allocation failed for %zu diagonals at n=%d", ps.size(), n);
    mqw_status st = group_energies(c, n, ps, o.energy_mode, precision, E);
    if (st) return st;

    // trajectories and their segments
    std::vector<Traj> trajs;
    for (size_t p = 0; p < ps.size(); p++) {
        Problem& pr = *ps[p];
        const int m = pr.m;
        auto seg = [&](int i, int j, int64_t slot) {
            Segment sg;
            sg.gamma = pr.sch.gamma[i];
            // spectral bound; |gamma| so schedules that cross zero keep a valid bound
            sg.onenorm = fmaf(fabsf(sg.gamma), mix_norm, pr.info.e_abs);
            sg.s = (double)(sg.onenorm * pr.times[(size_t)i * samples + j]);
            sg.prob_slot = slot;
            return sg;
        };
        if (m == 1) {
            Traj tr;
            tr.prob = (int)p;
            for (int j = 0; j < samples; j++) tr.segs.push_back(seg(0, j, (int64_t)(pr.prob_base + j)));
            trajs.push_back(std::move(tr));
        } else {
            for (int j = 0; j < samples; j++) {
                Traj tr;
                tr.prob = (int)p;
                for (int i = 0; i < m; i++)
                    tr.segs.push_back(seg(i, j, i + 1 == m ? (int64_t)(pr.prob_base + j) : -1));
                trajs.push_back(std::move(tr));
            }
        }
    }
    trace("energies");
    fill_coeffs(trajs, o.tol, precision == MQW_FP32);
    trace("coefficients of %zu trajectories", trajs.size());

    double* probs_d = (double*)dev_get(c, "probs", sizeof(double) * probs.size());
    if (!probs_d) return fail(c, MQW_ENOMEM, "device allocation failed for probabilities");
    HIPCHK(c, hipMemcpyAsync(probs_d, probs.data(), sizeof(double) * probs.size(), hipMemcpyHostToDevice, c->stream));
    const int small_max = o.small_n_max >= 0 ? o.small_n_max : 13;
    if (precision == MQW_FP32 && n <= small_max && lds_supported(n) && !mixer) {
        st = run_lds(c, n, trajs, ps, (const float*)E, probs_d, o.use_graph != 0);
        if (st) return st;
        HIPCHK(c, hipMemcpyAsync(probs.data(), probs_d, sizeof(double) * probs.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return MQW_OK;
    }

    const uint32_t nparts = k.parts;
    // Shared bases, per instance with m > 1 (its samples restart from |+>):
    // the first stage's V (kc vectors, combined per sample), the second
    // stage's W (kc x kc2, from m = 3, when it costs less than the samples'
    // own terms) and the last stage's adjoint X (kx).  0 = not shared.
    const bool basis_path =
        std::is_same<real, float>::value && k.half && !mixer && n >= k.tb && (N / nparts) % 256 == 0;
    const size_t vcap = o.max_device_bytes ? o.max_device_bytes / 4 : ((size_t)16 << 30);
    if (basis_path && o.direct_m1) {
        st = direct_m1(c, n, k, ps, trajs, (const float*)E, o.tol, o.use_graph != 0, vcap, probs, probs_d);
        if (st) return st;
        trace("direct m = 1 chains");
    }
    std::vector<int> kc1(ps.size(), 0), kc2(ps.size(), 0), kx(ps.size(), 0), ky(ps.size(), 0), k3(ps.size(), 0);
    std::vector<double> own_pen(ps.size(), 0.0); // bytes per amplitude of the samples' own stage m - 1 terms
    std::vector<double> own3(ps.size(), 0.0);    // and of their own third-stage terms
    if (basis_path) {
        std::vector<double> own2(ps.size(), 0.0); // bytes per amplitude of the samples' own second-stage terms
        for (const Traj& tr : trajs) {
            const int pm = ps[tr.prob]->m;
            if (pm < 2) continue;
            if (o.shared_first_stage) kc1[tr.prob] = std::max(kc1[tr.prob], (int)tr.segs.front().c.size());
            if (o.final_readout) kx[tr.prob] = std::max(kx[tr.prob], (int)tr.segs.back().c.size());
            if (pm >= 3 && o.shared_second_stage) {
                kc2[tr.prob] = std::max(kc2[tr.prob], (int)tr.segs[1].c.size());
                own2[tr.prob] += 36.0 * (double)std::max<size_t>(tr.segs[1].c.size() - 1, 1);
            }
            if (pm >= 4 && o.final_readout && o.adjoint_two_stages) {
                const size_t K = tr.segs[(size_t)pm - 2].c.size();
                ky[tr.prob] = std::max(ky[tr.prob], (int)K);
                own_pen[tr.prob] += 36.0 * (double)std::max<size_t>(K - 1, 1);
            }
            if (pm >= 4 && o.shared_third_stage) {
                const size_t K = tr.segs[2].c.size();
                k3[tr.prob] = std::max(k3[tr.prob], (int)K);
                own3[tr.prob] += 36.0 * (double)std::max<size_t>(K - 1, 1);
            }
        }
        const bool split = o.second_stage_gemm == MQW_GEMM_SPLIT_FP16;
        const int kc2max = split ? kC2MaxK2 : kC2MaxK; // the GEMM's second-stage coefficients
        for (size_t p = 0; p < ps.size(); p++) {
            if (kc1[p] > kCombineMaxKX) kc1[p] = 0;
            if (kx[p] > (split && n <= kReadMfmaMaxN ? kReadMaxK : kReadMaxKF32)) kx[p] = 0;
            // the two-stage adjoint: a second-stage GEMM over Y (limits as W's)
            if (ky[p] < 2 || ky[p] > kc2max || kx[p] < 8 || kx[p] > kAdjMaxK ||
                (o.second_stage_gemm == MQW_GEMM_SPLIT_FP16 && (N / nparts) % 512 != 0))
                ky[p] = 0;
            // W: recurrences (16 B per vector and amplitude) plus one GEMM read (4 B), against the
            // samples' own terms, with a margin for the matrix work
            const double shared2 = 20.0 * (double)kc1[p] * (double)kc2[p];
            // (the split-FP16 kernel takes amplitude chunks of 512: always so with half tiles)
            // (kc1 beyond the GEMM's limit only where the compression may bring it within)
            if (kc1[p] < 1 || (kc1[p] > kC2MaxK && !o.compress_bases) || kc2[p] < 2 || kc2[p] > kc2max ||
                shared2 > 0.8 * own2[p] ||
                (o.second_stage_gemm == MQW_GEMM_SPLIT_FP16 && (N / nparts) % 512 != 0))
                kc2[p] = 0;
        }
    }
    // second-stage and two-stage adjoint bases: batches within wcap; an instance
    // whose own bases exceed it (Y at n = 22 is about 40 GiB) runs alone in a
    // batch of up to twice that
    const size_t wcap = o.max_device_bytes ? o.max_device_bytes / 4 : ((size_t)32 << 30);
    const size_t wcap1 = 2 * wcap;
    // compressed first-stage bases (kcq = 2R vectors from Q, else the kc1 V_k)
    // and compressed last-stage adjoint bases of the two-stage adjoint (kzq = 2R from Z)
    std::vector<int> kcq(ps.size(), 0), kzq(ps.size(), 0);
    std::vector<const float*> qbase(ps.size(), nullptr), zbase(ps.size(), nullptr);
    auto kcw = [&](size_t p) { return kcq[p] > 0 ? kcq[p] : kc1[p]; };
    auto wbytes = [&](size_t p) -> size_t {
        return kc2[p] ? sizeof(float) * N * (size_t)kcw(p) * (size_t)(kc2[p] - 1) : 0;
    };
    auto ybytes = [&](size_t p) -> size_t {
        return kzq[p] ? sizeof(float) * N * (size_t)kzq[p] * (size_t)(ky[p] - 1) : 0;
    };
    if (basis_path) {
        std::vector<int> kxa(ps.size(), 0);
        bool any = o.compress_bases != 0;
        for (size_t p = 0; p < ps.size(); p++) {
            // not even a quarter of the last-stage basis would pay: no compression
            kxa[p] = ky[p] > 0 && 20.0 * (kx[p] / 4) * ky[p] < 0.8 * own_pen[p] ? kx[p] : 0;
            any |= kxa[p] > 0;
        }
        if (any) {
            const std::vector<int> none(ps.size(), 0);
            st = compress_bases(c, n, k, ps, trajs, (const float*)E, kc1, o.compress_bases ? kc2 : none, kxa,
                                o.use_graph != 0, vcap, kcq, qbase, kzq, zbase);
            if (st) return st;
            trace("compressed bases");
        }
        // the two-stage adjoint where its Y recurrences and GEMM read (20 B per
        // vector and amplitude) cost less than the samples' own stage m - 1 terms
        for (size_t p = 0; p < ps.size(); p++)
            if (kzq[p] > 0 && (kzq[p] > kC2MaxK || 20.0 * kzq[p] * ky[p] >= 0.8 * own_pen[p])) kzq[p] = 0;
    }
    // second-stage bases that do not fit a batch stay per-sample terms (their
    // first stage then comes from V, recomputed in the batch); so do the
    // stage m - 1 terms of a two-stage adjoint that does not fit beside them
    for (size_t p = 0; p < ps.size(); p++) {
        // W's first-stage directions within the GEMM's limit (compressed, or kc1 itself)
        if (kc2[p] > 0 && kcw(p) > kC2MaxK) {
            kc2[p] = 0;
            kcq[p] = 0;
            qbase[p] = nullptr;
        }
        if (kzq[p] > 0 && wbytes(p) + ybytes(p) > wcap1) kzq[p] = 0;
        if (kzq[p] > 0) {
            c->stats.adjoint2_instances++;
            c->stats.adjoint2_vectors += (uint64_t)kzq[p];
            c->stats.adjoint2_uncompressed += (uint64_t)kx[p];
        }
        if (wbytes(p) + ybytes(p) > wcap1) {
            kc2[p] = 0;
            kcq[p] = 0;
            qbase[p] = nullptr;
        }
        if (kcq[p] > 0) {
            c->stats.compressed_instances++;
            c->stats.compressed_vectors += (uint64_t)kcq[p];
            c->stats.uncompressed_vectors += (uint64_t)kc1[p];
        }
    }
    // problem batches whose bases fit together: second-stage and two-stage
    // adjoint bases within wcap, first-stage and last-stage bases made in the
    // batch within vcap (so no instance's shared stage falls back to terms for
    // lack of room beside the others)
    auto vxbytes = [&](size_t p) -> size_t {
        if (!basis_path) return 0;
        const size_t bv = kcq[p] > 0 ? 0 : sizeof(float) * N * (size_t)std::max(kc1[p] - 1, 0);
        const size_t bx = kzq[p] > 0 ? 0 : sizeof(float) * N * (size_t)kx[p];
        return bv + bx;
    };
    std::vector<std::pair<size_t, size_t>> pbatches;
    {
        size_t b0 = 0, wsum = 0, vsum = 0;
        for (size_t p = 0; p < ps.size(); p++) {
            if (p > b0 && (wsum + wbytes(p) + ybytes(p) > wcap || vsum + vxbytes(p) > vcap)) {
                pbatches.emplace_back(b0, p);
                b0 = p;
                wsum = 0;
                vsum = 0;
            }
            wsum += wbytes(p) + ybytes(p);
            vsum += vxbytes(p);
        }
        pbatches.emplace_back(b0, ps.size());
    }
    std::vector<std::vector<size_t>> ptraj(ps.size());
    for (size_t t = 0; t < trajs.size(); t++) ptraj[trajs[t].prob].push_back(t);
    size_t max_batch_trajs = 0;
    for (const auto& pb : pbatches) {
        size_t cnt = 0;
        for (size_t p = pb.first; p < pb.second; p++) cnt += ptraj[p].size();
        max_batch_trajs = std::max(max_batch_trajs, cnt);
    }

    // trajectory slots, sized to the memory-side cache budget
    const size_t nst = k.hi > 0 ? 8 : 6; // planar states per slot (+ the high-bit sums)
    const size_t per_traj = sizeof(real) * nst * N + sizeof(double) * 2 * nparts;
    const size_t budget = o.group_bytes ? o.group_bytes : ((size_t)32 << 30);
    size_t G = std::max<size_t>(1, budget / per_traj);
    if (o.max_device_bytes) G = std::min(G, std::max<size_t>(1, o.max_device_bytes / 2 / per_traj));
    G = std::min(G, std::max<size_t>(1, max_batch_trajs));
    trace("%zu batches of bases", pbatches.size());
    real* state = (real*)dev_get(c, "state", sizeof(real) * nst * N * G);
    double* parts = (double*)dev_get(c, "norm_part", sizeof(double) * 2 * nparts * G);
    if (!state || !parts) return fail(c, MQW_ENOMEM, "device allocation failed for %zu trajectory slots at n=%d", G, n);
    std::vector<Slot<real>> slots(G);
    for (size_t i = 0; i < G; i++) {
        for (int q = 0; q < 3; q++) slots[i].st[q] = state + (i * 3 + (size_t)q) * 2 * N;
        if (k.hi > 0) slots[i].hs = state + (3 * G + i) * 2 * N;
        for (int q = 0; q < 2; q++) slots[i].part[q] = parts + (i * 2 + (size_t)q) * nparts;
    }

    // Each problem batch's bases: offsets in a basis buffer and the row tables
    // of its second-stage GEMM; the pointers are bound when the batch is given
    // a buffer.
    struct BatchBases {
        Bases b;
        std::vector<size_t> offv, offx, offw, qoff, offy, yoff;
        std::vector<uint32_t> rows;
        size_t total = 0, plus_off = 0;
        bool any_w = false, any_y = false, use = false;
    };
    std::vector<BatchBases> bbs(pbatches.size());
    size_t max_total = 0, max_rows = 0;
    if (basis_path) {
        for (size_t q = 0; q < pbatches.size(); q++) {
            const auto& pb = pbatches[q];
            BatchBases& bb = bbs[q];
            Bases& bases = bb.b;
            bases.split = o.second_stage_gemm == MQW_GEMM_SPLIT_FP16;
            bases.V.assign(ps.size(), nullptr);
            bases.kc.assign(ps.size(), 0);
            bases.X.assign(ps.size(), nullptr);
            bases.kx.assign(ps.size(), 0);
            bases.W.assign(ps.size(), nullptr);
            bases.kc2.assign(ps.size(), 0);
            bases.qtab.assign(ps.size(), nullptr);
            bases.q0.assign(ps.size(), 0);
            bases.q1.assign(ps.size(), 0);
            bb.offv.assign(ps.size(), 0);
            bb.offx.assign(ps.size(), 0);
            bb.offw.assign(ps.size(), 0);
            bb.qoff.assign(ps.size(), 0);
            bases.cmp.assign(ps.size(), 0);
            bases.plusv.assign(ps.size(), nullptr);
            bases.Z.assign(ps.size(), nullptr);
            bases.kz.assign(ps.size(), 0);
            bases.Y.assign(ps.size(), nullptr);
            bases.ky.assign(ps.size(), 0);
            bases.ytab.assign(ps.size(), nullptr);
            bases.y0.assign(ps.size(), 0);
            bases.y1.assign(ps.size(), 0);
            bb.offy.assign(ps.size(), 0);
            bb.yoff.assign(ps.size(), 0);
            size_t vbytes = 0;
            std::vector<uint32_t> rows;
            for (size_t p = pb.first; p < pb.second; p++) {
                const size_t bv = sizeof(float) * N * (size_t)std::max(kc1[p] - 1, 0);
                const size_t bx = sizeof(float) * N * (size_t)kx[p];
                if (kcq[p] > 0) { // the compressed basis Q, computed before the batches
                    bases.kc[p] = kcq[p];
                    bases.cmp[p] = 1;
                    bases.plusv[p] = qbase[p];
                    bases.V[p] = qbase[p] + N;
                } else if (kc1[p] > 0 && vbytes + bv <= vcap) {
                    bases.kc[p] = kc1[p];
                    bb.offv[p] = bb.total;
                    bb.total += bv;
                    vbytes += bv;
                }
                if (kzq[p] > 0) { // the two-stage adjoint: Z from before the batch, Y in it, no X
                    bases.kz[p] = kzq[p];
                    bases.ky[p] = ky[p];
                    bases.Z[p] = zbase[p];
                    bb.offy[p] = bb.total;
                    bb.total += ybytes(p);
                    second_stage_rows(kzq[p], ky[p], rows, bases.y0[p], bases.y1[p]);
                    bb.yoff[p] = bb.rows.size();
                    bb.rows.insert(bb.rows.end(), rows.begin(), rows.end());
                    bb.any_y = true;
                } else if (kx[p] > 0 && vbytes + bx <= vcap) {
                    bases.kx[p] = kx[p];
                    bb.offx[p] = bb.total;
                    bb.total += bx;
                    vbytes += bx;
                }
                if (kc2[p] > 0 && bases.kc[p] > 0) {
                    bases.kc2[p] = kc2[p];
                    bb.offw[p] = bb.total;
                    bb.total += wbytes(p);
                    second_stage_rows(bases.kc[p], kc2[p], rows, bases.q0[p], bases.q1[p]);
                    bb.qoff[p] = bb.rows.size();
                    bb.rows.insert(bb.rows.end(), rows.begin(), rows.end());
                    bb.any_w = true;
                }
                bb.use |= bases.kc[p] > 0 || bases.kx[p] > 0 || bases.kz[p] > 0;
            }
            bb.plus_off = bb.total;
            if (bb.any_w) bb.total += sizeof(float) * N;
            max_total = std::max(max_total, bb.total);
            max_rows = std::max(max_rows, bb.rows.size());
        }
    }
    // Several batches: two basis buffers, the next batch's recurrences on the
    // second stream while this batch's plans run on the main one.
    // (two buffers of bases within a third of the 288 GB part: the n = 22
    // instances with m >= 4, about 60 GB of bases each, run one buffer)
    // (not with on-chip recurrences: a launch of them needs every CU of the
    // device, and two on two streams could wait for each other's workgroups)
    const bool overlap = o.overlap_bases && !k.chains && c->stream2 && pbatches.size() > 1 && max_total > 0 &&
                         (!o.max_device_bytes || 2 * max_total <= o.max_device_bytes / 2) &&
                         2 * max_total <= ((size_t)96 << 30);
    const int nbuf = overlap ? 2 : 1;
    char* bbase[2] = {nullptr, nullptr};
    uint32_t* bq[2] = {nullptr, nullptr};
    trace("trajectory slots: %zu", G);
    for (int b = 0; b < nbuf && max_total > 0; b++) {
        bbase[b] = (char*)dev_get(c, "basis" + std::to_string(b), max_total);
        if (!bbase[b]) return fail(c, MQW_ENOMEM, "device allocation failed for the shared bases");
        if (max_rows) {
            bq[b] = (uint32_t*)dev_get(c, "basis_rows" + std::to_string(b), sizeof(uint32_t) * max_rows);
            if (!bq[b]) return fail(c, MQW_ENOMEM, "device allocation failed for second-stage rows");
        }
    }
    // the second stream's work in flight: its plans (whose host tables the
    // copies read), graphs and events; drained before any of them is freed
    std::deque<Plan<float>> keep[2];
    Timed timed[2];
    bool recorded[2] = {false, false};
    struct Drain {
        mqw_ctx* c;
        bool on;
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~Drain()
        {
            if (on) (void)hipStreamSynchronize(c->stream2);
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
        }
    } drain{c, overlap};
    if (overlap)
        for (int b = 0; b < 2; b++) HIPCHK(c, hipEventCreateWithFlags(&drain.ev[b], hipEventDisableTiming));
    // binds batch q's bases to its buffer and runs their recurrences: at once
    // on the main stream, or queued on the second stream when overlapping
    // On-chip recurrences of a batch run while the host prepares its plan; the
    // plan runs once they are checked (settle).  (Declared after the host
    // buffers of bind's copies: on an early return it waits for the stream
    // before they go.)
    ChainDefer cdefer;
    auto settle = [&](size_t q) -> mqw_status {
        bool ok = true;
        mqw_status s = finish_chains(c, cdefer, &ok);
        if (s || ok) return s;
        SweepCfg k1 = k;
        k1.chains = false;
        return run_bases(c, n, k1, ps, trajs, (const float*)E, bbs[q].b, o.use_graph != 0);
    };
    auto bind = [&](size_t q) -> mqw_status {
        BatchBases& bb = bbs[q];
        Bases& bases = bb.b;
        const int b = overlap ? (int)(q & 1) : 0;
        const hipStream_t bs = overlap ? c->stream2 : c->stream;
        char* base = bbase[b];
        const auto& pb = pbatches[q];
        if (bb.any_w) bases.plus = (const float*)(base + bb.plus_off);
        if (!bb.rows.empty())
            HIPCHK(c, hipMemcpyAsync(bq[b], bb.rows.data(), sizeof(uint32_t) * bb.rows.size(), hipMemcpyHostToDevice,
                                     bs));
        for (size_t p = pb.first; p < pb.second; p++) {
            if (bases.kc[p] > 0 && !bases.cmp[p])
                bases.V[p] = (const float*)(base + bb.offv[p]); // unused when kc = 1 (V_0 only)
            if (bases.kx[p] > 0) bases.X[p] = (const float*)(base + bb.offx[p]);
            if (bases.kc2[p] > 0) {
                bases.W[p] = (const float*)(base + bb.offw[p]);
                bases.qtab[p] = bq[b] + bb.qoff[p];
            }
            if (bases.kz[p] > 0) {
                bases.Y[p] = (const float*)(base + bb.offy[p]);
                bases.ytab[p] = bq[b] + bb.yoff[p];
            }
        }
        mqw_status rs = MQW_OK;
        if (bb.use) {
            if (overlap) {
                rs = run_bases(c, n, k, ps, trajs, (const float*)E, bases, o.use_graph != 0, &timed[b], c->stream2,
                               &keep[b]);
                if (rs) return rs;
                HIPCHK(c, hipEventRecord(drain.ev[b], c->stream2));
                recorded[b] = true;
            } else {
                rs = run_bases(c, n, k, ps, trajs, (const float*)E, bases, o.use_graph != 0, nullptr, nullptr, nullptr,
                               "bases", &cdefer);
                if (rs) return rs;
            }
        }
        if (!bb.any_w) {
            bases.W.clear();
            bases.kc2.clear();
        }
        if (!bb.any_y) {
            bases.Z.clear();
            bases.kz.clear();
            bases.Y.clear();
            bases.ky.clear();
        }
        return MQW_OK;
    };
    if (overlap) {
        st = bind(0);
        if (st) return st;
    }
    for (size_t q = 0; q < pbatches.size(); q++) {
        const auto& pb = pbatches[q];
        if (overlap) {
            if (q + 1 < pbatches.size()) {
                st = bind(q + 1);
                if (st) return st;
            }
            if (recorded[q & 1]) HIPCHK(c, hipStreamWaitEvent(c->stream, drain.ev[q & 1], 0));
        } else if (basis_path) {
            st = bind(q);
            if (st) return st;
        }
        trace("batch %zu of %zu: bases bound", q, pbatches.size());
        Bases& bases = bbs[q].b;
        const bool use_bases = bbs[q].use;
        c->stats.basis_batches += use_bases;
        c->stats.overlap_batches += recorded[q & 1];
        // the batch's trajectories, longest first
        std::vector<size_t> order;
        for (size_t p = pb.first; p < pb.second; p++)
            for (size_t t : ptraj[p]) {
                Traj& tr = trajs[t];
                if (tr.direct_done) continue;
                tr.second_done = use_bases && !bases.kc2.empty() && bases.kc2[p] > 0;
                tr.first_done = use_bases && !tr.second_done && bases.kc[p] > 0;
                tr.adjoint2 = use_bases && !bases.kz.empty() && bases.kz[p] > 0;
                tr.adjoint = use_bases && !tr.adjoint2 && bases.kx[p] > 0;
                order.push_back(t);
            }
        // the compressed third stage of the batch's instances where it applies
        if (use_bases && o.shared_third_stage && !bases.kc2.empty() &&
            o.second_stage_gemm == MQW_GEMM_SPLIT_FP16) {
            std::vector<size_t> cand;
            for (size_t p = pb.first; p < pb.second; p++) {
                const int pm = ps[p]->m;
                const size_t S = ptraj[p].size();
                const bool adj3 = !bases.kz.empty() && bases.kz[p] > 0 && pm == 4; // stage 3 read out by the adjoint
                if (pm < 4 || adj3 || bases.kc2[p] <= 0 || S < 2 || 2 * S > (size_t)kGramMaxK || k3[p] < 2 ||
                    k3[p] > kC2MaxK2)
                    continue;
                bool all = true; // every sample still to run (none evaluated otherwise)
                for (size_t t : ptraj[p]) all &= trajs[t].second_done;
                // at most 2S directions (and at most the GEMM's 128)
                const double ku = (double)std::min<size_t>(2 * S, (size_t)kC2MaxK2);
                if (all && 20.0 * 0.5 * ku * k3[p] < 0.8 * own3[p]) cand.push_back(p);
            }
            if (!cand.empty()) {
                st = settle(q); // (the third stage reads results back to the host)
                if (st) return st;
                st = middle_stage(c, n, k, ps, trajs, ptraj, cand, (const float*)E, bases, k3, own3,
                                  o.use_graph != 0, wcap1);
                if (st) return st;
            }
        }
        std::vector<size_t> cost(trajs.size(), 0);
        for (size_t t : order) {
            const Traj& tr = trajs[t];
            const size_t s1 = tr.segs.size() - (tr.adjoint2 ? 2 : tr.adjoint ? 1 : 0);
            for (size_t s = tr.mid_done ? 3 : tr.second_done ? 2 : tr.first_done ? 1 : (size_t)tr.seg0; s < s1; s++)
                cost[t] += std::max<size_t>(tr.segs[s].c.size() - 1, 1);
        }
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
        std::vector<std::vector<size_t>> groups;
        for (size_t i = 0; i < order.size(); i += G)
            groups.emplace_back(order.begin() + i, order.begin() + std::min(order.size(), i + G));
        // bound the host job table: run the groups in batches
        const size_t jobs_cap = (size_t)1 << 22;
        size_t g0 = 0;
        while (g0 < groups.size()) {
            size_t g1 = g0, njobs = 0;
            while (g1 < groups.size()) {
                size_t gj = 0;
                for (size_t t : groups[g1]) gj += cost[t];
                if (g1 > g0 && njobs + gj > jobs_cap) break;
                njobs += gj;
                g1++;
            }
            std::vector<std::vector<size_t>> batch(groups.begin() + g0, groups.begin() + g1);
            Plan<real> plan;
            build_plan<real>(n, nparts, trajs, batch, ps, E, slots, probs_d, plan, k.hi,
                             use_bases ? &bases : nullptr);
            trace("batch %zu: plan of %zu ops", q, plan.ops.size());
            st = settle(q);
            if (st) return st;
            double ms_sweep = 0.0, ms_aux = 0.0, ms_basis = 0.0, ms_gemm = 0.0;
            st = run_plan<real>(c, precision, n, k, plan, o.use_graph != 0, &ms_sweep, &ms_aux, &ms_basis, &ms_gemm);
            if (st) return st;
            trace("batch %zu: plan ran", q);
            c->stats.gpu_ms += ms_sweep + ms_aux;
            c->stats.sweep_ms += ms_sweep;
            c->stats.aux_ms += ms_aux;
            c->stats.gemm_ms += ms_gemm;
            c->stats.gemm_flops += plan.gemm_flops;
            c->stats.gemm_useful_flops += plan.gemm_useful_flops;
            c->stats.gemm_bytes += plan.gemm_bytes;
            for (const Op& op : plan.ops) c->stats.gemm_launches += op.kind == 7;
            c->stats.aux_bytes += plan.aux_bytes;
            c->stats.aux_launches += plan.aux_launches;
            c->stats.sweeps += plan.sweeps;
            c->stats.term_units += plan.terms;
            c->stats.amp_terms += plan.amp_terms;
            c->stats.psi_const_amp_terms += plan.psi_const_amp_terms;
            g0 = g1;
        }
        st = settle(q); // (a batch that ran no plan)
        if (st) return st;
        if (recorded[q & 1]) {
            // the batch's recurrences ran on the second stream (their event
            // times overlap the main stream's)
            Timed& tm = timed[q & 1];
            HIPCHK(c, hipEventSynchronize(drain.ev[q & 1]));
            double ms_sweep = 0.0, ms_aux = 0.0, ms_basis = 0.0;
            collect(tm, &ms_sweep, &ms_aux, &ms_basis, nullptr);
            c->stats.gpu_ms += ms_sweep + ms_aux;
            c->stats.aux_ms += ms_sweep + ms_aux;
            c->stats.basis_ms += ms_basis;
            tm.evs.clear();
            tm.graphs.clear();
            tm.cls.clear();
            keep[q & 1].clear();
            recorded[q & 1] = false;
        }
    }
    HIPCHK(c, hipMemcpyAsync(probs.data(), probs_d, sizeof(double) * probs.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MQW_OK;
}

// Options are checked before any problem is prepared (prepare_problem sizes
// the sample tables from o.samples).
mqw_status check_opts(mqw_ctx* c, const mqw_opts& o)
{
    if (o.samples < 1) return fail(c, MQW_EINVAL, "samples must be >= 1 (got %d)", o.samples);
    if (!(o.tol > 0.0)) return fail(c, MQW_EINVAL, "tol must be > 0");
    if (o.precision != MQW_FP32 && o.precision != MQW_FP64) return fail(c, MQW_EINVAL, "bad precision");
    if (o.energy_mode != MQW_ENERGY_COMPAT && o.energy_mode != MQW_ENERGY_EXACT)
        return fail(c, MQW_EINVAL, "bad energy mode");
    if (o.second_stage_gemm != MQW_GEMM_SPLIT_FP16 && o.second_stage_gemm != MQW_GEMM_FP32)
        return fail(c, MQW_EINVAL, "bad second_stage_gemm %d", o.second_stage_gemm);
    if (o.mixer_count < 0 || o.mixer_count > 64 || (o.mixer_count > 0 && !o.mixer_masks))
        return fail(c, MQW_EINVAL, "mixer_count must be 0..64 with masks given");
    for (int q = 0; q < o.mixer_count; q++)
        if (o.mixer_masks[q] == 0u) return fail(c, MQW_EINVAL, "mixer mask %d is empty", q);
    if (o.mixer_kind != MQW_MIXER_XSTRING && o.mixer_kind != MQW_MIXER_XY)
        return fail(c, MQW_EINVAL, "bad mixer_kind %d", o.mixer_kind);
    if (o.chain_wait_polls < -1)
        return fail(c, MQW_EINVAL, "chain_wait_polls must be >= -1 (got %d)", o.chain_wait_polls);
    if (o.mixer_kind == MQW_MIXER_XY)
        for (int q = 0; q < o.mixer_count; q++)
            if (__builtin_popcount(o.mixer_masks[q]) != 2)
                return fail(c, MQW_EINVAL, "XY mixer mask %d (0x%x) must have exactly two bits", q, o.mixer_masks[q]);
    return MQW_OK;
}

mqw_status solve_problems(mqw_ctx* c, std::vector<Problem>& probs_in, const mqw_opts& o, float* prob,
                          float* total_t)
{
    c->stats = mqw_stats{};
    if (mqw_status st = check_opts(c, o)) return st;
    std::map<int, std::vector<Problem*>> groups;
    size_t slots = 0;
    for (Problem& p : probs_in) {
        p.prob_base = slots;
        slots += (size_t)o.samples;
        groups[p.n].push_back(&p);
    }
    std::vector<double> sp(slots, 0.0);
    for (auto& kv : groups) {
        mqw_status st = o.precision == MQW_FP32 ? solve_group<float>(c, kv.first, kv.second, o, sp)
                                                : solve_group<double>(c, kv.first, kv.second, o, sp);
        if (st) return st;
    }
    c->stats.algorithmic_bytes = (double)c->stats.amp_terms * (o.precision == MQW_FP32 ? 36.0 : 72.0);
    c->stats.required_bytes = c->stats.algorithmic_bytes -
                              (double)c->stats.psi_const_amp_terms * (o.precision == MQW_FP32 ? 8.0 : 16.0);
    for (Problem& p : probs_in) {
        double s = 0.0;
        for (int j = 0; j < o.samples; j++) s += sp[p.prob_base + j];
        prob[p.out] = (float)(s / o.samples);
        if (total_t) total_t[p.out] = p.sch.total_t;
        if (o.sample_probs)
            for (int j = 0; j < o.samples; j++)
                o.sample_probs[(size_t)p.out * o.samples + j] = (float)sp[p.prob_base + j];
    }
    return MQW_OK;
}

mqw_status prepare_problem(mqw_ctx* c, Problem& p, int n, int m, const double* rec, const float* gamma,
                           const float* short_t, int samples, int fixed_times)
{
    if (n < MQW_MIN_N || n > MQW_MAX_N) return fail(c, MQW_EINVAL, "n=%d outside [%d, %d]", n, MQW_MIN_N, MQW_MAX_N);
    if (m < 1 || m > MQW_MAX_STAGES) return fail(c, MQW_EINVAL, "m=%d outside [1, %d]", m, MQW_MAX_STAGES);
    if (!rec) return fail(c, MQW_EINVAL, "null record");
    if (!finite_rec(n, rec)) return fail(c, MQW_EINVAL, "record holds NaN or Inf");
    p.n = n;
    p.m = m;
    p.rec = rec;
    unpack(n, rec, p.J, p.h);
    schedule_heuristic(n, m, p.J, p.h, &p.sch);
    if (gamma && short_t) {
        float tot = 0;
        for (int i = 0; i < m; i++) {
            if (!isfinite(gamma[i]) || !(short_t[i] > 0) || !isfinite(short_t[i]))
                return fail(c, MQW_EINVAL, "schedule override: gamma finite, times positive and finite");
            p.sch.gamma[i] = gamma[i];
            p.sch.short_t[i] = short_t[i];
            tot += short_t[i];
        }
        p.sch.total_t = tot;
    }
    p.times.assign((size_t)m * samples, 0.0f);
    if (fixed_times) {
        // every sample runs each stage for exactly short_t (a deterministic schedule)
        for (int i = 0; i < m; i++)
            for (int j = 0; j < samples; j++) p.times[(size_t)i * samples + j] = p.sch.short_t[i];
    } else {
        sample_times(m, samples, p.sch.short_t, p.times.data());
    }
    return MQW_OK;
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int mqw_abi_version(void) { return MQW_ABI_VERSION; }

const char* mqw_status_string(mqw_status s)
{
    switch (s) {
    case MQW_OK: return "ok";
    case MQW_EINVAL: return "invalid argument";
    case MQW_ENOMEM: return "out of memory";
    case MQW_EIO: return "i/o error";
    case MQW_EDEVICE: return "device error";
    case MQW_ENOTSUP: return "not supported";
    }
    return "unknown status";
}

void mqw_opts_default(mqw_opts* o)
{
    memset(o, 0, sizeof *o);
    o->samples = MQW_DEFAULT_SAMPLES;
    o->tol = MQW_DEFAULT_TOL;
    o->precision = MQW_FP32;
    o->energy_mode = MQW_ENERGY_COMPAT;
    o->tile_bits = 0;
    o->chunk_bits = -1;
    o->small_n_max = -1;
    o->use_graph = 1;
    o->max_device_bytes = 0;
    o->group_bytes = 0;
    o->variant = -1;
    o->high_pass = -1;
    o->shared_first_stage = 1;
    o->final_readout = 1;
    o->shared_second_stage = 1;
    o->overlap_bases = 1;
    o->mixer_kind = MQW_MIXER_XSTRING;
    o->second_stage_gemm = MQW_GEMM_SPLIT_FP16;
    o->compress_bases = 1;
    o->direct_m1 = 1;
    o->adjoint_two_stages = 1;
    o->shared_third_stage = 1;
    o->chain_recurrences = 1;
    o->chain_wait_polls = 0;
    o->sample_probs = nullptr;
}

mqw_status mqw_create(int device, mqw_ctx** out)
{
    if (!out) return MQW_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        hipGetLastError();
        return MQW_EDEVICE;
    }
    if (device < 0 || device >= count) return MQW_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MQW_EDEVICE;
    std::unique_ptr<mqw_ctx> c(new mqw_ctx());
    c->device = device;
    for (int i = 0; i <= MQW_MAX_N; i++) {
        c->tuned_tb[i] = 0;
        c->tuned_lc[i] = -1;
        c->tuned_var[i] = 0;
        c->tuned_hp[i] = -1;
    }
    // this device's copy of the no-allocation scratch (mqw_clenshaw_noalloc);
    // a failed lookup is reported by that entry point, not here
    (void)noalloc_scratch_init(device);
    // the on-chip recurrences take one workgroup per CU of a 256-CU (8 x 32) part
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus == kChainWorkgroups && hipMalloc((void**)&c->chain_err, sizeof(uint32_t)) == hipSuccess) {
        c->chain_capable = hipMemset(c->chain_err, 0, sizeof(uint32_t)) == hipSuccess;
    }
    hipGetLastError();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return MQW_EDEVICE;
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        return MQW_EDEVICE;
    }
    *out = c.release();
    return MQW_OK;
}

void mqw_destroy(mqw_ctx* c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    release_bufs(c);
    if (c->chain_err) (void)hipFree(c->chain_err);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->stream2) hipStreamDestroy(c->stream2);
    delete c;
}

const char* mqw_last_error(const mqw_ctx* c) { return c ? c->err.c_str() : "no context"; }

mqw_status mqw_last_stats(const mqw_ctx* c, mqw_stats* out)
{
    if (!c || !out) return MQW_EINVAL;
    *out = c->stats;
    return MQW_OK;
}

mqw_status mqw_unpack(int n, const double* rec, float* J, float* h)
{
    if (n < 2 || n > MQW_MAX_N || !rec || !J || !h) return MQW_EINVAL;
    unpack(n, rec, J, h);
    return MQW_OK;
}

mqw_status mqw_schedule_heuristic(int n, int m, const double* rec, mqw_schedule* out)
{
    if (n < MQW_MIN_N || n > MQW_MAX_N || m < 1 || m > MQW_MAX_STAGES || !rec || !out) return MQW_EINVAL;
    if (!finite_rec(n, rec)) return MQW_EINVAL;
    float J[kMaxN * kMaxN], h[kMaxN];
    unpack(n, rec, J, h);
    schedule_heuristic(n, m, J, h, out);
    return MQW_OK;
}

mqw_status mqw_schedule_heuristic_moments(int n, int m, const double* rec, double skewness, double excess_kurtosis,
                                          mqw_schedule* out)
{
    if (n < MQW_MIN_N || n > MQW_MAX_N || m < 1 || m > MQW_MAX_STAGES || !rec || !out) return MQW_EINVAL;
    if (!finite_rec(n, rec) || !std::isfinite(skewness) || !std::isfinite(excess_kurtosis)) return MQW_EINVAL;
    float J[kMaxN * kMaxN], h[kMaxN];
    unpack(n, rec, J, h);
    const double shape[2] = {skewness, excess_kurtosis};
    schedule_heuristic(n, m, J, h, out, shape);
    return MQW_OK;
}

mqw_status mqw_sample_times(int m, int samples, const float* short_t, float* times)
{
    if (m < 1 || m > MQW_MAX_STAGES || samples < 1 || !short_t || !times) return MQW_EINVAL;
    sample_times(m, samples, short_t, times);
    return MQW_OK;
}

mqw_status mqw_chebyshev_coeffs(double s, double tol, double* cf, int cap, int* K)
{
    if (!cf || !K || cap < 1 || !(tol > 0) || !isfinite(s) || s < 0) return MQW_EINVAL;
    const int k = cheb_coeffs(s, tol, cf, cap);
    if (k < 0) return MQW_EINVAL;
    *K = k;
    return MQW_OK;
}

mqw_status mqw_generate_sk(int n, int count, uint64_t seed, double* out)
{
    if (n < 2 || n > MQW_MAX_N || count < 0 || (!out && count > 0)) return MQW_EINVAL;
    generate_sk(n, count, seed, out);
    return MQW_OK;
}

mqw_status mqw_energy(mqw_ctx* c, int n, const double* rec, int mode, float* hp, double* e, mqw_energy_info* info)
{
    if (!c) return MQW_EINVAL;
    if (n < MQW_MIN_N || n > MQW_MAX_N) return fail(c, MQW_EINVAL, "n=%d outside [%d, %d]", n, MQW_MIN_N, MQW_MAX_N);
    if (!rec || !info) return fail(c, MQW_EINVAL, "null argument");
    if (mode != MQW_ENERGY_COMPAT && mode != MQW_ENERGY_EXACT) return fail(c, MQW_EINVAL, "bad energy mode");
    if (!finite_rec(n, rec)) return fail(c, MQW_EINVAL, "record holds NaN or Inf");
    hipSetDevice(c->device);
    Problem p;
    p.n = n;
    p.m = 1;
    unpack(n, rec, p.J, p.h);
    std::vector<Problem*> ps{&p};
    const size_t N = (size_t)1 << n;
    float* hpd = (float*)dev_get(c, "energy_hp", sizeof(float) * N);
    if (!hpd) return fail(c, MQW_ENOMEM, "device allocation failed for the diagonal");
    double* e64 = nullptr;
    if (e) {
        e64 = (double*)dev_get(c, "energy_e64", sizeof(double) * N);
        if (!e64) return fail(c, MQW_ENOMEM, "device allocation failed for FP64 energies");
    }
    // FP32 diagonal always; FP64 energies unshifted when requested.
    const size_t nb = energy_blocks(n);
    float* Jh = (float*)dev_get(c, "energy_Jh", sizeof(float) * (size_t)(n * n + n));
    double* scr = (double*)dev_get(c, "energy_scratch", sizeof(double) * 4 * nb);
    EnergyDev* infod = (EnergyDev*)dev_get(c, "energy_info", sizeof(EnergyDev));
    if (!Jh || !scr || !infod) return fail(c, MQW_ENOMEM, "device allocation failed in energy build");
    std::vector<float> hJh((size_t)(n * n + n));
    memcpy(hJh.data(), p.J, sizeof(float) * n * n);
    memcpy(hJh.data() + n * n, p.h, sizeof(float) * n);
    HIPCHK(c, hipMemcpyAsync(Jh, hJh.data(), sizeof(float) * hJh.size(), hipMemcpyHostToDevice, c->stream));
    EnergyScratch es{scr, scr + nb, (uint32_t*)(scr + 2 * nb), scr + 3 * nb};
    const double e_init = mode == MQW_ENERGY_EXACT ? e_init_exact(n, p.J, p.h) : e_init_compat(n, p.J, p.h);
    if (launch_energy(n, Jh, Jh + n * n, e_init, mode == MQW_ENERGY_EXACT, hpd, e64, 0, es, infod, c->stream))
        return fail(c, MQW_EDEVICE, "energy kernels: %s", last_kernel_error());
    EnergyDev hi;
    HIPCHK(c, hipMemcpyAsync(&hi, infod, sizeof hi, hipMemcpyDeviceToHost, c->stream));
    if (hp) HIPCHK(c, hipMemcpyAsync(hp, hpd, sizeof(float) * N, hipMemcpyDeviceToHost, c->stream));
    if (e) HIPCHK(c, hipMemcpyAsync(e, e64, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    info->e0 = hi.e0;
    info->emax = hi.emax;
    info->eloc = hi.eloc;
    info->shift = hi.shift;
    info->e_abs = hi.e_abs;
    info->e_init = e_init;
    return MQW_OK;
}

// Moments of the walk's diagonal (the float energies H_P + shift): the
// skewness and kurtosis MultiQW.cpp:379-380 leaves commented out, from FP64
// power sums on the device in two passes (mean, then central sums).
mqw_status mqw_diag_moments(mqw_ctx* c, int n, const double* rec, int mode, double* out)
{
    if (!c) return MQW_EINVAL;
    if (!out) return fail(c, MQW_EINVAL, "null output");
    mqw_energy_info info;
    mqw_status st = mqw_energy(c, n, rec, mode, nullptr, nullptr, &info);
    if (st) return st;
    const size_t N = (size_t)1 << n;
    const float* hpd = (const float*)dev_get(c, "energy_hp", sizeof(float) * N);
    const int nb = diag_moment_blocks(N);
    double* part = (double*)dev_get(c, "moments_part", sizeof(double) * 4 * (size_t)nb);
    if (!hpd || !part) return fail(c, MQW_ENOMEM, "device allocation failed for moments");
    std::vector<double> hpart((size_t)4 * nb);
    auto sums = [&](double mu, double s[4]) -> mqw_status {
        if (diag_power_sums(hpd, N, (double)info.shift, mu, part, c->stream))
            return fail(c, MQW_EDEVICE, "moments kernel: %s", last_kernel_error());
        HIPCHK(c, hipMemcpyAsync(hpart.data(), part, sizeof(double) * hpart.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < 4; k++) s[k] = 0.0;
        for (int b = 0; b < nb; b++)
            for (int k = 0; k < 4; k++) s[k] += hpart[(size_t)b * 4 + k];
        return MQW_OK;
    };
    double s[4];
    if ((st = sums(0.0, s))) return st;
    const double mean = s[0] / (double)N;
    if ((st = sums(mean, s))) return st;
    const double var = s[1] / (double)N;
    out[0] = mean;
    out[1] = var;
    out[2] = var > 0 ? (s[2] / (double)N) / (var * sqrt(var)) : 0.0;
    out[3] = var > 0 ? (s[3] / (double)N) / (var * var) - 3.0 : 0.0;
    return MQW_OK;
}

mqw_status mqw_features(mqw_ctx* c, const mqw_instance* inst, int count, float* out)
{
    if (!c) return MQW_EINVAL;
    if (count < 0 || (count > 0 && (!inst || !out))) return fail(c, MQW_EINVAL, "bad count or null output");
    hipSetDevice(c->device);
    constexpr int kBins = 24;
    std::vector<Problem> ps((size_t)count);
    std::map<int, std::vector<Problem*>> groups;
    for (int k = 0; k < count; k++) {
        const mqw_instance& in = inst[k];
        if (in.n < MQW_MIN_N || in.n > MQW_MAX_N || in.m < 1 || in.m > MQW_MAX_STAGES || !in.rec)
            return fail(c, MQW_EINVAL, "instance %d: bad n, m or record", k);
        if (!finite_rec(in.n, in.rec)) return fail(c, MQW_EINVAL, "instance %d holds NaN or Inf", k);
        ps[k].n = in.n;
        ps[k].m = in.m;
        ps[k].rec = in.rec;
        ps[k].out = k;
        unpack(in.n, in.rec, ps[k].J, ps[k].h);
        groups[in.n].push_back(&ps[k]);
    }
    std::vector<uint32_t> counts((size_t)count * kBins);
    for (auto& kv : groups) {
        const int n = kv.first;
        auto& g = kv.second;
        const size_t N = (size_t)1 << n;
        float* E = (float*)dev_get(c, "feat_E", sizeof(float) * N * g.size());
        uint32_t* cnt = (uint32_t*)dev_get(c, "feat_counts", sizeof(uint32_t) * kBins * g.size());
        if (!E || !cnt) return fail(c, MQW_ENOMEM, "device allocation failed for features at n=%d", n);
        mqw_status st = group_energies(c, n, g, MQW_ENERGY_COMPAT, MQW_FP32, E);
        if (st) return st;
        HIPCHK(c, hipMemsetAsync(cnt, 0, sizeof(uint32_t) * kBins * g.size(), c->stream));
        for (size_t i = 0; i < g.size(); i++)
            if (diag_histogram(E + i * N, N, g[i]->info.e_abs, kBins, cnt + i * kBins, c->stream))
                return fail(c, MQW_EDEVICE, "histogram launch failed");
        std::vector<uint32_t> hc(kBins * g.size());
        HIPCHK(c, hipMemcpyAsync(hc.data(), cnt, sizeof(uint32_t) * hc.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < g.size(); i++)
            std::copy(hc.begin() + i * kBins, hc.begin() + (i + 1) * kBins, counts.begin() + (size_t)g[i]->out * kBins);
    }
    for (int k = 0; k < count; k++) {
        const Problem& p = ps[k];
        const int n = p.n;
        float* f = out + (size_t)k * MQW_FEATURES;
        double js = 0, jq = 0, jm = 0, hs = 0, hq = 0, hm = 0;
        int np = 0;
        for (int i = 1; i < n; i++)
            for (int j = 0; j < i; j++) {
                const double t = 2.0 * (double)p.J[i * n + j]; // the file's coupling (MultiQW.cpp:318-319 halves it)
                js += t;
                jq += t * t;
                jm = std::max(jm, fabs(t));
                np++;
            }
        for (int i = 0; i < n; i++) {
            const double v = p.h[i];
            hs += v;
            hq += v * v;
            hm = std::max(hm, fabs(v));
        }
        const double jmean = js / np, hmean = hs / n;
        f[0] = (float)(n / 20.0);
        f[1] = (float)(p.m / 5.0);
        f[2] = (float)jmean;
        f[3] = (float)(jq / np - jmean * jmean);
        f[4] = (float)jm;
        f[5] = (float)hmean;
        f[6] = (float)(hq / n - hmean * hmean);
        f[7] = (float)hm;
        const double N = (double)((size_t)1 << n);
        for (int b = 0; b < kBins; b++) f[8 + b] = (float)(counts[(size_t)k * kBins + b] / N);
    }
    return MQW_OK;
}

mqw_status mqw_clenshaw_step_dev(mqw_ctx* c, const float* b1, float* b2, const float* hp, const float* psi,
                                 unsigned n, float scale, float gamma, float coef)
{
    if (!c) return MQW_EINVAL;
    if ((int)n < MQW_MIN_N || (int)n > MQW_MAX_N) return fail(c, MQW_EINVAL, "n=%u outside range", n);
    if (!b1 || !b2 || !hp || (!psi && coef != 0.0f)) return fail(c, MQW_EINVAL, "null device pointer");
    if (!(scale != 0.0f) || !isfinite(scale) || !isfinite(gamma))
        return fail(c, MQW_EINVAL, "scale must be finite and nonzero");
    hipSetDevice(c->device);
    const float alpha = (float)(2.0 / (double)scale), beta = (float)(2.0 * (double)gamma / (double)scale);
    if (launch_half_step(b1, b2, hp, psi, (int)n, alpha, beta, coef, c->stream))
        return fail(c, MQW_EDEVICE, "half step: %s", last_kernel_error());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MQW_OK;
}

mqw_status mqw_clenshaw(mqw_ctx* c, int n, const float* coeffs, int K, float* psi, const float* hp, float gamma,
                        float scale, int psi_real)
{
    (void)psi_real; // the imaginary injections of a real state are zero either way
    if (!c) return MQW_EINVAL;
    if (n < MQW_MIN_N || n > MQW_MAX_N) return fail(c, MQW_EINVAL, "n=%d outside range", n);
    if (!coeffs || K < 1 || !psi || !hp) return fail(c, MQW_EINVAL, "null argument or K < 1");
    if (!(scale > 0.0f) || !isfinite(scale) || !isfinite(gamma)) return fail(c, MQW_EINVAL, "bad scale");
    hipSetDevice(c->device);
    const size_t N = (size_t)1 << n;
    const SweepCfg k = choose_sweep(c, MQW_FP32, n, 0, -1, -1, 0);
    const uint32_t ntiles = k.parts;
    float* E = (float*)dev_get(c, "cl_E", sizeof(float) * N);
    float* state = (float*)dev_get(c, "cl_state", sizeof(float) * 8 * N);
    double* np = (double*)dev_get(c, "cl_norm", sizeof(double) * 2 * ntiles);
    if (!E || !state || !np) return fail(c, MQW_ENOMEM, "device allocation failed");
    HIPCHK(c, hipMemcpyAsync(E, hp, sizeof(float) * N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(state, psi, sizeof(float) * 2 * N, hipMemcpyHostToDevice, c->stream));
    // one trajectory, one stage, input uploaded to `state`, buffers after it
    Traj tr;
    tr.prob = 0;
    tr.psi_in = state;
    Segment sg;
    sg.gamma = gamma;
    sg.onenorm = scale;
    sg.c.assign(coeffs, coeffs + K);
    sg.prob_slot = 0; // a readout gives the reduce job that names the output buffer
    tr.segs.push_back(sg);
    Slot<float> sl;
    for (int k = 0; k < 3; k++) sl.st[k] = state + (size_t)(2 + 2 * k) * N;
    sl.part[0] = np;
    sl.part[1] = np + ntiles;
    double* dummy = (double*)dev_get(c, "cl_prob", sizeof(double));
    if (!dummy) return fail(c, MQW_ENOMEM, "device allocation failed");
    Plan<float> plan;
    ReduceJob fin;
    bool has = false;
    traj_jobs<float>(n, ntiles, tr, sl, E, 0, dummy, plan.jobs, fin, has);
    for (size_t q = 0; q < plan.jobs.size(); q++) plan.ops.push_back(Op{0, q, 1});
    plan.rjobs.push_back(fin);
    plan.ops.push_back(Op{1, 0, 1});
    double ms = 0, ms_aux = 0;
    mqw_status st = run_plan<float>(c, MQW_FP32, n, k, plan, false, &ms, &ms_aux);
    if (st) return st;
    const float* outp = (const float*)plan.rjobs.back().out;
    HIPCHK(c, hipMemcpyAsync(psi, outp, sizeof(float) * 2 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MQW_OK;
}

mqw_status mqw_clenshaw_noalloc(mqw_ctx* c, int n, const float* coeffs, int K, float* psi, const float* hp,
                                float gamma, float scale, int psi_real)
{
    // mqw_clenshaw's jobs and kernels, with every device buffer carved from
    // the fixed per-device scratch (mqw_noalloc.hip) instead of the context's
    // workspace: no device allocation, no graph, no events.
    (void)psi_real;
    if (!c) return MQW_EINVAL;
    if (n < MQW_MIN_N || n > MQW_MAX_N) return fail(c, MQW_EINVAL, "n=%d outside range", n);
    if (n > kNoAllocMaxN)
        return fail(c, MQW_EINVAL, "n=%d exceeds the no-allocation limit n <= %d (%zu-byte device scratch)", n,
                    kNoAllocMaxN, kNoAllocScratchBytes);
    if (!coeffs || K < 1 || !psi || !hp) return fail(c, MQW_EINVAL, "null argument or K < 1");
    if (!(scale > 0.0f) || !isfinite(scale) || !isfinite(gamma)) return fail(c, MQW_EINVAL, "bad scale");
    char* base = (char*)noalloc_scratch(c->device);
    if (!base) {
        const hipError_t e = noalloc_scratch_init(c->device);
        return fail(c, MQW_EDEVICE, "no-allocation scratch unavailable on device %d: %s", c->device,
                    hipGetErrorString(e));
    }
    hipSetDevice(c->device);
    const size_t N = (size_t)1 << n;
    const SweepCfg k = choose_sweep(c, MQW_FP32, n, 0, -1, -1, 0);
    const uint32_t ntiles = k.parts;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        char* p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    float* E = (float*)carve(sizeof(float) * N);
    float* state = (float*)carve(sizeof(float) * 8 * N);
    double* np = (double*)carve(sizeof(double) * 2 * ntiles);
    double* dummy = (double*)carve(sizeof(double));
    Traj tr;
    tr.prob = 0;
    tr.psi_in = state;
    Segment sg;
    sg.gamma = gamma;
    sg.onenorm = scale;
    sg.c.assign(coeffs, coeffs + K);
    sg.prob_slot = 0;
    tr.segs.push_back(sg);
    Slot<float> sl;
    for (int q = 0; q < 3; q++) sl.st[q] = state + (size_t)(2 + 2 * q) * N;
    sl.part[0] = np;
    sl.part[1] = np + ntiles;
    thread_local std::vector<TermJob<float>> jobs; // host table, reused across calls on this thread
    jobs.clear();
    ReduceJob fin{};
    bool has = false;
    traj_jobs<float>(n, ntiles, tr, sl, E, 0, dummy, jobs, fin, has);
    TermJob<float>* jd = (TermJob<float>*)carve(sizeof(TermJob<float>) * jobs.size());
    ReduceJob* rd = (ReduceJob*)carve(sizeof(ReduceJob));
    if (off > kNoAllocScratchBytes)
        return fail(c, MQW_EINVAL, "n=%d with K=%d needs %zu bytes of device scratch; the no-allocation limit is %zu",
                    n, K, off, kNoAllocScratchBytes);
    HIPCHK(c, hipMemcpyAsync(E, hp, sizeof(float) * N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(state, psi, sizeof(float) * 2 * N, hipMemcpyHostToDevice, c->stream));
    if (!jobs.empty())
        HIPCHK(c, hipMemcpyAsync(jd, jobs.data(), sizeof(TermJob<float>) * jobs.size(), hipMemcpyHostToDevice,
                                 c->stream));
    HIPCHK(c, hipMemcpyAsync(rd, &fin, sizeof(ReduceJob), hipMemcpyHostToDevice, c->stream));
    for (size_t q = 0; q < jobs.size(); q++) {
        const int rc = k.half ? launch_half_term(k.tb, k.variant, jd + q, 1, n, k.lc, k.hi, c->stream)
                              : launch_term(MQW_FP32, k.tb, k.variant, jd + q, 1, n, k.lc, c->stream);
        if (rc) return fail(c, MQW_EDEVICE, "term launch: %s", last_kernel_error());
    }
    if (launch_reduce(MQW_FP32, rd, 1, c->stream)) return fail(c, MQW_EDEVICE, "reduce: %s", last_kernel_error());
    HIPCHK(c, hipMemcpyAsync(psi, fin.out, sizeof(float) * 2 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MQW_OK;
}

mqw_status mqw_solve(mqw_ctx* c, int n, int m, const double* recs, int count, const mqw_opts* opts, float* prob,
                     float* total_t)
{
    if (!c) return MQW_EINVAL;
    if (count < 0 || (count > 0 && (!recs || !prob))) return fail(c, MQW_EINVAL, "bad count or null output");
    mqw_opts o;
    if (opts)
        o = *opts;
    else
        mqw_opts_default(&o);
    if (mqw_status st = check_opts(c, o)) return st;
    hipSetDevice(c->device);
    const int P = n * (n + 1) / 2;
    std::vector<Problem> ps((size_t)count);
    for (int k = 0; k < count; k++) {
        mqw_status st = prepare_problem(c, ps[k], n, m, recs + (size_t)k * P, nullptr, nullptr, o.samples, o.fixed_times);
        if (st) return st;
        ps[k].out = k;
    }
    if (count == 0) return MQW_OK;
    return solve_problems(c, ps, o, prob, total_t);
}

mqw_status mqw_solve_batch(mqw_ctx* c, const mqw_instance* inst, int count, const mqw_opts* opts, float* prob,
                           float* total_t)
{
    if (!c) return MQW_EINVAL;
    if (count < 0 || (count > 0 && (!inst || !prob))) return fail(c, MQW_EINVAL, "bad count or null output");
    mqw_opts o;
    if (opts)
        o = *opts;
    else
        mqw_opts_default(&o);
    if (mqw_status st = check_opts(c, o)) return st;
    hipSetDevice(c->device);
    std::vector<Problem> ps((size_t)count);
    for (int k = 0; k < count; k++) {
        mqw_status st = prepare_problem(c, ps[k], inst[k].n, inst[k].m, inst[k].rec, inst[k].gamma, inst[k].short_t,
                                        o.samples, o.fixed_times);
        if (st) return st;
        ps[k].out = k;
    }
    if (count == 0) return MQW_OK;
    return solve_problems(c, ps, o, prob, total_t);
}

} // extern "C"
