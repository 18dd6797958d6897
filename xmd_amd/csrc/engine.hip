// engine.hip -- host side of the MI355X engine and its C ABI
// (include/xmd_hip.h).  Replaces the functor wiring and the OpenMP step
// loop of xmd/src/freeform/main.cpp:141-1077 with device buffers, the
// kernels of kernels.hpp and captured HIP graphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xmd_hip.h"
#include "kernels.hpp"

using namespace xmd_k;

namespace {

thread_local std::string g_create_err;

struct Fail {
    xmd_status st;
    std::string msg;
};

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw Fail{XMD_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_) + " (line " +    \
                                        std::to_string(__LINE__) + ")"};                            \
    } while (0)

// ---------------------------------------------------------------- GF(2) ---
// rand_gen's state update (rand_gen.h:24-26) is linear over GF(2); M holds
// the images of the 64 basis vectors (columns).
struct Mat {
    uint64_t col[64];
};
uint64_t apply(const Mat &M, uint64_t x) {
    uint64_t y = 0;
    for (int b = 0; b < 64; ++b)
        if ((x >> b) & 1ull) y ^= M.col[b];
    return y;
}
Mat mul(const Mat &A, const Mat &B) {
    Mat C;
    for (int b = 0; b < 64; ++b) C.col[b] = apply(A, B.col[b]);
    return C;
}
Mat identity() {
    Mat I;
    for (int b = 0; b < 64; ++b) I.col[b] = 1ull << b;
    return I;
}
Mat step_matrix() {
    Mat J;
    for (int b = 0; b < 64; ++b) J.col[b] = xmd_dev::rng_step(1ull << b);
    return J;
}
Mat mpow(Mat M, uint64_t k) {
    Mat R = identity();
    while (k) {
        if (k & 1) R = mul(M, R);
        M = mul(M, M);
        k >>= 1;
    }
    return R;
}
void rows_of(const Mat &M, uint64_t *rows) {
    for (int r = 0; r < 64; ++r) {
        uint64_t v = 0;
        for (int b = 0; b < 64; ++b)
            if ((M.col[b] >> r) & 1ull) v |= 1ull << b;
        rows[r] = v;
    }
}

// --------------------------------------------------- glibc fix table ---
// One byte per reachable draw value: corrections of logf / sinf / cosf
// (bm_math.hpp), and per bucket of 2^FIX_BUCKET_SHIFT draw values the
// margins below which a draw may need one (bead_noise).  Built once per
// process with the host's glibc.
struct FixTable {
    std::vector<uint8_t> tab;
    std::vector<float> thr;  // [bucket][log, sin, cos]
};
std::mutex g_fix_mu;
FixTable g_fix;

uint8_t fix_code(float mine, float ref) {
    if (mine == ref || (mine != mine && ref != ref)) return 0;
    if (ref == xmd_bm::next_up(mine)) return 1;
    if (ref == xmd_bm::next_down(mine)) return 2;
    throw Fail{XMD_ERR_STATE, "glibc libm differs from the engine's by more than one ulp"};
}

// a / spacing of boundary_margin (exact: the spacing is a power of two)
double margin_ratio(double y, float f) {
    double a, sp;
    boundary_margin(y, f, a, sp);
    return a / sp;
}

const FixTable &fix_table() {
    std::lock_guard<std::mutex> lk(g_fix_mu);
    if (!g_fix.tab.empty()) return g_fix;
    const uint32_t N = xmd_bm::NUM_DRAW_VALUES, NB = (N >> FIX_BUCKET_SHIFT) + 1;
    std::vector<uint8_t> tab(N);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    std::vector<int> bad(nt, 0);
    std::vector<std::vector<double>> mx(nt, std::vector<double>(3 * (size_t)NB, -1.0));
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            try {
                for (uint32_t k = t; k < N; k += nt) {
                    // the draw's float value v with draw_index(v) == k
                    float v;
                    if (k < 16777216u) v = (float)k;
                    else {
                        uint32_t q = k - 16777216u, e = q >> 23, m = q & 0x7fffffu;
                        v = xmd_bm::bitsf(((151u + e) << 23) | m);
                    }
                    const float x = v * (1.0f / 4294967296.0f);
                    const float two_pi_x = (float)(2.0 * M_PI) * x;
                    const double yl = bm_log_d(v);
                    double sd, cd;
                    bm_sincos_d(v, sd, cd);
                    const float fl = (float)yl, fs = (float)sd, fc = (float)cd;
                    const uint8_t cl = fix_code(fl, logf(x)), cs = fix_code(fs, sinf(two_pi_x)),
                                  cc = fix_code(fc, cosf(two_pi_x));
                    tab[k] = (uint8_t)(cl | (cs << 2) | (cc << 4));
                    double *m = mx[t].data() + 3 * (size_t)(k >> FIX_BUCKET_SHIFT);
                    if (cl) m[0] = std::max(m[0], margin_ratio(yl, fl));
                    if (cs) m[1] = std::max(m[1], margin_ratio(sd, fs));
                    if (cc) m[2] = std::max(m[2], margin_ratio(cd, fc));
                }
            } catch (...) {
                bad[t] = 1;
            }
        });
    }
    for (auto &x : th) x.join();
    for (int b : bad)
        if (b) throw Fail{XMD_ERR_STATE, "glibc libm differs from the engine's by more than one ulp"};
    std::vector<float> thr(3 * (size_t)NB);
    for (size_t j = 0; j < thr.size(); ++j) {
        double m = -1.0;
        for (unsigned t = 0; t < nt; ++t) m = std::max(m, mx[t][j]);
        float f = (float)m;
        if ((double)f < m) f = xmd_bm::next_up(f);  // rounded up: the test stays conservative
        thr[j] = f;
    }
    g_fix.tab.swap(tab);
    g_fix.thr.swap(thr);
    return g_fix;
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }

}  // namespace

// ================================================================ engine ===
// One part of an engine: device state and captured step graphs for a
// contiguous range of replicas, on its own HIP stream.
struct Part {
    xmd_params P;
    int nrep = 0, nbeads = 0, nslot = 0, nchains = 0;
    std::vector<int> rep_off, rep_slot_off, rep_n, slot_of_bead, bead_of_slot, slot_rep_h;
    Derived G;
    TermTab<double> TTd;
    TermTab<float> TTf;
    Dev D;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    std::vector<size_t> alloc_bytes;  // size of each allocation, in allocs' order
    int cur = 0;
    std::string err;
    int rep0 = 0;               // global index of the part's first replica
    std::vector<int> reported;  // error bits of each replica already reported
    // Graph launches queued on the stream and not yet known complete: at
    // most MAX_INFLIGHT, the host waits for the oldest before queueing
    // another, so one xmd_step call never holds more than that many
    // launches (16 steps each) in flight however many steps it asks for.
    static constexpr int MAX_INFLIGHT = 4;
    hipEvent_t ring[MAX_INFLIGHT] = {};
    int ring_head = 0, ring_n = 0;
    int nfp_chunks = 0, nct_chunks = 0, qa_grid = 0;
    uint32_t mask = 0;
    hipGraphExec_t g_chunk[2] = {nullptr, nullptr}, g_one[2] = {nullptr, nullptr};
    int chunk_steps = 16;
    int lpb = FORCE_LPB;  // lanes per bead of the force kernel
    // Use `l` lanes per bead from now on (graphs are captured again).
    void set_lpb(int l) {
        if (l == lpb) return;
        drop_graphs();
        lpb = l;
    }
    int64_t steps = 0, catchup_steps = 0;  // graph steps launched; those spent catching a replica up
    std::vector<char> pending_rb;
    std::vector<int> fp_base_h, fp_cap_h, ct_base_h, ct_cap_h;
    long long nl_total = 0;
    int fp_total = 0, ct_total = 0, nl_w = 0;
    // per-kernel event timing (eager profiling pass only)
    bool prof_on = false;
    std::vector<std::string> prof_names;
    std::vector<double> prof_ms;
    std::vector<int64_t> prof_n;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_pending;
    template <typename L> void launch(const char *name, L fn) {
        if (!prof_on) {
            fn();
            return;
        }
        auto ev = [&]() {
            if (ev_pool.empty()) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                return e;
            }
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        };
        hipEvent_t a = ev(), b = ev();
        HIPCHK(hipEventRecord(a, stream));
        fn();
        HIPCHK(hipEventRecord(b, stream));
        int idx = -1;
        for (size_t k = 0; k < prof_names.size(); ++k)
            if (prof_names[k] == name) idx = (int)k;
        if (idx < 0) {
            prof_names.push_back(name);
            prof_ms.push_back(0.0);
            prof_n.push_back(0);
            idx = (int)prof_names.size() - 1;
        }
        prof_pending.push_back({idx, {a, b}});
    }
    void prof_flush() {
        HIPCHK(hipStreamSynchronize(stream));
        for (auto &pe : prof_pending) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
            prof_ms[pe.first] += ms;
            prof_n[pe.first] += 1;
            ev_pool.push_back(pe.second.first);
            ev_pool.push_back(pe.second.second);
        }
        prof_pending.clear();
    }

    template <typename T> T *alloc(size_t n) {
        void *p = nullptr;
        if (n == 0) n = 1;
        HIPCHK(hipMalloc(&p, n * sizeof(T)));
        HIPCHK(hipMemsetAsync(p, 0, n * sizeof(T), stream));
        allocs.push_back(p);
        alloc_bytes.push_back(n * sizeof(T));
        return (T *)p;
    }
    // Host-to-device copies are synchronous: the source vectors are
    // temporaries and the stream may still hold earlier work on `d`.
    template <typename T> void put(T *d, const std::vector<T> &h) {
        if (h.empty()) return;
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    template <typename T> T *upload(const std::vector<T> &h) {
        T *d = alloc<T>(h.size());
        put(d, h);
        return d;
    }
    template <typename T> std::vector<T> get(const T *d, size_t n) {
        std::vector<T> h(n);
        if (n) HIPCHK(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return h;
    }
    ~Part() {
        // nothing queued may still use what is released below
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        for (auto e : ring)
            if (e) (void)hipEventDestroy(e);
        for (auto &g : g_chunk)
            if (g) (void)hipGraphExecDestroy(g);
        for (auto &g : g_one)
            if (g) (void)hipGraphExecDestroy(g);
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int nblk() const { return nslot / BLK; }
    size_t qa_lds() const { return (2 * (size_t)nrep + 1) * sizeof(int); }
    // one resident round of blocks of the grid-stride rebuild kernels (a
    // partial second round would leave most of the chip idle on a full rebuild)
    int grid_count = 0, grid_fill = 0;
    // the QA pass: one resident round of blocks striding over the live
    // chunks (the chunk capacity is about twice what is live, and idle
    // blocks still cost their dispatch and prefix scan)
    int grid_qa = 0;
    int qa_grid_for(int chunks) const { return std::max(1, std::min(chunks, grid_qa)); }
    bool fused_cells = false;  // every replica fits k_cells (CELLS_LDS cells, FUSED_MAX_BEADS beads)
    int nl_grid(bool fill = false) const { return std::min(nslot / NL_WAVES, fill ? grid_fill : grid_count); }

    // ------------------------------------------------------------ steps ---
    // Finalize (mode 1: a step's -- generator, time, verify decision; 0: the
    // decision from force_rb only) and the cell lists of the replicas that
    // rebuild: one launch when every replica fits k_cells.
    void cells_seq(const double *Ppos, int p, int mode) {
        if (fused_cells) {
            launch("k_cells", [&] { k_cells<<<nrep, 1024, 0, stream>>>(D, G, Ppos, p, mode); });
            return;
        }
        const int nb = nblk();
        launch("k_finalize", [&] { k_finalize<<<nrep, WAVE, 0, stream>>>(D, G, p, mode); });
        launch("k_cell_count", [&] { k_cell_count<<<nb, BLK, 0, stream>>>(D, Ppos, p); });
        launch("k_seg_scan", [&] { k_seg_scan<<<nrep, 1024, 0, stream>>>(D, p); });
        launch("k_cell_fill", [&] { k_cell_fill<<<nb, BLK, 0, stream>>>(D, p); });
    }
    // Rows and QA free pairs of the replicas that rebuild.
    void rows_seq(const double *Ppos, int p) {
        launch("k_nl_fill", [&] { k_nl_fill<<<nl_grid(true), BLK, 0, stream>>>(D, G, Ppos, p); });
        launch("k_fp_write", [&] { k_fp_write<<<nl_grid(), BLK, 0, stream>>>(D, p); });
    }
    void step_seq(int p) {
        if (G.qa_on) {
            launch("k_qa_pairs", [&] { k_qa_pairs<<<qa_grid, BLK, qa_lds(), stream>>>(D, G, p, 1); });
            launch("k_qa_serial", [&] { k_qa_serial<<<nrep, BLK, 0, stream>>>(D, p); });
        }
        launch("k_force", [&] {
            launch_force(P.precision == XMD_PREC_MIXED, false, lpb, nslot, stream, D, G, TTf, TTd, p, 1);
        });
        launch("k_integrate", [&] { k_integrate<<<nblk(), BLK, 0, stream>>>(D, G, p); });
        cells_seq(D.pos[1 - p], p, 1);
        rows_seq(D.pos[1 - p], p);
    }
    hipGraphExec_t capture(int p, int nsteps) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < nsteps; ++k) step_seq((p + k) & 1);
        HIPCHK(hipStreamEndCapture(stream, &g));
        hipGraphExec_t ex;
        HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
        return ex;
    }
    // Rebuilds the lists of the replicas marked pending (host-forced).
    void rebuild_pending() {
        bool any = false;
        for (char c : pending_rb) any |= c != 0;
        if (!any) return;
        std::vector<int> f(nrep);
        for (int k = 0; k < nrep; ++k) f[k] = pending_rb[k];
        put(D.force_rb, f);
        cells_seq(D.pos[cur], cur, 0);
        rows_seq(D.pos[cur], cur);
        std::fill(pending_rb.begin(), pending_rb.end(), 0);
    }
    void ensure_lists() {
        rebuild_pending();
        settle();
    }
    // After the stream has drained: grow the lists of replicas halted by a
    // rebuild that outgrew them (grow_lists), then report any error left.
    void settle() {
        while (grow_lists()) {
        }
        check_err();
    }
    // --------------------------------------------------- list growth ---
    // Replicas halted by a list rebuild that outgrew its capacity and by
    // nothing else -- a row longer than the row width (E_NL_CAP, with the
    // width it wanted) or more QA free pairs than the replica's region
    // holds (E_FP_CAP, with the count it found) -- have the state at the
    // start of the next step, with incomplete lists.  The rows are widened
    // in place of the old ones and the free-pair regions moved into larger
    // ones (the other replicas' lists keep their contents); those replicas
    // are rebuilt from the positions they halted at -- the lists an engine
    // created with the larger capacities built at that step -- and each
    // then steps alone, the others held, until it has taken as many steps
    // as they have.  False when there is nothing to grow (a row beyond
    // MAXSORT, or a free-pair overflow of the QA pass, stays an error).
    bool grow_lists() {
        HIPCHK(hipStreamSynchronize(stream));
        auto e = get(D.err, nrep);
        auto nw = get(D.need_w, nrep);
        auto nf = get(D.need_fp, nrep);
        std::vector<int> rec, cap2(fp_cap_h.begin(), fp_cap_h.begin() + nrep);
        int want = 0;
        bool fp_grow = false;
        for (int k = 0; k < nrep; ++k) {
            if (!e[k] || (e[k] & ~(E_NL_CAP | E_FP_CAP))) continue;
            if ((e[k] & E_NL_CAP) && nw[k] > MAXSORT) continue;
            if ((e[k] & E_FP_CAP) && nf[k] <= 0) continue;
            rec.push_back(k);
            if (e[k] & E_NL_CAP) want = std::max(want, nw[k]);
            if (e[k] & E_FP_CAP) {
                const long long c = std::max<long long>(nf[k] + nf[k] / 4LL, 2LL * fp_cap_h[k]);
                if (c >= (1LL << 30)) throw Fail{XMD_ERR_CAPACITY, "too many QA free pairs in one replica"};
                cap2[k] = round_up((int)c, BLK);
                fp_grow = true;
            }
        }
        if (rec.empty()) return false;
        if (want > 0) widen_rows(std::min(MAXSORT, round_up(std::max(want + want / 4, nl_w + WAVE), WAVE)));
        if (fp_grow) grow_free_pairs(cap2);
        put(D.need_w, std::vector<int>(nrep, 0));
        put(D.need_fp, std::vector<int>(nrep, 0));
        auto rb = get(D.n_rebuild, nrep);
        for (int k : rec) {
            clear_err(k);
            pending_rb[k] = 1;
            rb[k] -= 1;  // the rebuild that overflowed; the one below counts
        }
        put(D.n_rebuild, rb);
        rebuild_pending();
        for (int k : rec) catch_up(k);
        return true;
    }
    // Moves every replica's QA free pairs (pairs and vacancy flags) into
    // regions of `cap` entries each (multiples of BLK, none smaller than
    // before); the per-step candidate arrays start empty.
    void grow_free_pairs(const std::vector<int> &cap) {
        std::vector<int> base(nrep + 1), capv(nrep + 1);
        long long fp = 0;
        for (int k = 0; k < nrep; ++k) {
            base[k] = (int)fp;
            capv[k] = cap[k];
            fp += cap[k];
        }
        if (fp >= (1LL << 31)) throw Fail{XMD_ERR_CAPACITY, "too many QA free pairs for int32 offsets"};
        base[nrep] = capv[nrep] = (int)fp;
        int *i1 = alloc<int>(fp), *i2 = alloc<int>(fp);
        uint8_t *vac = alloc<uint8_t>(fp);
        for (int k = 0; k < nrep; ++k) {
            const size_t o = fp_base_h[k], n = fp_cap_h[k], b = base[k];
            HIPCHK(hipMemcpyAsync(i1 + b, D.fp_i1 + o, n * sizeof(int), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(i2 + b, D.fp_i2 + o, n * sizeof(int), hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(vac + b, D.fp_vac + o, n, hipMemcpyDeviceToDevice, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
        for (void *q : {(void *)D.fp_i1, (void *)D.fp_i2, (void *)D.fp_vac, (void *)D.cand_cnt, (void *)D.cand_fp,
                        (void *)D.cand_type, (void *)D.cand_d1, (void *)D.cand_d2, (void *)D.fp_chunk_rep})
            release(q);
        D.fp_i1 = i1;
        D.fp_i2 = i2;
        D.fp_vac = vac;
        D.cand_cnt = alloc<int>(fp / BLK);
        D.cand_fp = alloc<int>(fp);
        D.cand_type = alloc<int>(fp);
        D.cand_d1 = alloc<uint32_t>(fp);
        D.cand_d2 = alloc<uint32_t>(fp);
        std::vector<int> fprep(fp / BLK, -1);
        for (int k = 0; k < nrep; ++k)
            for (int c = base[k] / BLK; c < (base[k] + capv[k]) / BLK; ++c) fprep[c] = k;
        D.fp_chunk_rep = upload(fprep);
        fp_base_h = base;
        fp_cap_h = capv;
        put(const_cast<int *>(D.fp_base), base);
        put(const_cast<int *>(D.fp_cap), capv);
        fp_total = (int)fp;
        nfp_chunks = (int)(fp / BLK);
        qa_grid = qa_grid_for(nfp_chunks + nct_chunks);
        drop_graphs();  // they hold the old Dev
    }
    void widen_rows(int W2) {
        const int nsl = nslot / WAVE;
        if ((long long)nslot * W2 >= (1LL << 31)) throw Fail{XMD_ERR_CAPACITY, "rows too wide for int32 offsets"};
        uint32_t *ne = alloc<uint32_t>((size_t)nslot * W2);
        HIPCHK(hipMemcpy2DAsync(ne, (size_t)WAVE * W2 * sizeof(uint32_t), D.nl_e, (size_t)WAVE * nl_w * sizeof(uint32_t),
                                (size_t)WAVE * nl_w * sizeof(uint32_t), nsl, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        release(D.nl_e);
        D.nl_e = ne;
        D.nl_w = nl_w = W2;
        std::vector<int> ss(nsl);
        for (int sl = 0; sl < nsl; ++sl) ss[sl] = sl * WAVE * W2;
        put(D.slice_start, ss);
        nl_total = (long long)nslot * W2;
        drop_graphs();  // they hold the old Dev
    }
    // Steps replica k alone until it has taken as many steps as the engine
    // (every step of this part not spent catching a replica up); the
    // others are held meanwhile.
    void catch_up(int k) {
        auto done = get(D.n_done, nrep);
        auto e = get(D.err, nrep);
        long long lag = (steps - catchup_steps) - done[k];
        if (lag <= 0 || e[k]) return;
        std::vector<int> h = e;
        for (int j = 0; j < nrep; ++j)
            if (j != k && !e[j]) h[j] = E_HOLD;
        put(D.err, h);
        catchup_steps += lag;
        while (lag > 0) lag -= launch_next(lag);
        drain();
        auto after = get(D.err, nrep);
        for (int j = 0; j < nrep; ++j)
            if (h[j] == E_HOLD) after[j] = 0;
        put(D.err, after);
    }
    void release(void *p) {
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it != allocs.end()) {
            alloc_bytes.erase(alloc_bytes.begin() + (it - allocs.begin()));
            allocs.erase(it);
        }
        HIPCHK(hipFree(p));
    }
    // ------------------------------------------------------ snapshots ---
    // A copy of every device allocation of the part and of the host-side
    // step state, so that steps taken only to time them can be undone.  A
    // snapshot restores only into the allocations it copied (steps that
    // grow a list reallocate; the tuning steps do not grow lists).
    struct Snapshot {
        std::vector<void *> of, copy;
        int cur = 0;
        int64_t steps = 0, catchup_steps = 0;
        std::vector<int> reported, err;
        std::vector<char> pending_rb;
    };
    // (sn.of is set last: a snapshot that failed part-way restores nothing)
    void take(Snapshot &sn) {
        HIPCHK(hipStreamSynchronize(stream));
        sn.of.clear();
        sn.copy.assign(allocs.size(), nullptr);
        for (size_t k = 0; k < allocs.size(); ++k) {
            HIPCHK(hipMalloc(&sn.copy[k], alloc_bytes[k]));
            HIPCHK(hipMemcpyAsync(sn.copy[k], allocs[k], alloc_bytes[k], hipMemcpyDeviceToDevice, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
        sn.cur = cur;
        sn.steps = steps;
        sn.catchup_steps = catchup_steps;
        sn.reported = reported;
        sn.pending_rb = pending_rb;
        sn.err = get(D.err, nrep);
        sn.of = allocs;
    }
    void restore(const Snapshot &sn) {
        HIPCHK(hipStreamSynchronize(stream));
        if (sn.of != allocs) throw Fail{XMD_ERR_STATE, "device allocations changed since the snapshot"};
        for (size_t k = 0; k < allocs.size(); ++k)
            HIPCHK(hipMemcpyAsync(allocs[k], sn.copy[k], alloc_bytes[k], hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        cur = sn.cur;
        steps = sn.steps;
        catchup_steps = sn.catchup_steps;
        reported = sn.reported;
        pending_rb = sn.pending_rb;
        ring_head = ring_n = 0;
    }
    static void discard(Snapshot &sn) {
        for (void *q : sn.copy)
            if (q) (void)hipFree(q);
        sn.copy.clear();
        sn.of.clear();
    }
    // True when a replica that was running at the snapshot is halted now.
    bool halted_since(const Snapshot &sn) {
        auto e = get(D.err, nrep);
        for (int k = 0; k < nrep; ++k)
            if (e[k] && !sn.err[k]) return true;
        return false;
    }
    void drop_graphs() {
        HIPCHK(hipStreamSynchronize(stream));
        for (auto &g : g_chunk)
            if (g) {
                HIPCHK(hipGraphExecDestroy(g));
                g = nullptr;
            }
        for (auto &g : g_one)
            if (g) {
                HIPCHK(hipGraphExecDestroy(g));
                g = nullptr;
            }
    }
    // Error words: one per replica.  A kernel that finds an error sets its
    // replica's word; every kernel then leaves that replica as it is
    // (halted) until the host sets its state again (clear_err).  Each error
    // is reported once, by the first call that finds it, naming the replica.
    static std::string describe(int e, xmd_status &st) {
        if (e & E_ZERO_DRAW) {
            st = XMD_ERR_NONFINITE;
            return "rand_gen produced a zero uniform draw for log(): infinite Langevin noise, as in the reference "
                   "(rand_gen.h:45-46); positions are not finite";
        }
        if (e & E_NONFINITE) {
            st = XMD_ERR_NONFINITE;
            return "non-finite positions";
        }
        st = XMD_ERR_CAPACITY;
        std::string m = "device capacity exceeded:";
        if (e & E_NL_CAP) m += " neighbour-list capacity exceeded;";
        if (e & E_FP_CAP) m += " QA free-pair capacity exceeded;";
        if (e & E_CT_CAP) m += " QA contact capacity exceeded;";
        if (e & E_SLOT) m += " more QA contacts on a bead than slots;";
        return m;
    }
    void check_err() {
        auto e = get(D.err, nrep);
        xmd_status first = XMD_OK;
        std::string msg;
        for (int k = 0; k < nrep; ++k) {
            if (!e[k] || reported[k] == e[k]) continue;
            reported[k] = e[k];
            xmd_status st;
            const std::string d = describe(e[k], st);
            if (first == XMD_OK) first = st;
            msg += "replica " + std::to_string(rep0 + k) + ": " + d + " (replica halted until its state is set); ";
        }
        if (first != XMD_OK) throw Fail{first, msg};
    }
    // Refuses a call whose result would cover a halted replica.
    void require_running() {
        check_err();
        auto e = get(D.err, nrep);
        for (int k = 0; k < nrep; ++k)
            if (e[k]) {
                xmd_status st;
                throw Fail{XMD_ERR_STATE, "replica " + std::to_string(rep0 + k) + " is halted by an earlier error (" +
                                              describe(e[k], st) + "); set its state first"};
            }
    }
    void clear_err(int rep) {
        const int zero = 0;
        HIPCHK(hipMemcpy(D.err + rep, &zero, sizeof(int), hipMemcpyHostToDevice));
        reported[rep] = 0;
    }
    // A state set from the host: the replica is running again, level with
    // the engine's step count.
    void reset_replica(int rep) {
        clear_err(rep);
        const long long done = steps - catchup_steps;
        HIPCHK(hipMemcpy(D.n_done + rep, &done, sizeof(long long), hipMemcpyHostToDevice));
    }
    // Waits for the oldest queued graph launch when MAX_INFLIGHT are queued.
    void throttle() {
        if (ring_n < MAX_INFLIGHT) return;
        HIPCHK(hipEventSynchronize(ring[ring_head]));
        ring_head = (ring_head + 1) % MAX_INFLIGHT;
        --ring_n;
    }
    void mark_launch() {
        hipEvent_t &ev = ring[(ring_head + ring_n) % MAX_INFLIGHT];
        if (!ev) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(ev, stream));
        ++ring_n;
    }
    void drain() {
        HIPCHK(hipStreamSynchronize(stream));
        ring_head = ring_n = 0;
    }
    // Launches the graph for the next min(n, chunk_steps) steps (no wait);
    // returns how many steps it covers.
    int64_t launch_next(int64_t n) {
        throttle();
        if (n >= chunk_steps) {
            if (!g_chunk[cur]) g_chunk[cur] = capture(cur, chunk_steps);
            HIPCHK(hipGraphLaunch(g_chunk[cur], stream));
            mark_launch();
            if (chunk_steps & 1) cur ^= 1;
            steps += chunk_steps;
            return chunk_steps;
        }
        if (!g_one[cur]) g_one[cur] = capture(cur, 1);
        HIPCHK(hipGraphLaunch(g_one[cur], stream));
        mark_launch();
        cur ^= 1;
        steps += 1;
        return 1;
    }
};

namespace {

// Derived constants, main.cpp:345-907 expressions in double.
// go_cut: 2 * the longest native contact length over the whole topology
// (every stream group gets the same list radius).
void derive(Part &E, double go_cut) {
    const xmd_params &p = E.P;
    Derived &G = E.G;
    G.dt = p.dt;
    G.dt_inv = 1.0 / p.dt;
    G.half_dt2 = p.dt * p.dt / 2.0;
    G.gamma = p.gamma_factor * 1.0;
    G.lang_on = p.langevin_enabled != 0;
    G.H1 = p.tether_H1;
    G.H2 = p.tether_H2;
    double bb_rmin = p.lj_bb_rmin, bs_rmin = p.lj_bs_rmin;
    std::vector<double> ss_rmax(p.ss_rmax, p.ss_rmax + XMD_NUM_SS);
    double es_cut = p.const_es_enabled ? 2.0 * p.const_es_screen : 2.0 * p.rel_es_screen;
    double pid_cut = 0.0;
    pid_cut = std::max(pid_cut, 2.0 * bb_rmin);
    pid_cut = std::max(pid_cut, 2.0 * bs_rmin);
    pid_cut = std::max(pid_cut, 2.0 * bs_rmin);
    for (double r : ss_rmax) pid_cut = std::max(pid_cut, 2.0 * r);
    double mfmd = 0.0;
    mfmd = std::max(mfmd, bb_rmin);
    mfmd = std::max(mfmd, bs_rmin);
    for (double r : ss_rmax) mfmd = std::max(mfmd, r);
    double mc = 0.0;
    if (p.pauli_enabled) mc = std::max(mc, p.pauli_r_excl);
    if (p.go_enabled) mc = std::max(mc, go_cut);
    // each enabled variant adds its own cutoff (main.cpp:628,642); the
    // pair list keeps the constant variant's (D31)
    if (p.const_es_enabled) mc = std::max(mc, 2.0 * p.const_es_screen);
    if (p.rel_es_enabled) mc = std::max(mc, 2.0 * p.rel_es_screen);
    if (p.qa_enabled) mc = std::max(mc, mfmd);
    if (p.pid_enabled) mc = std::max(mc, pid_cut);
    // native_ss.cpp: harmonic::cutoff(equilibrium distance) = 2 * it
    G.ss_on = p.nat_ss_enabled;
    G.ss_H1 = p.nat_ss_H1;
    G.ss_equil = p.nat_ss_equil_dist;
    G.ss_cut = 2.0 * p.nat_ss_equil_dist;
    if (p.nat_ss_enabled) mc = std::max(mc, G.ss_cut);
    // walls/solid.cpp, walls/lj_attr.cpp (main.cpp:836-842), velocity_afm.cpp
    G.solid_eps = p.solid_walls_depth;
    G.solid_cut = p.solid_walls_cutoff;
    G.ljw_min = p.lj_walls_min_dist;
    G.ljw_factor = p.lj_walls_breaking_factor * std::pow(2.0, -1.0 / 6.0);
    G.ljw_cyc = p.lj_walls_phase_duration;
    G.ljw_cyc_inv = 1.0 / p.lj_walls_phase_duration;
    G.vafm_H1 = p.vel_afm_H1;
    G.vafm_H2 = p.vel_afm_H2;
    if (p.max_cutoff_override > 0) mc = p.max_cutoff_override;
    G.pad = p.pad_factor * mc;
    G.req_r = mc + G.pad;
    G.pauli_min_inv = 1.0 / (p.pauli_r_excl + G.pad);
    G.es_min_inv = 1.0 / (es_cut + G.pad);
    G.pid_min_inv = 1.0 / (pid_cut + G.pad);
    G.qa_min_inv = 1.0 / (mfmd + G.pad);
    G.pauli_on = p.pauli_enabled;
    G.es_on = p.const_es_enabled || p.rel_es_enabled;
    G.pid_on = p.pid_enabled;
    G.qa_on = p.qa_enabled;
    G.go_on = p.go_enabled;
    QaParams &Q = G.qa;
    Q.min_cos_hr = p.qa_min_cos_hr;
    Q.min_cos_hh = p.qa_min_cos_hh;
    Q.max_cos_nr = p.qa_max_cos_nr;
    Q.cycle_time = p.qa_phase_duration;
    Q.cycle_time_inv = 1.0 / Q.cycle_time;
    Q.factor = p.qa_breaking_factor * (double)powf(2.0f, -1.0f / 6.0f);
    Q.bb_depth = p.lj_bb_depth;
    Q.bb_rmin = bb_rmin;
    Q.bs_depth = p.lj_bs_depth;
    Q.bs_rmin = bs_rmin;
    Q.req_max = mfmd;

    auto fill = [&](auto &TT) {
        using T = std::remove_reference_t<decltype(TT.P.H1)>;
        TermParams<T> &Pt = TT.P;
        Pt.H1 = (T)p.tether_H1;
        Pt.H2 = (T)p.tether_H2;
        Pt.e_chi = (T)p.chir_e;
        Pt.ang_k = (T)p.nat_ang_k;
        Pt.CDA = (T)p.cnd_CDA;
        Pt.CDB = (T)p.cnd_CDB;
        Pt.CDH = (T)p.snd_CDH;
        Pt.pauli_r_excl = (T)p.pauli_r_excl;
        Pt.pauli_depth = (T)p.pauli_depth;
        Pt.go_depth = (T)p.go_depth;
        // main.cpp:618-642: V_factor = 1/(4 pi eps), screen_dist_inv per variant
        Pt.es_relative = 0;
        Pt.es_Vf = (T)(1.0 / (4.0 * M_PI * p.const_es_permittivity));
        Pt.es_sdi = (T)(1.0 / p.const_es_screen);
        Pt.res_Vf = (T)(1.0 / (4.0 * M_PI * p.rel_es_factor));
        Pt.res_sdi = (T)(1.0 / p.rel_es_screen);
        Pt.bbp_psi0 = (T)p.pid_bbp_psi0;
        Pt.bbp_alpha = (T)p.pid_bbp_alpha;
        Pt.bbp_rmin = (T)p.pid_bbp_rmin;
        Pt.bbp_depth = (T)p.pid_bbp_depth;
        Pt.bbm_psi0 = (T)p.pid_bbm_psi0;
        Pt.bbm_alpha = (T)p.pid_bbm_alpha;
        Pt.bbm_rmin = (T)p.pid_bbm_rmin;
        Pt.bbm_depth = (T)p.pid_bbm_depth;
        Pt.ss_psi0 = (T)p.pid_ss_psi0;
        Pt.ss_alpha = (T)p.pid_ss_alpha;
        Pt.lambda_version = p.pid_lambda_version;
        Pt.set_derived();
    };
    fill(E.TTd);
    fill(E.TTf);
    uint32_t m = 0;
    if (p.tether_enabled) m |= 1u << T_TETHER;
    if (p.chir_enabled) m |= 1u << T_CHIRAL;
    if (p.nat_ang_enabled) m |= 1u << T_NAT_ANG;
    if (p.cnd_enabled) m |= 1u << T_CND;
    if (p.snd_enabled) m |= 1u << T_SND;
    if (p.pauli_enabled) m |= 1u << T_PAULI;
    if (p.go_enabled) m |= 1u << T_GO;
    if (p.const_es_enabled) m |= 1u << T_CONST_ES;
    if (p.rel_es_enabled) m |= 1u << T_REL_ES;
    if (p.qa_enabled) m |= 1u << T_QA;
    if (p.pid_enabled) m |= 1u << T_PID;
    if (p.nat_ss_enabled) m |= 1u << T_NAT_SS;
    if (p.solid_walls_enabled) m |= 1u << T_SOLID_WALL;
    if (p.lj_walls_enabled) m |= 1u << T_LJ_WALL;
    if (p.vel_afm_enabled) m |= 1u << T_VEL_AFM;
    if (p.force_afm_enabled) m |= 1u << T_FORCE_AFM;
    if (p.heur_ang_enabled) m |= 1u << T_HEUR_ANG;
    if (p.heur_dih_enabled) m |= 1u << T_HEUR_DIH;
    E.mask = m;
    E.TTd.mask = m;
    E.TTf.mask = m;
}

void build(Part &E, const xmd_params *p, const xmd_topology *t, double go_cut) {
    if (!p || !t) throw Fail{XMD_ERR_ARG, "null params or topology"};
    E.P = *p;
    if (!(p->dt > 0)) throw Fail{XMD_ERR_ARG, "dt must be positive"};
    if (p->precision != XMD_PREC_FP64 && p->precision != XMD_PREC_MIXED) throw Fail{XMD_ERR_ARG, "bad precision"};
    if (p->force_lanes != 0 && p->force_lanes != 4 && p->force_lanes != 8 && p->force_lanes != 16)
        throw Fail{XMD_ERR_ARG, "force_lanes must be 0, 4, 8 or 16"};
    E.lpb = p->force_lanes ? p->force_lanes : FORCE_LPB;
    if (t->n_replicas <= 0 || t->n_beads <= 0) throw Fail{XMD_ERR_ARG, "empty topology"};
    if (t->replica_offset[0] != 0 || t->replica_offset[t->n_replicas] != t->n_beads)
        throw Fail{XMD_ERR_ARG, "replica_offset must span [0, n_beads]"};
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Fail{XMD_ERR_NO_DEVICE, "no HIP device"};
    HIPCHK(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
    {
        int dev = 0, ncu = 0, bc = 0, bf = 0, bq = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bc, k_nl_count, BLK, 0));
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bf, k_nl_fill, BLK, 0));
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bq, k_qa_pairs, BLK,
                                                            (2 * (size_t)std::max(t->n_replicas, 1) + 1) * sizeof(int)));
        E.grid_count = std::max(1, bc) * std::max(1, ncu);
        E.grid_fill = std::max(1, bf) * std::max(1, ncu);
        E.grid_qa = std::max(1, bq) * std::max(1, ncu);
    }
    E.nrep = t->n_replicas;
    E.nbeads = t->n_beads;
    E.nchains = t->n_chains;
    E.rep_off.assign(t->replica_offset, t->replica_offset + E.nrep + 1);
    E.rep_n.resize(E.nrep);
    E.rep_slot_off.resize(E.nrep + 1);
    int so = 0;
    for (int k = 0; k < E.nrep; ++k) {
        E.rep_n[k] = E.rep_off[k + 1] - E.rep_off[k];
        if (E.rep_n[k] <= 0) throw Fail{XMD_ERR_ARG, "empty replica"};
        E.rep_slot_off[k] = so;
        so += round_up(E.rep_n[k], WAVE);
    }
    E.rep_slot_off[E.nrep] = so;
    E.nslot = round_up(std::max(so, BLK), BLK);
    if (E.nslot > (int)NLE_JMASK)
        throw Fail{XMD_ERR_CAPACITY, "too many beads for one engine part (a row entry holds a slot in 26 bits)"};
    const int n = E.nslot;
    E.slot_of_bead.resize(E.nbeads);
    E.bead_of_slot.assign(n, -1);
    E.slot_rep_h.assign(n, 0);
    for (int k = 0; k < E.nrep; ++k) {
        int end = (k + 1 < E.nrep) ? E.rep_slot_off[k + 1] : n;
        for (int s = E.rep_slot_off[k]; s < end; ++s) E.slot_rep_h[s] = k;
        for (int i = 0; i < E.rep_n[k]; ++i) {
            E.slot_of_bead[E.rep_off[k] + i] = E.rep_slot_off[k] + i;
            E.bead_of_slot[E.rep_slot_off[k] + i] = E.rep_off[k] + i;
        }
    }
    // chains
    std::vector<int> chain_of(E.nbeads, -1);
    for (int c = 0; c < t->n_chains; ++c)
        for (int i = t->chain_offset[c]; i < t->chain_offset[c + 1]; ++i) {
            if (i < 0 || i >= E.nbeads) throw Fail{XMD_ERR_ARG, "chain_offset out of range"};
            chain_of[i] = c;
        }
    std::vector<int> chain_s(n, -1);
    std::vector<uint8_t> bfl(n, 0);
    std::vector<int8_t> at(n, 0);
    for (int g = 0; g < E.nbeads; ++g) {
        if (chain_of[g] < 0) throw Fail{XMD_ERR_ARG, "bead outside every chain"};
        if (t->atype[g] < 0 || t->atype[g] >= XMD_NUM_AA) throw Fail{XMD_ERR_ARG, "atype out of range"};
        int s = E.slot_of_bead[g];
        chain_s[s] = chain_of[g];
        at[s] = t->atype[g];
        uint8_t f = B_ACTIVE;
        if (g > 0 && chain_of[g - 1] == chain_of[g]) f |= B_PREV;
        if (g + 1 < E.nbeads && chain_of[g + 1] == chain_of[g]) f |= B_NEXT;
        int k = E.slot_rep_h[s];
        if (((f & B_PREV) && g - 1 < E.rep_off[k]) || ((f & B_NEXT) && g + 1 >= E.rep_off[k + 1]))
            throw Fail{XMD_ERR_ARG, "chain spans replicas"};
        // heuristic angle / dihedral entries of the model
        const uint8_t hb = t->heur_bonded ? t->heur_bonded[g] : 0;
        auto same_chain = [&](int m) { return g + m < E.nbeads && chain_of[g + m] == chain_of[g]; };
        if (hb & 1) {
            if (!same_chain(2)) throw Fail{XMD_ERR_ARG, "heuristic angle does not join consecutive residues"};
            f |= B_HANG;
        }
        if (hb & 2) {
            if (!same_chain(3)) throw Fail{XMD_ERR_ARG, "heuristic dihedral does not join consecutive residues"};
            f |= B_HDIH;
        }
        bfl[s] = f;
    }
    derive(E, go_cut);
    const xmd_params &P = E.P;
    // per-replica arrays
    std::vector<Box> boxes(E.nrep);
    std::vector<double> nsd(E.nrep), temp(E.nrep), t0(E.nrep, 0.0);
    std::vector<uint64_t> rng(E.nrep);
    std::vector<int> cell_off(E.nrep + 1, 0), cell_dim(3 * E.nrep, 0);
    int ncells = 0;
    for (int k = 0; k < E.nrep; ++k) {
        Box b{t->box[3 * k], t->box[3 * k + 1], t->box[3 * k + 2], 0, 0, 0};
        if (b.cx != 0.0) b.ix = 1.0 / b.cx;
        if (b.cy != 0.0) b.iy = 1.0 / b.cy;
        if (b.cz != 0.0) b.iz = 1.0 / b.cz;
        boxes[k] = b;
        temp[k] = t->temperature[k];
        // lang_pc.cpp:7-13 with mass 1
        double noise_factor = std::sqrt(2.0 * 1.0 * temp[k]);
        double gfs = std::sqrt(P.gamma_factor);
        nsd[k] = noise_factor * gfs * 1.0;
        rng[k] = t->rng_state[k];
        cell_off[k] = ncells;
        double side = E.G.req_r * (1.0 + 1e-9);
        int nx = b.cx > 0 ? (int)std::floor(b.cx / side) : 0;
        int ny = b.cy > 0 ? (int)std::floor(b.cy / side) : 0;
        int nz = b.cz > 0 ? (int)std::floor(b.cz / side) : 0;
        if (nx >= 3 && ny >= 3 && nz >= 3 && (long long)nx * ny * nz <= 64LL * E.rep_n[k] + 4096) {
            cell_dim[3 * k] = nx;
            cell_dim[3 * k + 1] = ny;
            cell_dim[3 * k + 2] = nz;
            ncells += nx * ny * nz;
        }
    }
    cell_off[E.nrep] = ncells;
    E.fused_cells = true;
    for (int k = 0; k < E.nrep; ++k)
        if (cell_off[k + 1] - cell_off[k] > CELLS_LDS || E.rep_n[k] > FUSED_MAX_BEADS) E.fused_cells = false;
    // native contacts: rows of both ends, sorted
    std::vector<std::vector<std::pair<int, double>>> nrow(n);
    for (int c = 0; c < t->n_native; ++c) {
        int a = t->nat_i1[c], b = t->nat_i2[c];
        if (a < 0 || b >= E.nbeads || !(a < b)) throw Fail{XMD_ERR_ARG, "native contact needs 0 <= i1 < i2 < n"};
        int sa = E.slot_of_bead[a], sb = E.slot_of_bead[b];
        if (E.slot_rep_h[sa] != E.slot_rep_h[sb]) throw Fail{XMD_ERR_ARG, "native contact crosses replicas"};
        nrow[sa].push_back({sb, t->nat_dist[c]});
        nrow[sb].push_back({sa, t->nat_dist[c]});
    }
    std::vector<int> nat_start(n + 1, 0), nat_j;
    std::vector<double> nat_d;
    for (int s = 0; s < n; ++s) {
        std::sort(nrow[s].begin(), nrow[s].end());
        nat_start[s] = (int)nat_j.size();
        for (auto &e : nrow[s]) {
            nat_j.push_back(e.first);
            nat_d.push_back(e.second);
        }
    }
    nat_start[n] = (int)nat_j.size();
    // disulfide bonds: rows of both ends, sorted; never a native contact
    std::vector<std::vector<int>> srow(n);
    for (int c = 0; c < t->n_ssbonds; ++c) {
        int a = t->ss_i1[c], b = t->ss_i2[c];
        if (a < 0 || b >= E.nbeads || !(a < b)) throw Fail{XMD_ERR_ARG, "ssbond needs 0 <= i1 < i2 < n"};
        int sa = E.slot_of_bead[a], sb = E.slot_of_bead[b];
        if (E.slot_rep_h[sa] != E.slot_rep_h[sb]) throw Fail{XMD_ERR_ARG, "ssbond crosses replicas"};
        for (auto &e : nrow[sa])
            if (e.first == sb) throw Fail{XMD_ERR_ARG, "pair is both a native contact and an ssbond"};
        srow[sa].push_back(sb);
        srow[sb].push_back(sa);
    }
    std::vector<int> ss_start(n + 1, 0), ss_j;
    for (int s = 0; s < n; ++s) {
        std::sort(srow[s].begin(), srow[s].end());
        ss_start[s] = (int)ss_j.size();
        ss_j.insert(ss_j.end(), srow[s].begin(), srow[s].end());
    }
    ss_start[n] = (int)ss_j.size();
    // walls (the same planes in every replica) and AFM tips by slot
    if (t->n_solid_walls < 0 || t->n_lj_walls < 0 || (t->n_solid_walls > 0 && !t->solid_walls) ||
        (t->n_lj_walls > 0 && !t->lj_walls))
        throw Fail{XMD_ERR_ARG, "bad wall arrays"};
    std::vector<double> solid_w(t->solid_walls, t->solid_walls + 6 * (size_t)t->n_solid_walls);
    std::vector<double> lj_w(t->lj_walls, t->lj_walls + 6 * (size_t)t->n_lj_walls);
    E.G.n_solid = t->n_solid_walls;
    E.G.n_ljw = t->n_lj_walls;
    std::vector<std::vector<int>> vtip(n), ftip(n);
    for (int c = 0; c < t->n_vel_afm; ++c) {
        int g = t->vel_afm_idx[c];
        if (g < 0 || g >= E.nbeads) throw Fail{XMD_ERR_ARG, "velocity AFM tip outside the beads"};
        vtip[E.slot_of_bead[g]].push_back(c);
    }
    for (int c = 0; c < t->n_force_afm; ++c) {
        int g = t->force_afm_idx[c];
        if (g < 0 || g >= E.nbeads) throw Fail{XMD_ERR_ARG, "force AFM tip outside the beads"};
        ftip[E.slot_of_bead[g]].push_back(c);
    }
    std::vector<int> vafm_start(n + 1, 0), fafm_start(n + 1, 0);
    std::vector<double> vafm_orig, vafm_v, fafm_F;
    for (int s = 0; s < n; ++s) {
        vafm_start[s] = (int)vafm_v.size() / 3;
        for (int c : vtip[s]) {
            const double *o = t->vel_afm_orig + 3 * c, *v = t->vel_afm_v + 3 * c;
            int g = t->vel_afm_idx[c];
            for (int d = 0; d < 3; ++d) {
                // no origin: the bead's initial position (D27)
                vafm_orig.push_back(std::isnan(o[0]) ? t->r0[(size_t)d * E.nbeads + g] : o[d]);
                vafm_v.push_back(v[d]);
            }
        }
        fafm_start[s] = (int)fafm_F.size() / 3;
        for (int c : ftip[s])
            for (int d = 0; d < 3; ++d) fafm_F.push_back(t->force_afm_F[3 * c + d]);
    }
    vafm_start[n] = (int)vafm_v.size() / 3;
    fafm_start[n] = (int)fafm_F.size() / 3;
    // bonded natives in slot space
    std::vector<double> teth(n, P.tether_def_length), th(n, NAN), ph(n, NAN), ch(n, NAN), cf(n, NAN);
    for (int g = 0; g < E.nbeads; ++g) {
        int s = E.slot_of_bead[g];
        if (t->tether_len && !std::isnan(t->tether_len[g])) teth[s] = t->tether_len[g];
        if (t->nat_theta) th[s] = t->nat_theta[g];
        if (t->nat_phi) ph[s] = t->nat_phi[g];
        if (t->nat_chir) ch[s] = t->nat_chir[g];
        if (t->nat_chir_factor) cf[s] = t->nat_chir_factor[g];
    }
    // positions and state
    std::vector<double> pos(4 * (size_t)n, 0.0);  // AoS (x, y, z, 0) per slot
    for (int g = 0; g < E.nbeads; ++g) {
        int s = E.slot_of_bead[g];
        for (int d = 0; d < 3; ++d) pos[4 * (size_t)s + d] = t->r0[d * E.nbeads + g];
    }
    // QA sync (D7)
    std::vector<uint32_t> sync(n, 0);
    for (int s = 0; s < n; ++s)
        if (bfl[s] & B_ACTIVE) {
            const int8_t *L = P.aa_limits[at[s]];
            sync[s] = xmd_dev::sync_pack(L[0], L[1], L[2], L[3]);
        }
    // tables
    std::vector<double> ss_depth(P.ss_depth, P.ss_depth + XMD_NUM_SS), ss_rmax(P.ss_rmax, P.ss_rmax + XMD_NUM_SS),
        ss_rmin(XMD_NUM_SS);
    for (int k = 0; k < XMD_NUM_SS; ++k) ss_rmin[k] = P.ss_sinking ? P.lj_bb_rmin : ss_rmax[k];
    std::vector<float> sspf(4 * XMD_NUM_SS, 0.0f);
    std::vector<double> ssp(4 * XMD_NUM_SS, 0.0);
    for (int k = 0; k < XMD_NUM_SS; ++k) {
        ssp[4 * k] = ss_depth[k];
        ssp[4 * k + 1] = ss_rmin[k];
        ssp[4 * k + 2] = ss_rmax[k];
        for (int c = 0; c < 3; ++c) sspf[4 * k + c] = (float)ssp[4 * k + c];
    }
    std::vector<double> rmd(XMD_QA_NUM_TYPES, 0.0);
    rmd[1] = P.lj_bb_rmin;
    rmd[2] = P.lj_bs_rmin;
    rmd[3] = P.lj_bs_rmin;
    for (int a1 = 0; a1 < XMD_NUM_AA; ++a1)
        for (int a2 = 0; a2 < XMD_NUM_AA; ++a2) rmd[4 + a1 + a2 * XMD_NUM_AA] = ss_rmax[a1 * XMD_NUM_AA + a2];
    for (double r : rmd) E.G.qa.req_max = std::max(E.G.qa.req_max, r);
    std::vector<double> charge(P.aa_charge, P.aa_charge + XMD_NUM_AA);
    std::vector<int8_t> ptype(P.aa_polarization, P.aa_polarization + XMD_NUM_AA);
    Mat J = step_matrix();
    std::vector<uint64_t> jrow(64 * 24), jnib(16 * 16 * 64);
    Mat J256 = mpow(J, 256);
    for (int j = 0; j < 24; ++j) {
        rows_of(J256, jrow.data() + 64 * j);
        J256 = mul(J256, J256);
    }
    // J^(256 w) for every wave offset of the largest replica
    int jwave_n = 1;
    for (int k = 0; k < E.nrep; ++k) jwave_n = std::max(jwave_n, (E.rep_n[k] + WAVE - 1) / WAVE);
    std::vector<uint64_t> jwave(64 * (size_t)jwave_n);
    {
        Mat Jw = identity(), Jstep = mpow(J, 256);
        for (int w = 0; w < jwave_n; ++w) {
            rows_of(Jw, jwave.data() + 64 * (size_t)w);
            Jw = mul(Jstep, Jw);
        }
    }
    std::vector<uint64_t> jadv(64 * (size_t)E.nrep);
    for (int k = 0; k < E.nrep; ++k) rows_of(mpow(J, 4ull * (uint64_t)E.rep_n[k]), jadv.data() + 64 * (size_t)k);
    Mat J4 = mpow(J, 4), Jl = identity();
    for (int l = 0; l < 64; ++l) {
        for (int k = 0; k < 16; ++k)
            for (int v = 0; v < 16; ++v) {
                uint64_t x = 0;
                for (int i = 0; i < 4; ++i)
                    if ((v >> i) & 1) x ^= Jl.col[4 * k + i];
                jnib[64 * (16 * k + v) + l] = x;
            }
        Jl = mul(J4, Jl);
    }

    Dev &D = E.D;
    std::memset(&D, 0, sizeof(D));
    D.nslot = n;
    D.nrep = E.nrep;
    D.rep_slot_off = E.upload(E.rep_slot_off);
    D.rep_n = E.upload(E.rep_n);
    D.box = E.upload(boxes);
    D.noise_sd = E.upload(nsd);
    D.rng[0] = E.upload(rng);
    D.rng[1] = E.upload(rng);
    D.tt[0] = E.upload(t0);
    D.tt[1] = E.upload(t0);
    D.slice_maxd[0] = E.alloc<unsigned long long>(n / WAVE);
    D.slice_maxd[1] = E.alloc<unsigned long long>(n / WAVE);
    D.rb_now = E.alloc<int>(E.nrep);
    D.force_rb = E.alloc<int>(E.nrep);
    D.any_rb = E.alloc<int>(2);
    D.err = E.alloc<int>(E.nrep);
    E.reported.assign(E.nrep, 0);
    D.n_done = E.alloc<long long>(E.nrep);
    D.need_w = E.alloc<int>(E.nrep);
    D.need_fp = E.alloc<int>(E.nrep);
    D.n_rebuild = E.alloc<long long>(E.nrep);
    D.cell_off = E.upload(cell_off);
    D.cell_dim = E.upload(cell_dim);
    D.cell_cnt = E.alloc<int>(ncells);
    D.cell_start = E.alloc<int>(ncells);
    D.cell_end = E.alloc<int>(ncells);
    D.cell_cursor = E.alloc<int>(ncells);
    D.cell_beads = E.alloc<int>(n);
    D.bead_cell = E.alloc<int>(n);
    D.slot_rep = E.upload(E.slot_rep_h);
    D.chain = E.upload(chain_s);
    D.bflags = E.upload(bfl);
    D.atype = E.upload(at);
    D.pos[0] = E.upload(pos);
    D.pos[1] = E.upload(pos);
    D.y = E.alloc<double>(15 * (size_t)n);
    D.orig_r = E.alloc<double>(4 * (size_t)n);
    D.Fbuf = E.alloc<double>(3 * (size_t)n);
    D.pid_side = E.alloc<double>(6 * (size_t)n);
    D.Vbead = E.alloc<double>(n);
    D.Vterm = E.alloc<double>((size_t)NUM_TERMS * n);
    D.tether_nat = E.upload(teth);
    D.theta_nat = E.upload(th);
    D.heur_ang_c = E.upload(std::vector<double>(&P.heur_ang_coeffs[0][0], &P.heur_ang_coeffs[0][0] + 63));
    D.heur_dih_c = E.upload(std::vector<double>(&P.heur_dih_coeffs[0][0], &P.heur_dih_coeffs[0][0] + 54));
    D.phi_nat = E.upload(ph);
    D.chir_nat = E.upload(ch);
    D.chir_fac = E.upload(cf);
    D.nat_start = E.upload(nat_start);
    D.nat_j = E.upload(nat_j);
    D.nat_d = E.upload(nat_d);
    D.nat_act = E.alloc<uint8_t>(nat_j.size());
    D.ss_start = E.upload(ss_start);
    D.ss_j = E.upload(ss_j);
    D.ss_act = E.alloc<uint8_t>(ss_j.size());
    D.solid_w = E.upload(solid_w);
    D.lj_w = E.upload(lj_w);
    D.ljw_status = E.alloc<int8_t>((size_t)n * t->n_lj_walls);
    D.ljw_joint = E.alloc<double>(3 * (size_t)n * t->n_lj_walls);
    D.ljw_ref = E.alloc<double>((size_t)n * t->n_lj_walls);
    D.vafm_start = E.upload(vafm_start);
    D.fafm_start = E.upload(fafm_start);
    D.vafm_orig = E.upload(vafm_orig);
    D.vafm_v = E.upload(vafm_v);
    D.fafm_F = E.upload(fafm_F);
    D.sync = E.upload(sync);
    D.qslot = E.alloc<int>((size_t)MAXQ * n);
    D.qn = E.alloc<int>(n);
    D.ss_depth = E.upload(ss_depth);
    D.ss_rmin = E.upload(ss_rmin);
    D.ss_rmax = E.upload(ss_rmax);
    D.ss_pack_f = E.upload(sspf);
    D.ss_pack = E.upload(ssp);
    D.req_min_dist = E.upload(rmd);
    D.charge = E.upload(charge);
    D.ptype = E.upload(ptype);
    D.jrow = E.upload(jrow);
    D.jwave = E.upload(jwave);
    D.jwave_n = jwave_n;
    D.jnib = E.upload(jnib);
    D.jadv = E.upload(jadv);
    if (P.rng_exact) {
        const FixTable &F = fix_table();
        D.bm_fix = E.upload(F.tab);
        D.bm_thr = E.upload(F.thr);
    }
    int nsl = n / WAVE;
    D.slice_start = E.alloc<int>(nsl);
    D.nl_len = E.alloc<int>(n);
    D.fp_slice_cnt = E.alloc<int>(nsl);
    D.fp_n = E.alloc<int>(n);
    // QA contacts: at most MAXQ per bead -> 8n live; twice that for vacancies
    E.ct_base_h.resize(E.nrep + 1);
    E.ct_cap_h.resize(E.nrep + 1);
    int ct = 0;
    for (int k = 0; k < E.nrep; ++k) {
        E.ct_base_h[k] = ct;
        E.ct_cap_h[k] = !P.qa_enabled       ? BLK
                        : P.ct_capacity > 0 ? round_up(std::min(P.ct_capacity, 1 << 29), BLK)
                                            : round_up(16 * E.rep_n[k] + 1024, BLK);
        ct += E.ct_cap_h[k];
    }
    E.ct_base_h[E.nrep] = ct;
    E.ct_cap_h[E.nrep] = ct;  // total, read by kernels as the SoA stride
    E.ct_total = ct;
    D.ct_base = E.upload(E.ct_base_h);
    D.ct_cap = E.upload(E.ct_cap_h);
    std::vector<int> ctrep(ct / BLK, -1);
    for (int k = 0; k < E.nrep; ++k)
        for (int c = E.ct_base_h[k] / BLK; c < (E.ct_base_h[k] + E.ct_cap_h[k]) / BLK; ++c) ctrep[c] = k;
    D.ct_chunk_rep = E.upload(ctrep);
    E.nct_chunks = ct / BLK;
    D.ct_count = E.alloc<int>(E.nrep);
    D.ct_i1 = E.alloc<int>(ct);
    D.ct_i2 = E.alloc<int>(ct);
    D.ct_type = E.alloc<int>(ct);
    D.ct_status = E.alloc<int>(ct);
    D.ct_ref = E.alloc<double>(ct);
    D.ct_F = E.alloc<double>(3 * (size_t)ct);
    D.ct_V = E.alloc<double>(ct);
    D.ct_d1 = E.alloc<uint32_t>(ct);
    D.ct_d2 = E.alloc<uint32_t>(ct);
    D.ct_vac = E.alloc<uint8_t>(ct);
    D.rm_cnt = E.alloc<int>(ct / BLK);
    D.rm_idx = E.alloc<int>(ct);

    // first rebuild, sized in two passes: count rows, then allocate rows
    E.pending_rb.assign(E.nrep, 1);
    E.put(D.force_rb, std::vector<int>(E.nrep, 1));
    // placeholders so kernels see valid pointers for the count pass
    std::vector<int> zero_base(E.nrep + 1, 0);
    D.fp_base = E.upload(zero_base);
    D.fp_cap = E.upload(zero_base);
    D.fp_count = E.alloc<int>(E.nrep);
    E.cells_seq(D.pos[0], 0, 0);
    k_nl_count<<<E.nl_grid(), BLK, 0, E.stream>>>(D, E.G, D.pos[0], 0);
    HIPCHK(hipGetLastError());
    auto len = E.get(D.nl_len, n);
    std::vector<int> w(nsl, 0);
    int maxlen = 0;
    for (int s = 0; s < n; ++s) {
        w[s / WAVE] = std::max(w[s / WAVE], len[s]);
        maxlen = std::max(maxlen, len[s]);
    }
    // rows are fixed-width: three times the longest first row plus 64
    // (chains compact after the start), at least 256, at most the LDS sort
    // buffer; a longer row later is a capacity error
    if (maxlen > MAXSORT) throw Fail{XMD_ERR_CAPACITY, "a neighbour row is longer than the largest supported row"};
    if (P.row_width > 0) E.nl_w = std::min(MAXSORT, std::max(P.row_width, std::max(maxlen, 1)));
    else E.nl_w = std::min(MAXSORT, std::max(256, round_up(3 * maxlen + 64, WAVE)));
    D.nl_w = E.nl_w;
    {
        std::vector<int> ss(nsl);
        for (int sl = 0; sl < nsl; ++sl) ss[sl] = sl * WAVE * E.nl_w;
        E.put(D.slice_start, ss);
    }
    E.fp_base_h.resize(E.nrep + 1);
    E.fp_cap_h.resize(E.nrep + 1);
    long long nl = 0, fp = 0;
    for (int k = 0; k < E.nrep; ++k) {
        long long need = 0;
        for (int sl = E.rep_slot_off[k] / WAVE; sl < (E.rep_slot_off[k] + E.rep_n[k] + WAVE - 1) / WAVE; ++sl)
            need += 64LL * w[sl];
        long long cap = round_up((int)std::min<long long>(2 * need + 64LL * 64, 1LL << 30), BLK);
        long long fcap = !P.qa_enabled ? BLK
                         : P.fp_capacity > 0 ? round_up(std::min(P.fp_capacity, 1 << 29), BLK)
                                             : round_up((int)(cap / 2 + E.ct_cap_h[k]), BLK);
        E.fp_base_h[k] = (int)fp;
        E.fp_cap_h[k] = (int)fcap;
        fp += fcap;
    }
    nl = (long long)n * E.nl_w;
    if (nl >= (1LL << 31) || fp >= (1LL << 31)) throw Fail{XMD_ERR_CAPACITY, "system too large for int32 list offsets"};
    E.fp_base_h[E.nrep] = (int)fp;
    E.fp_cap_h[E.nrep] = (int)fp;
    E.nl_total = nl;
    E.fp_total = (int)fp;
    E.put(const_cast<int *>(D.fp_base), E.fp_base_h);
    E.put(const_cast<int *>(D.fp_cap), E.fp_cap_h);
    D.nl_e = E.alloc<uint32_t>(nl);
    D.fp_i1 = E.alloc<int>(fp);
    D.fp_i2 = E.alloc<int>(fp);
    D.fp_vac = E.alloc<uint8_t>(fp);
    D.cand_cnt = E.alloc<int>(fp / BLK);
    D.cand_fp = E.alloc<int>(fp);
    D.cand_type = E.alloc<int>(fp);
    D.cand_d1 = E.alloc<uint32_t>(fp);
    D.cand_d2 = E.alloc<uint32_t>(fp);
    std::vector<int> fprep(fp / BLK, -1);
    for (int k = 0; k < E.nrep; ++k)
        for (int c = E.fp_base_h[k] / BLK; c < (E.fp_base_h[k] + E.fp_cap_h[k]) / BLK; ++c) fprep[c] = k;
    D.fp_chunk_rep = E.upload(fprep);
    E.nfp_chunks = (int)(fp / BLK);
    E.qa_grid = E.qa_grid_for(E.nfp_chunks + E.nct_chunks);
    if (P.qa_enabled && E.qa_lds() > 60 * 1024)
        throw Fail{XMD_ERR_CAPACITY, "too many replicas for one engine with QA enabled (at most 7680)"};
    // finish the first rebuild
    E.rows_seq(D.pos[0], 0);
    HIPCHK(hipGetLastError());
    std::fill(E.pending_rb.begin(), E.pending_rb.end(), 0);
    E.settle();  // a first list beyond a given fp_capacity grows it
}

xmd_status run_guarded(std::string &err, const std::function<void()> &fn) {
    try {
        fn();
        return XMD_OK;
    } catch (const Fail &f) {
        err = f.msg;
        return f.st;
    } catch (const std::exception &ex) {
        err = ex.what();
        return XMD_ERR_STATE;
    }
}
}  // namespace

namespace {
// Device buffer for the duration of one API call (after the stream is
// drained, so no queued kernel still uses it).
template <typename T> struct Scratch {
    Part *e;
    T *p = nullptr;
    Scratch(Part *e_, size_t n) : e(e_) {
        HIPCHK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
        HIPCHK(hipMemsetAsync(p, 0, (n ? n : 1) * sizeof(T), e->stream));
    }
    ~Scratch() {
        if (p) {
            (void)hipStreamSynchronize(e->stream);
            (void)hipFree(p);
        }
    }
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
};

// Positions of slots [s0, s0 + nb) from / to host SoA (x | y | z); the
// device keeps them AoS, 4 doubles per slot.
static void put_positions(Part *e, int s0, int nb, const double *r) {
    std::vector<double> a(4 * (size_t)nb, 0.0);
    for (int i = 0; i < nb; ++i)
        for (int d = 0; d < 3; ++d) a[4 * (size_t)i + d] = r[(size_t)d * nb + i];
    HIPCHK(hipMemcpy(e->D.pos[e->cur] + 4 * (size_t)s0, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice));
}
static void get_positions(Part *e, int s0, int nb, double *r) {
    std::vector<double> a(4 * (size_t)nb);
    HIPCHK(hipMemcpy(a.data(), e->D.pos[e->cur] + 4 * (size_t)s0, a.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < nb; ++i)
        for (int d = 0; d < 3; ++d) r[(size_t)d * nb + i] = a[4 * (size_t)i + d];
}

static xmd_status part_set_state(Part *e, int32_t rep, const double *y, uint64_t rng, double t) {
    if (!e) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        if (rep < 0 || rep >= e->nrep || !y) throw Fail{XMD_ERR_ARG, "bad replica or null state"};
        int nb = e->rep_n[rep], s0 = e->rep_slot_off[rep], n = e->nslot;
        HIPCHK(hipStreamSynchronize(e->stream));
        put_positions(e, s0, nb, y);
        for (int k = 1; k < 6; ++k)
            for (int d = 0; d < 3; ++d) {
                const double *src = y + (3 * k + d) * nb;
                double *dst = e->D.y + (size_t)(3 * (k - 1) + d) * n + s0;
                HIPCHK(hipMemcpy(dst, src, nb * sizeof(double), hipMemcpyHostToDevice));
            }
        HIPCHK(hipMemcpy(e->D.rng[e->cur] + rep, &rng, sizeof(uint64_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->D.tt[e->cur] + rep, &t, sizeof(double), hipMemcpyHostToDevice));
        e->reset_replica(rep);
        e->pending_rb[rep] = 1;
    });
}

static xmd_status part_get_state(Part *e, int32_t rep, double *y, double *v, uint64_t *rng, double *t) {
    if (!e) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        if (rep < 0 || rep >= e->nrep) throw Fail{XMD_ERR_ARG, "bad replica"};
        int nb = e->rep_n[rep], s0 = e->rep_slot_off[rep], n = e->nslot;
        HIPCHK(hipStreamSynchronize(e->stream));
        std::vector<double> y1(3 * (size_t)nb);
        if (y) get_positions(e, s0, nb, y);
        for (int k = 1; k < 6; ++k)
            for (int d = 0; d < 3; ++d) {
                const double *src = e->D.y + (size_t)(3 * (k - 1) + d) * n + s0;
                double *dst = y ? y + (3 * k + d) * nb : nullptr;
                if (k == 1) {
                    HIPCHK(hipMemcpy(y1.data() + d * nb, src, nb * sizeof(double), hipMemcpyDeviceToHost));
                    if (dst) std::memcpy(dst, y1.data() + d * nb, nb * sizeof(double));
                } else if (dst) {
                    HIPCHK(hipMemcpy(dst, src, nb * sizeof(double), hipMemcpyDeviceToHost));
                }
            }
        if (v)
            for (size_t i = 0; i < y1.size(); ++i) v[i] = y1[i] * e->G.dt_inv;
        if (rng) HIPCHK(hipMemcpy(rng, e->D.rng[e->cur] + rep, sizeof(uint64_t), hipMemcpyDeviceToHost));
        if (t) HIPCHK(hipMemcpy(t, e->D.tt[e->cur] + rep, sizeof(double), hipMemcpyDeviceToHost));
    });
}

static xmd_status part_set_positions(Part *e, int32_t rep, const double *r) {
    if (!e) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        if (rep < 0 || rep >= e->nrep || !r) throw Fail{XMD_ERR_ARG, "bad replica or null positions"};
        int nb = e->rep_n[rep], s0 = e->rep_slot_off[rep];
        HIPCHK(hipStreamSynchronize(e->stream));
        put_positions(e, s0, nb, r);
        e->reset_replica(rep);
        e->pending_rb[rep] = 1;
    });
}

static xmd_status part_rebuild_lists(Part *e) {
    if (!e) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        std::fill(e->pending_rb.begin(), e->pending_rb.end(), 1);
        e->ensure_lists();
    });
}

static xmd_status part_get_list(Part *e, int32_t rep, int32_t id, int32_t *i1, int32_t *i2, int64_t cap, int64_t *nout) {
    if (!e || !nout) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        if (rep < 0 || rep >= e->nrep) throw Fail{XMD_ERR_ARG, "bad replica"};
        e->ensure_lists();
        int s0 = e->rep_slot_off[rep], nb = e->rep_n[rep], n = e->nslot;
        std::vector<std::pair<int, int>> out;
        auto local = [&](int s) { return s - s0; };
        if (id == XMD_LIST_NL_ALL || id == XMD_LIST_NL_NATIVE || id == XMD_LIST_NL_NON_NATIVE ||
            id == XMD_LIST_NL_NAT_SS) {
            // scratch of this call only, freed on return
            Scratch<int> cnt(e, n);
            k_nl_export<<<e->nblk(), BLK, 0, e->stream>>>(e->D, e->G, cnt.p, nullptr, nullptr, nullptr, nullptr, 0);
            auto c = e->get(cnt.p, n);
            std::vector<int> oall(n + 1, 0);
            for (int s = 0; s < n; ++s) oall[s + 1] = oall[s] + c[s];
            Scratch<int> off(e, n + 1), oi1_s(e, oall[n] + 1), oi2_s(e, oall[n] + 1);
            Scratch<uint8_t> onat_s(e, oall[n] + 1);
            e->put(off.p, oall);
            int *oi1 = oi1_s.p, *oi2 = oi2_s.p;
            uint8_t *onat = onat_s.p;
            k_nl_export<<<e->nblk(), BLK, 0, e->stream>>>(e->D, e->G, cnt.p, off.p, oi1, oi2, onat, 1);
            HIPCHK(hipGetLastError());
            auto a = e->get(oi1, oall[n]), b = e->get(oi2, oall[n]);
            auto nat = e->get(onat, oall[n]);
            for (int s = s0; s < s0 + nb; ++s)
                for (int k = oall[s]; k < oall[s + 1]; ++k) {
                    bool keep = id == XMD_LIST_NL_ALL || (id == XMD_LIST_NL_NATIVE && nat[k] == 1) ||
                                (id == XMD_LIST_NL_NON_NATIVE && nat[k] == 0) || (id == XMD_LIST_NL_NAT_SS && nat[k] == 2);
                    if (keep) out.push_back({local(a[k]), local(b[k])});
                }
        } else if (id == XMD_LIST_GO) {
            auto st = e->get(e->D.nat_start, n + 1);
            auto j = e->get(e->D.nat_j, st[n]);
            auto act = e->get(e->D.nat_act, st[n]);
            for (int s = s0; s < s0 + nb; ++s)
                for (int k = st[s]; k < st[s + 1]; ++k)
                    if (j[k] > s && act[k]) out.push_back({local(s), local(j[k])});
        } else if (id == XMD_LIST_NAT_SS) {
            auto st = e->get(e->D.ss_start, n + 1);
            auto j = e->get(e->D.ss_j, st[n]);
            auto act = e->get(e->D.ss_act, st[n]);
            for (int s = s0; s < s0 + nb; ++s)
                for (int k = st[s]; k < st[s + 1]; ++k)
                    if (j[k] > s && act[k]) out.push_back({local(s), local(j[k])});
        } else if (id == XMD_LIST_PAULI || id == XMD_LIST_ES || id == XMD_LIST_PID) {
            uint8_t want = id == XMD_LIST_PAULI ? F_PAULI : (id == XMD_LIST_ES ? F_ES : F_PID);
            auto len = e->get(e->D.nl_len, n);
            // the replica's rows: slots [s0, s0 + nb) -> slices of width 64 nl_w
            const size_t r0 = (size_t)(s0 / WAVE) * WAVE * e->nl_w;
            const size_t rn = (size_t)((nb + WAVE - 1) / WAVE) * WAVE * e->nl_w;
            auto ent = e->get(e->D.nl_e + r0, rn);
            for (int s = s0; s < s0 + nb; ++s)
                for (int k = 0; k < len[s]; ++k) {
                    size_t idx = (size_t)(s / WAVE) * WAVE * e->nl_w - r0 + 64 * (size_t)k + s % WAVE;
                    const int j = nle_j(ent[idx]);
                    if (j > s && (nle_flag(ent[idx]) & want)) out.push_back({local(s), local(j)});
                }
        } else if (id == XMD_LIST_QA_FREE) {
            auto cntv = e->get(e->D.fp_count, e->nrep);
            int b = e->fp_base_h[rep];
            auto a = e->get(e->D.fp_i1 + b, cntv[rep]);
            auto c = e->get(e->D.fp_i2 + b, cntv[rep]);
            auto v = e->get(e->D.fp_vac + b, cntv[rep]);
            for (int k = 0; k < cntv[rep]; ++k)
                if (!v[k]) out.push_back({local(a[k]), local(c[k])});
        } else {
            throw Fail{XMD_ERR_ARG, "bad list id"};
        }
        *nout = (int64_t)out.size();
        if (i1 && i2)
            for (int64_t k = 0; k < std::min<int64_t>(cap, *nout); ++k) {
                i1[k] = out[k].first;
                i2[k] = out[k].second;
            }
    });
}

static xmd_status part_eval_forces(Part *e, uint32_t term_mask, double *F, double *V_terms) {
    if (!e || !F) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        e->ensure_lists();
        e->require_running();
        uint32_t m = term_mask & e->mask;
        int p = e->cur;
        if (e->G.qa_on && (m & (1u << T_QA)))
            k_qa_pairs<<<e->qa_grid, BLK, e->qa_lds(), e->stream>>>(e->D, e->G, p, 0);
        int nb = e->nblk();
        TermTab<float> Tf = e->TTf;
        TermTab<double> Td = e->TTd;
        Tf.mask = m;
        Td.mask = m;
        launch_force(e->P.precision == XMD_PREC_MIXED, true, e->lpb, e->nslot, e->stream, e->D, e->G, Tf, Td, p, 0);
        k_gather_force<<<nb, BLK, 0, e->stream>>>(e->D);
        HIPCHK(hipGetLastError());
        int n = e->nslot;
        auto Fh = e->get(e->D.Fbuf, 3 * (size_t)n);
        for (int g = 0; g < e->nbeads; ++g) {
            int s = e->slot_of_bead[g];
            for (int d = 0; d < 3; ++d) F[d * e->nbeads + g] = Fh[d * n + s];
        }
        if (V_terms) {
            auto Vt = e->get(e->D.Vterm, (size_t)NUM_TERMS * n);
            for (int k = 0; k < e->nrep; ++k)
                for (int t = 0; t < NUM_TERMS; ++t) {
                    double acc = 0.0;
                    if (m & (1u << t))
                        for (int s = e->rep_slot_off[k]; s < e->rep_slot_off[k] + e->rep_n[k]; ++s)
                            acc += Vt[(size_t)t * n + s];
                    V_terms[k * XMD_NUM_TERMS + t] = acc;
                }
        }
    });
}

static xmd_status part_qa_update(Part *e) {
    if (!e) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        e->ensure_lists();
        e->require_running();
        if (!e->G.qa_on) return;
        k_qa_pairs<<<e->qa_grid, BLK, e->qa_lds(), e->stream>>>(e->D, e->G, e->cur, 1);
        k_qa_serial<<<e->nrep, BLK, 0, e->stream>>>(e->D, e->cur);
        HIPCHK(hipGetLastError());
        e->check_err();
    });
}

static xmd_status part_get_qa_state(Part *e, int32_t rep, xmd_qa_contact *c, int64_t cap, int64_t *nout, int8_t *sync) {
    if (!e || !nout) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        if (rep < 0 || rep >= e->nrep) throw Fail{XMD_ERR_ARG, "bad replica"};
        int s0 = e->rep_slot_off[rep], nb = e->rep_n[rep];
        auto cnt = e->get(e->D.ct_count, e->nrep)[rep];
        int b = e->ct_base_h[rep];
        auto i1 = e->get(e->D.ct_i1 + b, cnt), i2 = e->get(e->D.ct_i2 + b, cnt), ty = e->get(e->D.ct_type + b, cnt),
             st = e->get(e->D.ct_status + b, cnt);
        auto rf = e->get(e->D.ct_ref + b, cnt);
        auto d1 = e->get(e->D.ct_d1 + b, cnt), d2 = e->get(e->D.ct_d2 + b, cnt);
        auto vac = e->get(e->D.ct_vac + b, cnt);
        int64_t k = 0;
        for (int q = 0; q < cnt; ++q) {
            if (vac[q]) continue;
            if (c && k < cap) {
                xmd_qa_contact &o = c[k];
                o.i1 = i1[q] - s0;
                o.i2 = i2[q] - s0;
                o.type = (int16_t)ty[q];
                o.status = (int8_t)st[q];
                o._pad = 0;
                std::memcpy(o.diff1, &d1[q], 4);
                std::memcpy(o.diff2, &d2[q], 4);
                o.ref_time = rf[q];
            }
            ++k;
        }
        *nout = k;
        if (sync) {
            auto sy = e->get(e->D.sync + s0, nb);
            std::memcpy(sync, sy.data(), 4 * (size_t)nb);
        }
    });
}

static xmd_status part_get_wall_state(Part *e, int32_t rep, int8_t *status, double *joint, double *ref_t) {
    return run_guarded(e->err, [&]() {
        if (rep < 0 || rep >= e->nrep) throw Fail{XMD_ERR_ARG, "bad replica"};
        const size_t w = (size_t)e->G.n_ljw, s0 = (size_t)e->rep_slot_off[rep], k = (size_t)e->rep_n[rep] * w;
        if (!k) return;
        auto st = e->get(e->D.ljw_status + s0 * w, k);
        auto jo = e->get(e->D.ljw_joint + 3 * s0 * w, 3 * k);
        auto rt = e->get(e->D.ljw_ref + s0 * w, k);
        if (status) std::memcpy(status, st.data(), k);
        if (joint) std::memcpy(joint, jo.data(), 3 * k * sizeof(double));
        if (ref_t) std::memcpy(ref_t, rt.data(), k * sizeof(double));
    });
}

static xmd_status part_set_wall_state(Part *e, int32_t rep, const int8_t *status, const double *joint,
                                      const double *ref_t) {
    return run_guarded(e->err, [&]() {
        if (rep < 0 || rep >= e->nrep) throw Fail{XMD_ERR_ARG, "bad replica"};
        const size_t w = (size_t)e->G.n_ljw, s0 = (size_t)e->rep_slot_off[rep], k = (size_t)e->rep_n[rep] * w;
        if (!k) return;
        if (!status || !joint || !ref_t) throw Fail{XMD_ERR_ARG, "null wall state"};
        for (size_t q = 0; q < k; ++q)
            if (status[q] < 0 || status[q] > 2) throw Fail{XMD_ERR_ARG, "bad wall status"};
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipMemcpy(e->D.ljw_status + s0 * w, status, k, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->D.ljw_joint + 3 * s0 * w, joint, 3 * k * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->D.ljw_ref + s0 * w, ref_t, k * sizeof(double), hipMemcpyHostToDevice));
    });
}

static xmd_status part_set_qa_state(Part *e, int32_t rep, const xmd_qa_contact *c, int64_t nc, const int8_t *sync) {
    if (!e) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        if (rep < 0 || rep >= e->nrep) throw Fail{XMD_ERR_ARG, "bad replica"};
        int s0 = e->rep_slot_off[rep], nb = e->rep_n[rep], n = e->nslot;
        if (nc > e->ct_cap_h[rep] / 2) throw Fail{XMD_ERR_CAPACITY, "too many contacts"};
        std::vector<int> i1(nc), i2(nc), ty(nc), st(nc);
        std::vector<double> rf(nc);
        std::vector<uint32_t> d1(nc), d2(nc);
        std::vector<uint8_t> vac(nc, 0);
        std::vector<int> qn(nb, 0), qslot((size_t)MAXQ * nb, 0);
        int b = e->ct_base_h[rep];
        for (int64_t k = 0; k < nc; ++k) {
            if (c[k].i1 < 0 || c[k].i2 >= nb || !(c[k].i1 < c[k].i2)) throw Fail{XMD_ERR_ARG, "bad contact"};
            i1[k] = c[k].i1 + s0;
            i2[k] = c[k].i2 + s0;
            ty[k] = c[k].type;
            st[k] = c[k].status;
            rf[k] = c[k].ref_time;
            std::memcpy(&d1[k], c[k].diff1, 4);
            std::memcpy(&d2[k], c[k].diff2, 4);
            for (int a : {c[k].i1, c[k].i2}) {
                if (qn[a] >= MAXQ) throw Fail{XMD_ERR_CAPACITY, "more QA contacts on a bead than slots"};
                qslot[(size_t)qn[a] * nb + a] = b + (int)k;
                qn[a]++;
            }
        }
        HIPCHK(hipStreamSynchronize(e->stream));
        auto up = [&](auto *dst, const auto &h) {
            if (!h.empty()) HIPCHK(hipMemcpy(dst, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice));
        };
        up(e->D.ct_i1 + b, i1);
        up(e->D.ct_i2 + b, i2);
        up(e->D.ct_type + b, ty);
        up(e->D.ct_status + b, st);
        up(e->D.ct_ref + b, rf);
        up(e->D.ct_d1 + b, d1);
        up(e->D.ct_d2 + b, d2);
        up(e->D.ct_vac + b, vac);
        int cnt = (int)nc;
        HIPCHK(hipMemcpy(e->D.ct_count + rep, &cnt, sizeof(int), hipMemcpyHostToDevice));
        up(e->D.qn + s0, qn);
        for (int q = 0; q < MAXQ; ++q)
            HIPCHK(hipMemcpy(e->D.qslot + (size_t)q * n + s0, qslot.data() + (size_t)q * nb, nb * sizeof(int),
                             hipMemcpyHostToDevice));
        if (sync) HIPCHK(hipMemcpy(e->D.sync + s0, sync, 4 * (size_t)nb, hipMemcpyHostToDevice));
        e->reset_replica(rep);
        e->pending_rb[rep] = 1;
    });
}


static xmd_status part_get_energy(Part *e, double *V) {
    if (!e || !V) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        auto Vb = e->get(e->D.Vbead, e->nslot);
        for (int k = 0; k < e->nrep; ++k) {
            double acc = 0.0;
            for (int s = e->rep_slot_off[k]; s < e->rep_slot_off[k] + e->rep_n[k]; ++s) acc += Vb[s];
            V[k] = acc;
        }
    });
}

static xmd_status part_get_counters(Part *e, int64_t *out) {
    if (!e || !out) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        auto rb = e->get(e->D.n_rebuild, e->nrep);
        auto len = e->get(e->D.nl_len, e->nslot);
        auto cnt = e->get(e->D.ct_count, e->nrep);
        int64_t r = 0, nl = 0, qa = 0;
        for (auto x : rb) r += x;
        for (int s = 0; s < e->nslot; ++s)
            if (e->bead_of_slot[s] >= 0) nl += len[s];
        for (int k = 0; k < e->nrep; ++k) {
            auto v = e->get(e->D.ct_vac + e->ct_base_h[k], cnt[k]);
            for (auto x : v) qa += !x;
        }
        out[0] = e->steps - e->catchup_steps;
        out[1] = r;
        out[2] = nl;
        out[3] = e->nl_total;
        out[4] = qa;
    });
}

static xmd_status part_profile(Part *e, int64_t nsteps, char *names, int64_t names_cap, int64_t *counts,
                       double *total_ms, int32_t cap, int32_t *n_kernels) {
    if (!e || nsteps < 0 || !n_kernels) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        e->ensure_lists();
        e->prof_names.clear();
        e->prof_ms.clear();
        e->prof_n.clear();
        e->prof_on = true;
        try {
            for (int64_t k = 0; k < nsteps; ++k) {
                e->step_seq(e->cur);
                e->cur ^= 1;
                e->steps += 1;
                if ((k & 63) == 63) e->prof_flush();
            }
            e->prof_flush();
        } catch (...) {
            e->prof_on = false;
            throw;
        }
        e->prof_on = false;
        HIPCHK(hipGetLastError());
        e->check_err();
        int nk = (int)e->prof_names.size();
        *n_kernels = nk;
        std::string all;
        for (int k = 0; k < nk; ++k) {
            if (k < cap) {
                if (counts) counts[k] = e->prof_n[k];
                if (total_ms) total_ms[k] = e->prof_ms[k];
            }
            all += e->prof_names[k] + "\n";
        }
        if (names && names_cap > 0) {
            size_t m = std::min<size_t>(all.size(), (size_t)names_cap - 1);
            std::memcpy(names, all.data(), m);
            names[m] = 0;
        }
    });
}

static xmd_status part_rand_normals(Part *e, uint64_t state, int64_t nb, float *out, uint64_t *state_out) {
    if (!e || nb < 0 || (nb > 0 && !out)) return XMD_ERR_ARG;
    return run_guarded(e->err, [&]() {
        if (nb > (1LL << 28)) throw Fail{XMD_ERR_ARG, "too many beads"};
        long long padded = (nb + BLK - 1) / BLK * BLK;
        float *d = nullptr;
        if (nb > 0) {
            HIPCHK(hipMalloc(&d, 3 * sizeof(float) * (size_t)padded));
            k_rand_normals<<<(int)(padded / BLK), BLK, 0, e->stream>>>(e->D, state, nb, d);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(out, d, 3 * sizeof(float) * (size_t)nb, hipMemcpyDeviceToHost, e->stream));
            HIPCHK(hipStreamSynchronize(e->stream));
            HIPCHK(hipFree(d));
        }
        if (state_out) {
            Mat J = step_matrix();
            *state_out = apply(mpow(J, 4ull * (uint64_t)nb), state);
        }
    });
}

}  // namespace

// ------------------------------------------------- no-alloc staging ---
// Staging of xmd_observables_noalloc: a file-scope device array, so that
// call allocates nothing.  Every GPU the library runs on holds its own copy
// until the process ends.  Host code never passes the array's name to a
// kernel or copy (on the host it is only a stand-in): the device address
// comes from hipGetSymbolAddress, once per device, with that device current
// (at engine creation, the library's set-up).
constexpr int OBS_NOALLOC_MAX_REPLICAS = XMD_OBS_NOALLOC_MAX_REPLICAS;
static __device__ __align__(256) double g_obs_scratch[8 * OBS_NOALLOC_MAX_REPLICAS];
namespace {
std::mutex g_obs_mu;
std::vector<double *> g_obs_ptr;  // device address of g_obs_scratch, by device ordinal
double *obs_scratch_of_current_device() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_obs_mu);
    if ((int)g_obs_ptr.size() <= dev) g_obs_ptr.resize(dev + 1, nullptr);
    if (!g_obs_ptr[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(g_obs_scratch)));
        g_obs_ptr[dev] = (double *)ptr;
    }
    return g_obs_ptr[dev];
}
}  // namespace

// ================================================================ parts ===
// An engine advances its replicas as one or more parts (contiguous replica
// ranges of about equal bead count), each with its own device state, HIP
// stream and captured step graphs.  With two parts, one part's
// latency-bound kernels (QA sift, integrator, list rebuild) run while the
// other's force kernel holds the chip: xmd_step launches their graphs
// interleaved and waits once.  Results per replica do not depend on the
// split (kernels never mix replicas).
struct xmd_engine {
    std::vector<Part *> parts;
    std::vector<int> rep_part, rep_local, part_rep0, part_bead0;
    int nrep = 0, nbeads = 0;
    std::string err;
    double *obs_scratch = nullptr;  // this GPU's g_obs_scratch (device address)
    ~xmd_engine() {
        for (Part *q : parts) delete q;
    }
};

namespace {

// Runs a part-level call and copies the part's error message.
xmd_status on_part(xmd_engine *e, Part *q, xmd_status st) {
    if (st != XMD_OK) e->err = q->err;
    return st;
}

// The sub-topology of replicas [r0, r1) with indices rebased to its first
// bead; `store` keeps the arrays it points into alive.
struct SubTopo {
    xmd_topology t{};
    std::vector<int32_t> rep_off, chain_off, i1, i2, ss1, ss2, vidx, fidx;
    std::vector<double> vorig, vvel, fF;
    std::vector<double> r0, dist;
};
void sub_topology(const xmd_topology *t, int r0, int r1, SubTopo &o) {
    const int b0 = t->replica_offset[r0], b1 = t->replica_offset[r1], nb = b1 - b0;
    for (int r = r0; r <= r1; ++r) o.rep_off.push_back(t->replica_offset[r] - b0);
    for (int c = 0; c < t->n_chains; ++c)
        if (t->chain_offset[c] >= b0 && t->chain_offset[c] < b1) o.chain_off.push_back(t->chain_offset[c] - b0);
    o.chain_off.push_back(nb);
    o.r0.resize(3 * (size_t)nb);
    for (int d = 0; d < 3; ++d)
        std::memcpy(o.r0.data() + (size_t)d * nb, t->r0 + (size_t)d * t->n_beads + b0, nb * sizeof(double));
    for (int k = 0; k < t->n_native; ++k)
        if (t->nat_i1[k] >= b0 && t->nat_i1[k] < b1) {
            if (t->nat_i2[k] < b0 || t->nat_i2[k] >= b1) throw Fail{XMD_ERR_ARG, "native contact spans replicas"};
            o.i1.push_back(t->nat_i1[k] - b0);
            o.i2.push_back(t->nat_i2[k] - b0);
            o.dist.push_back(t->nat_dist[k]);
        }
    for (int k = 0; k < t->n_ssbonds; ++k)
        if (t->ss_i1[k] >= b0 && t->ss_i1[k] < b1) {
            if (t->ss_i2[k] < b0 || t->ss_i2[k] >= b1) throw Fail{XMD_ERR_ARG, "ssbond spans replicas"};
            o.ss1.push_back(t->ss_i1[k] - b0);
            o.ss2.push_back(t->ss_i2[k] - b0);
        }
    auto at = [&](const double *a) { return a ? a + b0 : nullptr; };
    o.t.n_replicas = r1 - r0;
    o.t.n_beads = nb;
    o.t.n_chains = (int32_t)o.chain_off.size() - 1;
    o.t.n_native = (int32_t)o.i1.size();
    o.t.replica_offset = o.rep_off.data();
    o.t.chain_offset = o.chain_off.data();
    o.t.atype = t->atype + b0;
    o.t.r0 = o.r0.data();
    o.t.box = t->box + 3 * (size_t)r0;
    o.t.temperature = t->temperature + r0;
    o.t.rng_state = t->rng_state + r0;
    o.t.nat_i1 = o.i1.data();
    o.t.nat_i2 = o.i2.data();
    o.t.nat_dist = o.dist.data();
    o.t.tether_len = at(t->tether_len);
    o.t.nat_theta = at(t->nat_theta);
    o.t.nat_phi = at(t->nat_phi);
    o.t.nat_chir = at(t->nat_chir);
    o.t.nat_chir_factor = at(t->nat_chir_factor);
    o.t.n_ssbonds = (int32_t)o.ss1.size();
    o.t.ss_i1 = o.ss1.data();
    o.t.ss_i2 = o.ss2.data();
    o.t.n_solid_walls = t->n_solid_walls;
    o.t.solid_walls = t->solid_walls;
    o.t.n_lj_walls = t->n_lj_walls;
    o.t.lj_walls = t->lj_walls;
    for (int k = 0; k < t->n_vel_afm; ++k)
        if (t->vel_afm_idx[k] >= b0 && t->vel_afm_idx[k] < b1) {
            o.vidx.push_back(t->vel_afm_idx[k] - b0);
            o.vorig.insert(o.vorig.end(), t->vel_afm_orig + 3 * k, t->vel_afm_orig + 3 * k + 3);
            o.vvel.insert(o.vvel.end(), t->vel_afm_v + 3 * k, t->vel_afm_v + 3 * k + 3);
        }
    for (int k = 0; k < t->n_force_afm; ++k)
        if (t->force_afm_idx[k] >= b0 && t->force_afm_idx[k] < b1) {
            o.fidx.push_back(t->force_afm_idx[k] - b0);
            o.fF.insert(o.fF.end(), t->force_afm_F + 3 * k, t->force_afm_F + 3 * k + 3);
        }
    o.t.n_vel_afm = (int32_t)o.vidx.size();
    o.t.vel_afm_idx = o.vidx.data();
    o.t.vel_afm_orig = o.vorig.data();
    o.t.vel_afm_v = o.vvel.data();
    o.t.n_force_afm = (int32_t)o.fidx.size();
    o.t.force_afm_idx = o.fidx.data();
    o.t.force_afm_F = o.fF.data();
    o.t.heur_bonded = t->heur_bonded ? t->heur_bonded + b0 : nullptr;
}

// Every precondition of xmd_params / xmd_topology that the host can check,
// on the whole topology before it is split into stream groups and before
// any device call (the reference throws std::runtime_error from its
// loaders for the same kinds of input).
void validate(const xmd_params *p, const xmd_topology *t) {
    auto bad = [](const char *m) { throw Fail{XMD_ERR_ARG, m}; };
    if (!p || !t) bad("null params or topology");
    if (!(p->dt > 0)) bad("dt must be positive");
    if (p->precision != XMD_PREC_FP64 && p->precision != XMD_PREC_MIXED) bad("bad precision");
    if (p->force_lanes != 0 && p->force_lanes != 4 && p->force_lanes != 8 && p->force_lanes != 16)
        bad("force_lanes must be 0, 4, 8 or 16");
    if (p->stream_groups < 0) bad("stream_groups must be >= 0");
    if (p->row_width < 0) bad("row_width must be >= 0");
    if (p->fp_capacity < 0) bad("fp_capacity must be >= 0");
    if (p->ct_capacity < 0) bad("ct_capacity must be >= 0");
    if (t->n_replicas <= 0 || t->n_beads <= 0 || !t->replica_offset || !t->chain_offset || !t->r0)
        bad("empty topology");
    if (!t->atype || !t->box || !t->temperature || !t->rng_state) bad("null topology array");
    const int n = t->n_beads, R = t->n_replicas;
    const int32_t *ro = t->replica_offset, *co = t->chain_offset;
    if (ro[0] != 0 || ro[R] != n) bad("replica_offset must span [0, n_beads]");
    for (int r = 0; r < R; ++r)
        if (ro[r + 1] <= ro[r]) bad("empty replica");
    if (t->n_chains <= 0 || co[0] != 0 || co[t->n_chains] != n) bad("chain_offset must span [0, n_beads]");
    for (int c = 0, r = 0; c < t->n_chains; ++c) {
        if (co[c + 1] <= co[c]) bad("empty chain");
        while (ro[r + 1] <= co[c]) ++r;
        if (co[c + 1] > ro[r + 1]) bad("chain spans replicas");
    }
    for (int g = 0; g < n; ++g) {
        if (t->atype[g] < 0 || t->atype[g] >= XMD_NUM_AA) bad("atype out of range");
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(t->r0[(size_t)d * n + g])) bad("non-finite initial position");
    }
    for (int k = 0; k < 3 * R; ++k)
        if (!(t->box[k] >= 0) || !std::isfinite(t->box[k])) bad("box lengths must be finite and >= 0");
    auto rep_of = [&](int g) { return (int)(std::upper_bound(ro, ro + R + 1, g) - ro) - 1; };
    auto chain_of = [&](int g) { return (int)(std::upper_bound(co, co + t->n_chains + 1, g) - co) - 1; };
    if (t->n_native < 0 || (t->n_native > 0 && (!t->nat_i1 || !t->nat_i2 || !t->nat_dist)))
        bad("bad native contact arrays");
    for (int k = 0; k < t->n_native; ++k) {
        const int a = t->nat_i1[k], b = t->nat_i2[k];
        if (a < 0 || b >= n || !(a < b)) bad("native contact needs 0 <= i1 < i2 < n");
        if (rep_of(a) != rep_of(b)) bad("native contact crosses replicas");
        if (!(t->nat_dist[k] > 0) || !std::isfinite(t->nat_dist[k])) bad("native contact distance must be positive");
    }
    if (t->n_ssbonds < 0 || (t->n_ssbonds > 0 && (!t->ss_i1 || !t->ss_i2))) bad("bad ssbond arrays");
    for (int k = 0; k < t->n_ssbonds; ++k) {
        const int a = t->ss_i1[k], b = t->ss_i2[k];
        if (a < 0 || b >= n || !(a < b)) bad("ssbond needs 0 <= i1 < i2 < n");
        if (rep_of(a) != rep_of(b)) bad("ssbond crosses replicas");
    }
    if (t->n_solid_walls < 0 || t->n_lj_walls < 0 || (t->n_solid_walls > 0 && !t->solid_walls) ||
        (t->n_lj_walls > 0 && !t->lj_walls))
        bad("bad wall arrays");
    if (t->n_vel_afm < 0 || t->n_force_afm < 0 ||
        (t->n_vel_afm > 0 && (!t->vel_afm_idx || !t->vel_afm_orig || !t->vel_afm_v)) ||
        (t->n_force_afm > 0 && (!t->force_afm_idx || !t->force_afm_F)))
        bad("bad AFM arrays");
    for (int k = 0; k < t->n_vel_afm; ++k)
        if (t->vel_afm_idx[k] < 0 || t->vel_afm_idx[k] >= n) bad("velocity AFM tip outside the beads");
    for (int k = 0; k < t->n_force_afm; ++k)
        if (t->force_afm_idx[k] < 0 || t->force_afm_idx[k] >= n) bad("force AFM tip outside the beads");
    if (t->heur_bonded)
        for (int g = 0; g < n; ++g) {
            const uint8_t hb = t->heur_bonded[g];
            if ((hb & 1) && !(g + 2 < n && chain_of(g + 2) == chain_of(g)))
                bad("heuristic angle does not join consecutive residues");
            if ((hb & 2) && !(g + 3 < n && chain_of(g + 3) == chain_of(g)))
                bad("heuristic dihedral does not join consecutive residues");
        }
}

void create(xmd_engine &E, const xmd_params *p, const xmd_topology *t) {
    validate(p, t);
    const int R = t->n_replicas;
    const int G = std::min(R, p->stream_groups > 0 ? p->stream_groups : (R >= 2 ? 2 : 1));
    // go.cpp cutoff: lj::cutoff(nat_dist) = 2 nat_dist, max over all contacts
    double go_cut = 0.0;
    for (int k = 0; k < t->n_native; ++k) go_cut = std::max(go_cut, 2.0 * t->nat_dist[k]);
    // cut points: the first replica starting at or after g/G of the beads
    std::vector<int> cut{0};
    for (int g = 1; g < G; ++g) {
        long long target = (long long)t->n_beads * g / G;
        int r = cut.back() + 1;
        while (r < R - (G - g) && t->replica_offset[r] < target) ++r;
        cut.push_back(r);
    }
    cut.push_back(R);
    E.nrep = R;
    E.nbeads = t->n_beads;
    E.rep_part.resize(R);
    E.rep_local.resize(R);
    for (int g = 0; g < G; ++g) {
        SubTopo st;
        sub_topology(t, cut[g], cut[g + 1], st);
        Part *q = new Part();
        E.parts.push_back(q);
        q->rep0 = cut[g];
        build(*q, p, &st.t, go_cut);
        E.part_rep0.push_back(cut[g]);
        E.part_bead0.push_back(t->replica_offset[cut[g]]);
        for (int r = cut[g]; r < cut[g + 1]; ++r) {
            E.rep_part[r] = g;
            E.rep_local[r] = r - cut[g];
        }
    }
}

}  // namespace

extern "C" {

const char *xmd_version(void) { return "xmd-hip 0.1.0 gfx950"; }

int64_t xmd_rng_first_zero_draw(uint64_t state, int64_t n_draws, uint32_t slot_mask) {
    for (int64_t k = 0; k < n_draws; ++k) {
        if (((slot_mask >> (k & 3)) & 1u) && (xmd_dev::rng_out(state) >> 32) == 0) return k;
        state = xmd_dev::rng_step(state);
    }
    return -1;
}

const char *xmd_last_error(const xmd_engine *e) { return e ? e->err.c_str() : g_create_err.c_str(); }

xmd_status xmd_engine_create(const xmd_params *p, const xmd_topology *t, xmd_engine **out) {
    if (!out) {
        g_create_err = "null output pointer";
        return XMD_ERR_ARG;
    }
    *out = nullptr;
    xmd_engine *e = new xmd_engine();
    xmd_status st = run_guarded(g_create_err, [&]() {
        create(*e, p, t);
        e->obs_scratch = obs_scratch_of_current_device();
    });
    if (st != XMD_OK) {
        delete e;
        return st;
    }
    *out = e;
    return XMD_OK;
}

void xmd_engine_destroy(xmd_engine *e) { delete e; }

void *xmd_stream(xmd_engine *e) { return e ? (void *)e->parts[0]->stream : nullptr; }

#define XMD_REPLICA_CALL(call)                                                                                      \
    if (!e) return XMD_ERR_ARG;                                                                                     \
    if (rep < 0 || rep >= e->nrep) {                                                                                \
        e->err = "bad replica";                                                                                     \
        return XMD_ERR_ARG;                                                                                         \
    }                                                                                                               \
    Part *q = e->parts[e->rep_part[rep]];                                                                           \
    const int lr = e->rep_local[rep];                                                                               \
    (void)lr;                                                                                                       \
    return on_part(e, q, call);

xmd_status xmd_set_state(xmd_engine *e, int32_t rep, const double *y, uint64_t rng, double t) {
    XMD_REPLICA_CALL(part_set_state(q, lr, y, rng, t))
}
xmd_status xmd_get_state(xmd_engine *e, int32_t rep, double *y, double *v, uint64_t *rng, double *t) {
    XMD_REPLICA_CALL(part_get_state(q, lr, y, v, rng, t))
}
xmd_status xmd_set_positions(xmd_engine *e, int32_t rep, const double *r) {
    XMD_REPLICA_CALL(part_set_positions(q, lr, r))
}
xmd_status xmd_get_list(xmd_engine *e, int32_t rep, int32_t id, int32_t *i1, int32_t *i2, int64_t cap, int64_t *nout) {
    XMD_REPLICA_CALL(part_get_list(q, lr, id, i1, i2, cap, nout))
}
xmd_status xmd_get_qa_state(xmd_engine *e, int32_t rep, xmd_qa_contact *c, int64_t cap, int64_t *nout, int8_t *sync) {
    XMD_REPLICA_CALL(part_get_qa_state(q, lr, c, cap, nout, sync))
}
xmd_status xmd_set_qa_state(xmd_engine *e, int32_t rep, const xmd_qa_contact *c, int64_t nc, const int8_t *sync) {
    XMD_REPLICA_CALL(part_set_qa_state(q, lr, c, nc, sync))
}
xmd_status xmd_get_wall_state(xmd_engine *e, int32_t rep, int8_t *status, double *joint, double *ref_t) {
    XMD_REPLICA_CALL(part_get_wall_state(q, lr, status, joint, ref_t))
}
xmd_status xmd_set_wall_state(xmd_engine *e, int32_t rep, const int8_t *status, const double *joint,
                              const double *ref_t) {
    XMD_REPLICA_CALL(part_set_wall_state(q, lr, status, joint, ref_t))
}

xmd_status xmd_rebuild_lists(xmd_engine *e) {
    if (!e) return XMD_ERR_ARG;
    for (Part *q : e->parts) {
        xmd_status st = on_part(e, q, part_rebuild_lists(q));
        if (st != XMD_OK) return st;
    }
    return XMD_OK;
}

xmd_status xmd_qa_update(xmd_engine *e) {
    if (!e) return XMD_ERR_ARG;
    for (Part *q : e->parts) {
        xmd_status st = on_part(e, q, part_qa_update(q));
        if (st != XMD_OK) return st;
    }
    return XMD_OK;
}

xmd_status xmd_eval_forces(xmd_engine *e, uint32_t term_mask, double *F, double *V_terms) {
    if (!e || !F) return XMD_ERR_ARG;
    for (size_t g = 0; g < e->parts.size(); ++g) {
        Part *q = e->parts[g];
        std::vector<double> Fp(3 * (size_t)q->nbeads), Vp((size_t)q->nrep * XMD_NUM_TERMS);
        xmd_status st = on_part(e, q, part_eval_forces(q, term_mask, Fp.data(), V_terms ? Vp.data() : nullptr));
        if (st != XMD_OK) return st;
        for (int d = 0; d < 3; ++d)
            std::memcpy(F + (size_t)d * e->nbeads + e->part_bead0[g], Fp.data() + (size_t)d * q->nbeads,
                        q->nbeads * sizeof(double));
        if (V_terms) std::memcpy(V_terms + (size_t)e->part_rep0[g] * XMD_NUM_TERMS, Vp.data(), Vp.size() * sizeof(double));
    }
    return XMD_OK;
}

xmd_status xmd_step(xmd_engine *e, int64_t nsteps) {
    if (!e || nsteps < 0) return XMD_ERR_ARG;
    for (Part *q : e->parts) {
        xmd_status st = on_part(e, q, run_guarded(q->err, [&]() { q->ensure_lists(); }));
        if (st != XMD_OK) return st;
    }
    // interleave the parts' graph launches so their streams run side by side
    std::vector<int64_t> left(e->parts.size(), nsteps);
    bool more = nsteps > 0;
    while (more) {
        more = false;
        for (size_t g = 0; g < e->parts.size(); ++g) {
            Part *q = e->parts[g];
            if (left[g] <= 0) continue;
            xmd_status st = on_part(e, q, run_guarded(q->err, [&]() { left[g] -= q->launch_next(left[g]); }));
            if (st != XMD_OK) return st;
            more |= left[g] > 0;
        }
    }
    // wait for every group, then report new errors of all groups
    xmd_status first = XMD_OK;
    std::string msg;
    for (Part *q : e->parts) {
        xmd_status st = run_guarded(q->err, [&]() {
            q->drain();
            q->settle();
        });
        if (st != XMD_OK) {
            if (first == XMD_OK) first = st;
            msg += q->err;
        }
    }
    if (first != XMD_OK) e->err = msg;
    return first;
}

xmd_status xmd_replica_status(xmd_engine *e, int32_t rep) {
    if (!e) return XMD_ERR_ARG;
    if (rep < 0 || rep >= e->nrep) {
        e->err = "bad replica";
        return XMD_ERR_ARG;
    }
    Part *q = e->parts[e->rep_part[rep]];
    const int lr = e->rep_local[rep];
    return on_part(e, q, run_guarded(q->err, [&]() {
        HIPCHK(hipStreamSynchronize(q->stream));
        int w = 0;
        HIPCHK(hipMemcpy(&w, q->D.err + lr, sizeof(int), hipMemcpyDeviceToHost));
        if (w) {
            xmd_status st;
            const std::string d = Part::describe(w, st);
            throw Fail{st, "replica " + std::to_string(rep) + ": " + d + " (replica halted until its state is set)"};
        }
    }));
}

xmd_status xmd_get_energy(xmd_engine *e, double *V) {
    if (!e || !V) return XMD_ERR_ARG;
    for (size_t g = 0; g < e->parts.size(); ++g) {
        xmd_status st = on_part(e, e->parts[g], part_get_energy(e->parts[g], V + e->part_rep0[g]));
        if (st != XMD_OK) return st;
    }
    return XMD_OK;
}

xmd_status xmd_get_counters(xmd_engine *e, int64_t *out) {
    if (!e || !out) return XMD_ERR_ARG;
    for (int k = 0; k < 5; ++k) out[k] = 0;
    for (Part *q : e->parts) {
        int64_t c[5];
        xmd_status st = on_part(e, q, part_get_counters(q, c));
        if (st != XMD_OK) return st;
        out[0] = c[0];
        for (int k = 1; k < 5; ++k) out[k] += c[k];
    }
    return XMD_OK;
}

xmd_status xmd_profile(xmd_engine *e, int64_t nsteps, char *names, int64_t names_cap, int64_t *counts,
                       double *total_ms, int32_t cap, int32_t *n_kernels) {
    if (!e || nsteps < 0 || !n_kernels) return XMD_ERR_ARG;
    // each part in turn, eagerly; kernels merged by name
    std::vector<std::string> nm;
    std::vector<int64_t> cnt;
    std::vector<double> ms;
    for (Part *q : e->parts) {
        std::vector<char> nb(8192);
        std::vector<int64_t> c(256);
        std::vector<double> t(256);
        int32_t nk = 0;
        xmd_status st = on_part(e, q, part_profile(q, nsteps, nb.data(), (int64_t)nb.size(), c.data(), t.data(), 256, &nk));
        if (st != XMD_OK) return st;
        std::string all(nb.data());
        size_t pos = 0;
        for (int k = 0; k < nk && k < 256; ++k) {
            size_t end = all.find('\n', pos);
            std::string name = all.substr(pos, end - pos);
            pos = end + 1;
            size_t i = std::find(nm.begin(), nm.end(), name) - nm.begin();
            if (i == nm.size()) {
                nm.push_back(name);
                cnt.push_back(0);
                ms.push_back(0.0);
            }
            cnt[i] += c[k];
            ms[i] += t[k];
        }
    }
    *n_kernels = (int32_t)nm.size();
    std::string all;
    for (size_t k = 0; k < nm.size(); ++k) {
        if ((int32_t)k < cap) {
            if (counts) counts[k] = cnt[k];
            if (total_ms) total_ms[k] = ms[k];
        }
        all += nm[k] + "\n";
    }
    if (names && names_cap > 0) {
        size_t m = std::min<size_t>(all.size(), (size_t)names_cap - 1);
        std::memcpy(names, all.data(), m);
        names[m] = 0;
    }
    return XMD_OK;
}

xmd_status xmd_observables(xmd_engine *e, double active_threshold, double *out) {
    if (!e || !out) return XMD_ERR_ARG;
    for (size_t g = 0; g < e->parts.size(); ++g) {
        Part *q = e->parts[g];
        xmd_status st = on_part(e, q, run_guarded(q->err, [&]() {
            Scratch<double> d(q, 8 * (size_t)q->nrep);
            k_observables<<<q->nrep, BLK, 0, q->stream>>>(q->D, q->G, q->cur, active_threshold, d.p);
            HIPCHK(hipGetLastError());
            auto h = q->get(d.p, 8 * (size_t)q->nrep);
            std::memcpy(out + 8 * (size_t)e->part_rep0[g], h.data(), h.size() * sizeof(double));
        }));
        if (st != XMD_OK) return st;
    }
    return XMD_OK;
}

xmd_status xmd_observables_noalloc(xmd_engine *e, double active_threshold, double *out) {
    if (!e || !out) return XMD_ERR_ARG;
    if (e->nrep > OBS_NOALLOC_MAX_REPLICAS) {
        e->err = "xmd_observables_noalloc: " + std::to_string(e->nrep) + " replicas, at most " +
                 std::to_string(OBS_NOALLOC_MAX_REPLICAS);
        return XMD_ERR_ARG;
    }
    // each stream group writes its own replicas' rows of the staging array
    for (size_t g = 0; g < e->parts.size(); ++g) {
        Part *q = e->parts[g];
        xmd_status st = on_part(e, q, run_guarded(q->err, [&]() {
            k_observables<<<q->nrep, BLK, 0, q->stream>>>(q->D, q->G, q->cur, active_threshold,
                                                          e->obs_scratch + 8 * (size_t)e->part_rep0[g]);
            HIPCHK(hipGetLastError());
        }));
        if (st != XMD_OK) return st;
    }
    for (size_t g = 0; g < e->parts.size(); ++g) {
        Part *q = e->parts[g];
        xmd_status st = on_part(e, q, run_guarded(q->err, [&]() {
            const size_t off = 8 * (size_t)e->part_rep0[g];
            HIPCHK(hipMemcpyAsync(out + off, e->obs_scratch + off, 8 * (size_t)q->nrep * sizeof(double),
                                  hipMemcpyDeviceToHost, q->stream));
            HIPCHK(hipStreamSynchronize(q->stream));
        }));
        if (st != XMD_OK) return st;
    }
    return XMD_OK;
}

xmd_status xmd_autotune(xmd_engine *e, int32_t reps, int32_t *lpb_out, double *ms_out) {
    if (!e || reps <= 0) return XMD_ERR_ARG;
    static const int cand[3] = {4, 8, 16};
    constexpr int ROUNDS = 3;
    const size_t np = e->parts.size();
    const int lpb0 = e->parts[0]->lpb;  // every part runs the same tile
    // Each candidate times whole graph-captured steps of every stream group
    // side by side, as xmd_step runs them (the force kernel alone times
    // differently: with 4 lanes per bead it overlaps a second group's launch
    // better but the whole step is slower), from the same state: every
    // part's state is copied first and restored before each sample and at
    // the end.  Candidates in interleaved rounds, each keeping its median.
    std::vector<Part::Snapshot> snap(np);
    std::vector<double> t[3];
    xmd_status st = XMD_OK;
    for (size_t g = 0; g < np && st == XMD_OK; ++g)
        st = on_part(e, e->parts[g], run_guarded(e->parts[g]->err, [&]() {
                         e->parts[g]->ensure_lists();
                         e->parts[g]->take(snap[g]);
                     }));
    // n steps of every part, interleaved, then a wait for all; no list
    // growth (a capacity error just halts its replica for the sample)
    auto run = [&](int64_t n) {
        std::vector<int64_t> left(np, n);
        bool more = n > 0;
        while (more && st == XMD_OK) {
            more = false;
            for (size_t g = 0; g < np && st == XMD_OK; ++g) {
                Part *q = e->parts[g];
                if (left[g] <= 0) continue;
                st = on_part(e, q, run_guarded(q->err, [&]() { left[g] -= q->launch_next(left[g]); }));
                more |= left[g] > 0;
            }
        }
        for (size_t g = 0; g < np && st == XMD_OK; ++g)
            st = on_part(e, e->parts[g], run_guarded(e->parts[g]->err, [&]() { e->parts[g]->drain(); }));
    };
    const int64_t chunk = e->parts[0]->chunk_steps, timed = (int64_t)reps * chunk;
    for (int r = -1; r < ROUNDS && st == XMD_OK; ++r)  // round -1 warms up
        for (int c = 0; c < 3 && st == XMD_OK; ++c) {
            for (size_t g = 0; g < np && st == XMD_OK; ++g)
                st = on_part(e, e->parts[g], run_guarded(e->parts[g]->err, [&]() {
                                 e->parts[g]->restore(snap[g]);
                                 e->parts[g]->set_lpb(cand[c]);
                             }));
            run(chunk);  // captures the candidate's graphs (not timed)
            const auto t0 = std::chrono::steady_clock::now();
            run(timed);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            // a sample in which a replica halted stepped fewer replicas
            bool halted = false;
            for (size_t g = 0; g < np && st == XMD_OK; ++g)
                st = on_part(e, e->parts[g], run_guarded(e->parts[g]->err, [&]() {
                                 halted |= e->parts[g]->halted_since(snap[g]);
                             }));
            if (r >= 0) t[c].push_back(halted ? std::numeric_limits<double>::infinity() : ms / (double)timed);
        }
    // every step taken here undone, also after a failed sample (which
    // leaves the tile as it was)
    for (size_t g = 0; g < np; ++g) {
        if (!snap[g].of.empty()) {
            const xmd_status rs = on_part(e, e->parts[g], run_guarded(e->parts[g]->err, [&]() {
                                              e->parts[g]->restore(snap[g]);
                                              if (st != XMD_OK) e->parts[g]->set_lpb(lpb0);
                                          }));
            if (st == XMD_OK) st = rs;
        }
        Part::discard(snap[g]);
    }
    if (st != XMD_OK) return st;
    // the fastest tile, unless the default is within 3 % of it: tiles sum
    // a bead's partial forces in different orders, so a choice that follows
    // timing noise would make the trajectory differ from run to run
    double best = 0.0, dflt = 0.0;
    int pick = FORCE_LPB;
    for (int c = 0; c < 3; ++c) {
        std::sort(t[c].begin(), t[c].end());
        const double med = t[c][ROUNDS / 2];
        if (ms_out) ms_out[c] = med;
        if (cand[c] == FORCE_LPB) dflt = med;
        if (c == 0 || med < best) {
            best = med;
            pick = cand[c];
        }
    }
    if (dflt <= 1.03 * best) pick = FORCE_LPB;
    for (Part *q : e->parts) {
        st = on_part(e, q, run_guarded(q->err, [&]() { q->set_lpb(pick); }));
        if (st != XMD_OK) return st;
    }
    if (lpb_out) *lpb_out = pick;
    return XMD_OK;
}

xmd_status xmd_rand_normals(xmd_engine *e, uint64_t state, int64_t nb, float *out, uint64_t *state_out) {
    if (!e) return XMD_ERR_ARG;
    return on_part(e, e->parts[0], part_rand_normals(e->parts[0], state, nb, out, state_out));
}

}  // extern "C"
