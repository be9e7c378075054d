// Batched local search: the spin_glass_solver loop
// (pyqrackising/spin_glass_solver.py:603-727) run on many instances at once.
//
// Every phase is one launch over the list of instances that take that branch
// of the reference loop in the current round; the host reads back one small
// record per instance per phase and builds the next list.  Each instance's
// moves are exactly those pqi_local_search makes on it alone: the kernels are
// the same bodies, only indexed by instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pqi_internal.h"

namespace pqi {

// One job at a time on a thread of its own, kept for the context's lifetime
// (no thread creation, and the HIP runtime's per-thread set-up done once).
struct Worker {
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::function<void()> job;
    bool pending = false, done = true, quit = false;

    Worker() {
        th = std::thread([this]() {
            for (;;) {
                std::function<void()> f;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [this]() { return pending || quit; });
                    if (!pending) return;
                    f = std::move(job);
                    pending = false;
                }
                f();
                {
                    std::lock_guard<std::mutex> lk(mu);
                    done = true;
                }
                cv.notify_all();
            }
        });
    }
    void run(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> lk(mu);
            job = std::move(f);
            pending = true;
            done = false;
        }
        cv.notify_all();
    }
    void wait() {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [this]() { return done; });
    }
};

void worker_stop(Worker* w) {
    w->wait();
    {
        std::lock_guard<std::mutex> lk(w->mu);
        w->quit = true;
    }
    w->cv.notify_all();
    w->th.join();
    delete w;
}

namespace {

struct Arena {
    void* base = nullptr;
    size_t used = 0, cap = 0;
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

int64_t reference_steps(int64_t T, int64_t n) {
    const int64_t nb = (n + 63) >> 6;
    int64_t total = 0;
    for (int p = 0; p < 63; ++p) {
        const int64_t cnt = (T >> p) - (T >> (p + 1));
        if (!cnt) continue;
        int64_t valid = 0;
        for (int64_t b = 0; b < nb; ++b) valid += (p + 64 * b < n) ? 1 : 0;
        total += cnt * valid;
    }
    return total;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

}  // namespace pqi

using namespace pqi;

#define BCHK(expr)                                                                  \
    do {                                                                            \
        hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess) {                                                     \
            std::string _m = std::string(#expr) + ": " + hipGetErrorString(_e);     \
            return fail(_e == hipErrorOutOfMemory ? PQI_ERR_OOM : PQI_ERR_HIP, _m); \
        }                                                                           \
    } while (0)

namespace pqi {

// The batched loop.  With `net`, a learned phase runs first: up to `passes`
// rounds of (features, network scores, score order, one first-improvement pass
// with exact gains over the spins scored above min_score), stopping early when
// no instance flips.  With `targets`,
// an instance stops as soon as its value (energy in spin-glass mode, cut
// otherwise) reaches its target.  `set` selects the context's resources: 0 its
// stream, batch arena and staging; 1 the second stream's, so that a second
// search can run beside a first one.  `first_descent`, with targets, is called
// once after the first round's descent with the instances still below their
// targets (before the rest of the round).
using BelowFn = std::function<void(const std::vector<int32_t>&)>;
static int run_batched(pqi_ctx* c, pqi_graph* const* graphs, int64_t count, uint8_t* const* thetas,
                       const pqi_search_params* p, pqi_search_stats* stats, pqi_mpnn* net, const double* targets,
                       int64_t passes, float min_score, int set = 0, const BelowFn* first_descent = nullptr) {
    if (!c || !graphs || !thetas || !p || !stats) return fail(PQI_ERR_INVALID, "null argument");
    if (count < 1) return fail(PQI_ERR_INVALID, "count must be >= 1");
    if (p->walkers < 1 || p->seed_multiple < 1) return fail(PQI_ERR_INVALID, "walkers and seed_multiple must be >= 1");
    if (p->max_rounds < 0 && !net) return fail(PQI_ERR_INVALID, "max_rounds must be >= 0");
    if (p->double_mode != PQI_DOUBLE_REFERENCE)
        return fail(PQI_ERR_UNSUPPORTED, "the batched search implements the reference double-flip mode");
    const int kind = graphs[0]->kind;
    for (int64_t i = 0; i < count; ++i) {
        if (!graphs[i] || graphs[i]->ctx != c) return fail(PQI_ERR_INVALID, "graph of another context");
        if (graphs[i]->kind != kind) return fail(PQI_ERR_INVALID, "all graphs must share one weight kind");
        if (graphs[i]->n < 2) return fail(PQI_ERR_INVALID, "n must be >= 2");
        if (!thetas[i]) return fail(PQI_ERR_INVALID, "null theta");
        for (int64_t v = 0; v < graphs[i]->n; ++v)
            if (thetas[i][v] > 1) return fail(PQI_ERR_INVALID, "theta entries must be 0 or 1");
    }
    if (hipSetDevice(c->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = set ? c->stream2 : c->stream;
    if (!s) return fail(PQI_ERR_INVALID, "no second stream");
    void* arena = nullptr;
    const size_t fb = kind == PQI_WEIGHTS_I8 ? 4 : 8;
    const int64_t G = p->walkers;
    const double scale = p->is_spin_glass ? 2.0 : 1.0;

    // per-instance sizes and arena layout
    std::vector<BInstRaw> h_inst(count);
    std::vector<size_t> off_x(count), xoff(count);
    size_t total = 0, xbytes = 0;
    int64_t max_n = 0, max_seeds = 0;
    for (int64_t i = 0; i < count; ++i) {
        const pqi_graph* g = graphs[i];
        const int64_t n = g->n, nb = (n + 63) >> 6;
        const int64_t bs = G * p->seed_multiple, seeds = bs * nb;
        if (G > seeds) return fail(PQI_ERR_INVALID, "more walkers than seeds");
        const int64_t gi = p->gray_iterations < 0 ? n * n : p->gray_iterations;
        BInstRaw& r = h_inst[i];
        std::memset(&r, 0, sizeof(r));
        r.W = g->W;
        r.ld = g->n_pad;
        r.n = n;
        r.ap = g->adj_ptr;
        r.ai = g->adj_idx;
        r.aw = g->adj_w;
        r.use_adj = (g->use_adj || g->W == nullptr) ? 1 : 0;
        r.kadj = g->key_adj;
        r.kstride = g->key_adj ? g->key_adj_stride : 0;
        r.seed_count = seeds;
        r.bs = bs;
        r.G = G;
        r.T = (gi + G - 1) / G;
        max_n = std::max(max_n, n);
        max_seeds = std::max(max_seeds, seeds);
        off_x[i] = total;
        total += align256(n * fb) * 2 + align256(n) + align256(G * n) + align256(G * n * fb) + align256(G * 8) +
                 align256(seeds * 8) + align256(seeds * 4);
        if (net) total += align256(n * 4);
        xoff[i] = xbytes;  // best states, contiguous over the instances (one copy in, one out)
        xbytes += align256(n);
    }
    if (net) {
        int64_t off = 0;
        for (int64_t i = 0; i < count; ++i) {
            if (!graphs[i]->adj_ptr) return fail(PQI_ERR_UNSUPPORTED, "learned proposals need adjacency lists");
            if (mpnn_node_offset(net, i) != off) return fail(PQI_ERR_INVALID, "the network is bound to other graphs");
            off += graphs[i]->n;
        }
        if (mpnn_total_nodes(net) != off) return fail(PQI_ERR_INVALID, "the network is bound to other graphs");
    }
    int all_sparse = 1;
    for (int64_t i = 0; i < count; ++i) all_sparse &= h_inst[i].use_adj;
    // packed-key descent (int8 weights): one LDS size for the launch, the
    // largest staged layout when that fits, else the largest unstaged one;
    // each work group stages its adjacency when its own layout fits
    size_t key_lds = 0;
    {
        const size_t limit = b_key_lds_limit();
        size_t staged = 0, base = 0;
        bool keyed = kind == PQI_WEIGHTS_I8;
        for (int64_t i = 0; i < count && keyed; ++i) {
            h_inst[i].key_shift = key_shift(*graphs[i]);
            keyed = h_inst[i].key_shift > 0;
            staged = std::max(staged, key_lds_bytes(*graphs[i], SIZE_MAX));
            base = std::max(base, key_lds_bytes(*graphs[i], 0));
        }
        if (keyed) key_lds = staged <= limit ? staged : (base <= limit ? base : 0);
    }
    // register-resident descent when every instance qualifies: the largest
    // register count and table size over the instances
    int reg_r = 0;
    int64_t reg_stride = 0;
    size_t reg_lds = 0;
    if (key_lds > 0) {
        for (int64_t i = 0; i < count; ++i) {
            size_t t = 0;
            const int r = reg_descent_plan(*graphs[i], &t);
            if (!r) {
                reg_r = 0;
                break;
            }
            reg_r = std::max(reg_r, r);
            reg_stride = std::max(reg_stride, graphs[i]->key_adj_stride);
            reg_lds = std::max(reg_lds, t);
        }
    }
    // bit-plane descent when every instance qualifies: planes for the
    // largest row strength (max_n decides the words per lane)
    int plane_s = -1;
    if (reg_r > 0) {
        int S = -1;
        for (int64_t i = 0; i < count; ++i) {
            if (!plane_descent_plan(*graphs[i], nullptr)) {
                S = -1;
                break;
            }
            S = std::max<int>(S, (int)graphs[i]->max_strength);
        }
        if (S >= 0) {
            int half = 1;
            for (int64_t i = 0; i < count; ++i) half &= plane_half(*graphs[i], S);
            plane_s = 2 * S + half;
        }
    }
    const int64_t L0 = count;
    // shared records, one per instance in instance order, so that a phase's
    // results come back in one copy: phase records, top seed gains and walker
    // totals (G per instance), ages of the learned phase
    const size_t rec_out = align256(L0 * sizeof(PhaseOut)), rec_g = align256(L0 * G * 8);
    const size_t recs = xbytes + rec_out + 3 * rec_g + (net ? xbytes : 0);
    // learned passes between two synchronisations (one record per pass each)
    constexpr int64_t PG = 4;
    const size_t lp_gate = net ? align256(L0 * sizeof(LpGate)) : 0, lp_outs = net ? align256(L0 * PG * sizeof(PhaseOut)) : 0;
    const size_t misc = align256(L0 * 4) + align256(L0 * 16) + align256(L0 * 4) + align256(L0 * 8) +
                        align256(L0 * sizeof(BInstRaw)) + lp_gate + lp_outs;
    {
        const int e = batch_arena(c, set, total + recs + misc, &arena);
        if (e != PQI_OK) return e;
    }
    // pinned host side: the states, then the records, then a bump area for
    // the lists and moves uploaded between two synchronisations
    const size_t bump_cap = align256(L0 * 16) * 6;
    {
        const int e = ensure_pinned(c, set ? 3 : 2, xbytes + rec_out + 3 * rec_g + align256(L0 * 8) + bump_cap + lp_outs);
        if (e != PQI_OK) return e;
    }
    char* pin = (char*)c->pin[set ? 3 : 2].p;
    uint8_t* pin_x = (uint8_t*)pin;
    PhaseOut* pin_out = (PhaseOut*)(pin + xbytes);
    double* pin_seld = (double*)(pin + xbytes + rec_out);
    double* pin_we = (double*)(pin + xbytes + rec_out + rec_g);
    int64_t* pin_wev = (int64_t*)(pin + xbytes + rec_out + 2 * rec_g);
    double* pin_q = (double*)(pin + xbytes + rec_out + 3 * rec_g);
    char* pin_bump = pin + xbytes + rec_out + 3 * rec_g + align256(L0 * 8);
    PhaseOut* pin_lp = (PhaseOut*)(pin_bump + bump_cap);
    size_t bump_used = 0;
    char* a = (char*)arena;
    char* rec = a + total;
    uint8_t* d_xall = (uint8_t*)rec;
    PhaseOut* d_out = (PhaseOut*)(rec + xbytes);
    double* d_seld = (double*)(rec + xbytes + rec_out);
    double* d_we = (double*)(rec + xbytes + rec_out + rec_g);
    int64_t* d_wev = (int64_t*)(rec + xbytes + rec_out + 2 * rec_g);
    uint8_t* d_age = net ? (uint8_t*)(rec + xbytes + rec_out + 3 * rec_g) : nullptr;
    for (int64_t i = 0; i < count; ++i) {
        BInstRaw& r = h_inst[i];
        const int64_t n = r.n;
        char* q = a + off_x[i];
        auto take = [&](size_t bytes) {
            char* t = q;
            q += align256(bytes);
            return t;
        };
        r.x = d_xall + xoff[i];
        r.h = take(n * fb);
        r.rx = (uint8_t*)take(n);
        r.rh = take(n * fb);
        r.wx = (uint8_t*)take(G * n);
        r.wh = take(G * n * fb);
        r.we = d_we + i * G;
        r.wev = d_wev + i * G;
        r.sel = (int64_t*)take(G * 8);
        r.seld = d_seld + i * G;
        r.delta = (double*)take(r.seed_count * 8);
        r.rank = (int32_t*)take(r.seed_count * 4);
        r.out = d_out + i;
        if (net) {
            r.perm = (int32_t*)take(n * 4);
            r.age = d_age + xoff[i];
            r.feat = mpnn_features(net) + mpnn_node_offset(net, i) * 4;
            r.score = mpnn_scores(net) + mpnn_node_offset(net, i);
        }
    }
    char* m = a + total + recs;
    int32_t* d_list = (int32_t*)m;
    m += align256(L0 * 4);
    int64_t* d_moves = (int64_t*)m;
    m += align256(L0 * 16);
    int32_t* d_src = (int32_t*)m;
    m += align256(L0 * 4);
    double* d_q = (double*)m;
    m += align256(L0 * 8);
    BInstRaw* d_inst = (BInstRaw*)m;
    m += align256(L0 * sizeof(BInstRaw));
    LpGate* d_gate = (LpGate*)m;
    m += lp_gate;
    PhaseOut* d_lp = (PhaseOut*)m;
    BCHK(hipMemcpyAsync(d_inst, h_inst.data(), count * sizeof(BInstRaw), hipMemcpyHostToDevice, s));
    for (int64_t i = 0; i < count; ++i) std::memcpy(pin_x + xoff[i], thetas[i], h_inst[i].n);
    BCHK(hipMemcpyAsync(d_xall, pin_x, xbytes, hipMemcpyHostToDevice, s));

    std::vector<int32_t> list(count);
    for (int64_t i = 0; i < count; ++i) list[i] = (int32_t)i;
    std::vector<PhaseOut> outs(count);
    std::vector<int64_t> moves(2 * count);
    std::vector<int32_t> src(count);
    std::vector<double> qv(count);
    std::memset(stats, 0, sizeof(pqi_search_stats) * count);
    for (int64_t i = 0; i < count; ++i) stats[i].learned_min_gain = INFINITY;

    // Host -> device uploads go through the pinned bump area, so they are
    // asynchronous; the area is reused after each synchronisation (and, should
    // it fill up, after an extra one).  Device -> host results of a phase come
    // back in one copy of the instance-ordered records.
    auto sync = [&]() -> hipError_t {
        bump_used = 0;
        return hipStreamSynchronize(s);
    };
    auto upload = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        const size_t need = align256(bytes);
        if (bump_used + need > bump_cap) {
            const hipError_t e = sync();
            if (e != hipSuccess) return e;
        }
        char* hp = pin_bump + bump_used;
        bump_used += need;
        std::memcpy(hp, src, bytes);
        return hipMemcpyAsync(dst, hp, bytes, hipMemcpyHostToDevice, s);
    };
    // the list now in d_list (only this lambda writes it): a phase over the
    // same instances as the previous one uploads nothing
    std::vector<int32_t> list_on_device;
    bool list_known = false;
    auto upload_list = [&](const std::vector<int32_t>& l) -> hipError_t {
        if (list_known && l == list_on_device) return hipSuccess;
        const hipError_t e = upload(d_list, l.data(), l.size() * 4);
        list_known = e == hipSuccess;
        if (list_known) list_on_device = l;
        return e;
    };
    auto read_outs = [&](const std::vector<int32_t>& l) -> hipError_t {
        hipError_t e = hipMemcpyAsync(pin_out, d_out, count * sizeof(PhaseOut), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = sync();
        for (size_t t = 0; t < l.size() && e == hipSuccess; ++t) outs[t] = pin_out[l[t]];
        return e;
    };
    auto flips = [&](const std::vector<int32_t>& l, const std::vector<int64_t>& mv, int which) -> hipError_t {
        hipError_t e = upload_list(l);
        if (e == hipSuccess) e = upload(d_moves, mv.data(), l.size() * 16);
        if (e == hipSuccess) launch_b_flip(kind, d_inst, d_list, (int64_t)l.size(), d_moves, which, s);
        return e;
    };
    auto copies = [&](const std::vector<int32_t>& l, const std::vector<int32_t>& sr, int dst) -> hipError_t {
        hipError_t e = upload_list(l);
        if (e == hipSuccess) e = upload(d_src, sr.data(), l.size() * 4);
        if (e == hipSuccess) launch_b_copy(kind, d_inst, d_list, (int64_t)l.size(), max_n, d_src, dst, s);
        return e;
    };
    auto score = [&](std::vector<int32_t>& l) -> hipError_t {
        hipError_t e = upload_list(l);
        if (e != hipSuccess) return e;
        launch_b_quad(kind, d_inst, d_list, (int64_t)l.size(), d_q, s);
        e = hipMemcpyAsync(pin_q, d_q, l.size() * 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = sync();
        if (e == hipSuccess) std::memcpy(qv.data(), pin_q, l.size() * 8);
        return e;
    };

    // fields and starting values
    BCHK(upload_list(list));
    launch_b_field(kind, d_inst, d_list, count, max_n, all_sparse, s);
    BCHK(score(list));
    for (int64_t i = 0; i < count; ++i) {
        const pqi_graph* g = graphs[i];
        const double cut = kind == PQI_WEIGHTS_I8 ? (double)((2 * g->w_up_i - (long long)qv[i]) / 4)
                                                  : (2.0 * g->w_up - qv[i]) / 4.0;
        const double en = kind == PQI_WEIGHTS_I8 ? (double)(-(long long)qv[i] / 2) : -qv[i] / 2.0;
        stats[i].max_energy = p->is_spin_glass ? en : cut;
    }

    auto reached = [&](int32_t i) { return targets && stats[i].max_energy >= targets[i]; };
    int64_t learned_moves = 0;

    // learned phase
    if (net && passes > 0) {
        const double t0 = now_s();
        BCHK(hipMemsetAsync(d_age, 64, xbytes, s));  // every node starts "long ago flipped"
        std::vector<int32_t> run;
        for (int32_t i : list)
            if (!reached(i)) run.push_back(i);
        // Up to PG passes between two synchronisations: each instance's gate
        // (its value, target and activity) is kept on the device, a pass over
        // an inactive instance returns at once, and the host then replays the
        // group's records with the same rule (an instance goes on while its
        // pass moved and it is below its target).
        if (!run.empty()) {
            std::vector<LpGate> gates(count);
            for (int64_t i = 0; i < count; ++i) {
                gates[i].energy = stats[i].max_energy;
                gates[i].target = targets ? targets[i] : INFINITY;
                gates[i].scale = scale;
                gates[i].active = 0;
                gates[i].pad = 0;
                gates[i].outs = d_lp + i * PG;
            }
            for (int32_t i : run) gates[i].active = 1;
            BCHK(upload(d_gate, gates.data(), count * sizeof(LpGate)));
        }
        for (int64_t pass0 = 0; pass0 < passes && !run.empty(); pass0 += PG) {
            const int64_t group = std::min<int64_t>(PG, passes - pass0);
            const std::vector<int32_t> launched = run;  // a superset of every pass's actives in the group
            for (int64_t q = 0; q < group; ++q) {
                BCHK(upload_list(list));
                launch_b_features(kind, d_inst, d_list, count, max_n, s);
                mpnn_forward_device(net, s);
                BCHK(upload_list(launched));
                launch_b_order(d_inst, d_list, (int64_t)launched.size(), s);
                launch_b_learned_pass(kind, d_inst, d_list, (int64_t)launched.size(), max_n, min_score, d_gate, (int)q,
                                      s);
            }
            BCHK(hipMemcpyAsync(pin_lp, d_lp, count * PG * sizeof(PhaseOut), hipMemcpyDeviceToHost, s));
            BCHK(sync());
            for (int64_t q = 0; q < group && !run.empty(); ++q) {
                std::vector<int32_t> again;
                for (int32_t i : run) {
                    const PhaseOut& o = pin_lp[i * PG + q];
                    pqi_search_stats& st = stats[i];
                    st.learned_accepts += o.accepts;
                    learned_moves += o.accepts;
                    st.learned_passes++;
                    if (o.accepts > 0) st.learned_min_gain = std::min(st.learned_min_gain, o.gain);
                    st.max_energy += scale * o.gain_sum;
                    if (o.accepts > 0 && !reached(i)) again.push_back(i);
                }
                run.swap(again);
            }
        }
        const double dt = now_s() - t0;
        for (int64_t i = 0; i < count; ++i) stats[i].seconds_learned += dt / (double)count;
    }

    // with PQI_FLAG_TIME_KERNELS, every descent launch of the call is timed by
    // events it records itself (its duration, not the host's)
    const bool timing = (p->flags & PQI_FLAG_TIME_KERNELS) != 0;
    hipEvent_t dev0 = nullptr, dev1 = nullptr;
    if (timing && (hipEventCreateWithFlags(&dev0, hipEventDisableSystemFence) != hipSuccess ||
                   hipEventCreateWithFlags(&dev1, hipEventDisableSystemFence) != hipSuccess))
        return fail(PQI_ERR_HIP, "event creation failed");
    struct EvFree {
        hipEvent_t a, b;
        ~EvFree() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev_free{dev0, dev1};
    double desc_s = 0.0;
    int64_t desc_launches = 0, desc_pairs = 0, desc_writes = 0;
    auto descent_time = [&]() {
        float ms = 0.f;
        if (timing && hipEventElapsedTime(&ms, dev0, dev1) == hipSuccess) {
            desc_s += ms * 1e-3;
            ++desc_launches;
        }
    };
    std::vector<int32_t> active;
    for (int32_t i : list)
        if (!reached(i) && !(net && p->max_rounds < 0)) active.push_back(i);
    // the fields are the exact ones of the starts until a move is made
    bool field_exact = learned_moves == 0, first_round = true;
    std::vector<double> walk_gain(count), reheat(count);
    std::vector<double> we(G);
    std::vector<int64_t> wev(G);
    while (!active.empty()) {
        std::vector<int32_t> next;
        if (p->max_rounds > 0) {
            std::vector<int32_t> keep;
            for (int32_t i : active)
                if (stats[i].rounds < p->max_rounds) keep.push_back(i);
            active.swap(keep);
            if (active.empty()) break;
        }
        for (int32_t i : active) stats[i].rounds++;
        if (kind != PQI_WEIGHTS_I8) {
            // a new round: the fp64 fields are recomputed exactly (as the
            // single-instance search does), so drift never carries over.  In
            // the first round with no learned move before it, the fields are
            // still the exact ones computed from the starts, so the launch
            // (which would reproduce them bit for bit) is skipped
            if (!(field_exact && first_round)) {
                BCHK(upload_list(active));
                launch_b_field(kind, d_inst, d_list, (int64_t)active.size(), max_n, all_sparse, s);
            }
            for (int32_t i : active) stats[i].field_checks++;
        }
        first_round = false;

        // single-flip descent, bounded launches repeated while the bound is hit
        double t0 = now_s();
        std::vector<int32_t> run = active;
        while (!run.empty()) {
            const int64_t cap = 16 * max_n + 4096;
            BCHK(upload_list(run));
            launch_b_descent(kind, d_inst, d_list, (int64_t)run.size(), max_n, cap, 0, all_sparse, key_lds, reg_r,
                             reg_stride, reg_lds, plane_s, s, dev0, dev1);
            BCHK(read_outs(run));
            descent_time();
            desc_pairs += (int64_t)run.size();
            std::vector<int32_t> again;
            for (size_t t = 0; t < run.size(); ++t) {
                if (outs[t].accepts > 0) ++desc_writes;
                pqi_search_stats& st = stats[run[t]];
                st.single_evals += outs[t].evals;
                st.single_accepts += outs[t].accepts;
                st.max_energy += scale * outs[t].gain_sum;
                if (outs[t].accepts == cap) again.push_back(run[t]);
            }
            run.swap(again);
        }
        const double t_single = now_s() - t0;
        // with targets, an instance whose descent reached its target is done
        const std::vector<int32_t> round_set = active;
        if (targets) {
            std::vector<int32_t> open;
            for (int32_t i : active)
                if (!reached(i)) open.push_back(i);
            active.swap(open);
            if (first_descent) {
                (*first_descent)(active);
                first_descent = nullptr;
            }
            if (active.empty()) {
                for (int32_t i : round_set) stats[i].seconds_single += t_single / (double)round_set.size();
                break;
            }
        }

        // double flips
        t0 = now_s();
        BCHK(upload_list(active));
        launch_b_double(kind, d_inst, d_list, (int64_t)active.size(), 0, s);
        BCHK(read_outs(active));
        std::vector<int32_t> acc, rest;
        std::vector<int64_t> mv;
        for (size_t t = 0; t < active.size(); ++t) {
            const int32_t i = active[t];
            const int64_t n = h_inst[i].n, C = n * (n - 1) / 2;
            stats[i].double_evals += std::min(G, C);
            if (outs[t].gain > 0.0) {
                acc.push_back(i);
                mv.push_back(outs[t].k);
                mv.push_back(outs[t].l);
                stats[i].max_energy += scale * outs[t].gain;
                stats[i].double_accepts++;
            } else {
                rest.push_back(i);
            }
        }
        if (!acc.empty()) BCHK(flips(acc, mv, 0));
        next.insert(next.end(), acc.begin(), acc.end());
        const double t_double = now_s() - t0;

        // Gray seeds
        t0 = now_s();
        std::vector<int32_t> rest2, seeded;
        if (!rest.empty()) {
            BCHK(upload_list(rest));
            launch_b_seeds(kind, d_inst, d_list, (int64_t)rest.size(), max_seeds, G, s);
            BCHK(hipMemcpyAsync(pin_seld, d_seld, (size_t)count * G * 8, hipMemcpyDeviceToHost, s));
            BCHK(sync());
            std::vector<double> top(rest.size());
            for (size_t t = 0; t < rest.size(); ++t) top[t] = pin_seld[(size_t)rest[t] * G];
            for (size_t t = 0; t < rest.size(); ++t) {
                const int32_t i = rest[t];
                stats[i].seed_evals += h_inst[i].seed_count;
                if (top[t] > 0.0) {
                    seeded.push_back(i);
                    stats[i].max_energy += scale * top[t];
                    stats[i].seed_accepts++;
                } else {
                    rest2.push_back(i);
                }
            }
            if (!seeded.empty()) {
                BCHK(upload_list(seeded));
                launch_b_walk(kind, d_inst, d_list, (int64_t)seeded.size(), max_n, G, 0, 1, reg_r, reg_stride, reg_lds,
                              s);
                std::vector<int32_t> zero(seeded.size(), 0);
                BCHK(copies(seeded, zero, 0));
                next.insert(next.end(), seeded.begin(), seeded.end());
            }
        }
        const double t_seeds = now_s() - t0;

        // Gray walkers
        t0 = now_s();
        std::vector<int32_t> rest3, walked;
        std::vector<int32_t> best_walker(count, 0);
        if (!rest2.empty()) {
            BCHK(upload_list(rest2));
            launch_b_walk(kind, d_inst, d_list, (int64_t)rest2.size(), max_n, G, -1, -1, reg_r, reg_stride, reg_lds,
                          s);
            BCHK(hipMemcpyAsync(pin_we, d_we, (size_t)count * G * 8, hipMemcpyDeviceToHost, s));
            BCHK(hipMemcpyAsync(pin_wev, d_wev, (size_t)count * G * 8, hipMemcpyDeviceToHost, s));
            BCHK(sync());
            std::vector<double> all_we(rest2.size() * G);
            std::vector<int64_t> all_wev(rest2.size() * G);
            for (size_t t = 0; t < rest2.size(); ++t) {
                std::memcpy(&all_we[t * G], pin_we + (size_t)rest2[t] * G, G * 8);
                std::memcpy(&all_wev[t * G], pin_wev + (size_t)rest2[t] * G, G * 8);
            }
            std::vector<int32_t> wsrc;
            for (size_t t = 0; t < rest2.size(); ++t) {
                const int32_t i = rest2[t];
                int64_t bw = 0;
                for (int64_t w = 0; w < G; ++w) {
                    stats[i].gray_evals_computed += all_wev[t * G + w];
                    if (all_we[t * G + w] > all_we[t * G + bw]) bw = w;
                }
                stats[i].gray_evals += G * reference_steps(h_inst[i].T, h_inst[i].n);
                best_walker[i] = (int32_t)bw;
                walk_gain[i] = all_we[t * G + bw];
                if (walk_gain[i] > 0.0) {
                    walked.push_back(i);
                    wsrc.push_back((int32_t)bw);
                    stats[i].max_energy += scale * walk_gain[i];
                    stats[i].walk_accepts++;
                } else {
                    rest3.push_back(i);
                }
            }
            if (!walked.empty()) BCHK(copies(walked, wsrc, 0));
            next.insert(next.end(), walked.begin(), walked.end());
        }
        const double t_walk = now_s() - t0;

        // post-reheat phase
        t0 = now_s();
        if (!rest3.empty()) {
            std::vector<int32_t> wsrc;
            for (int32_t i : rest3) {
                wsrc.push_back(best_walker[i]);
                reheat[i] = walk_gain[i];
            }
            BCHK(copies(rest3, wsrc, 1));
            BCHK(upload_list(rest3));
            launch_b_descent(kind, d_inst, d_list, (int64_t)rest3.size(), max_n, 0, 1, all_sparse, key_lds, reg_r,
                             reg_stride, reg_lds, plane_s, s, dev0, dev1);
            BCHK(read_outs(rest3));
            descent_time();
            desc_pairs += (int64_t)rest3.size();  // best-only descents: no move, no write
            std::vector<int32_t> take1, improve1, rest4;
            std::vector<int64_t> mv1, mvi;
            for (size_t t = 0; t < rest3.size(); ++t) {
                const int32_t i = rest3[t];
                stats[i].single_evals += outs[t].evals;
                if (outs[t].gain + reheat[i] > 0.0) {
                    take1.push_back(i);
                    mv1.push_back(outs[t].k);
                    mv1.push_back(-1);
                    stats[i].max_energy += scale * (reheat[i] + outs[t].gain);
                    stats[i].reheat_accepts++;
                } else {
                    if (outs[t].gain > 0.0) {
                        improve1.push_back(i);
                        mvi.push_back(outs[t].k);
                        mvi.push_back(-1);
                        reheat[i] += outs[t].gain;
                    }
                    rest4.push_back(i);
                }
            }
            if (!take1.empty()) {
                std::vector<int32_t> rsrc(take1.size(), -1);
                BCHK(copies(take1, rsrc, 0));
                BCHK(flips(take1, mv1, 0));
                next.insert(next.end(), take1.begin(), take1.end());
            }
            if (!improve1.empty()) BCHK(flips(improve1, mvi, 1));
            if (!rest4.empty()) {
                BCHK(upload_list(rest4));
                launch_b_double(kind, d_inst, d_list, (int64_t)rest4.size(), 1, s);
                BCHK(read_outs(rest4));
                std::vector<int32_t> take2;
                std::vector<int64_t> mv2;
                for (size_t t = 0; t < rest4.size(); ++t) {
                    const int32_t i = rest4[t];
                    const int64_t n = h_inst[i].n, C = n * (n - 1) / 2;
                    stats[i].double_evals += std::min(G, C);
                    if (outs[t].gain + reheat[i] > 0.0) {
                        take2.push_back(i);
                        mv2.push_back(outs[t].k);
                        mv2.push_back(outs[t].l);
                        stats[i].max_energy += scale * (reheat[i] + outs[t].gain);
                        stats[i].reheat_accepts++;
                    }
                }
                if (!take2.empty()) {
                    std::vector<int32_t> rsrc(take2.size(), -1);
                    BCHK(copies(take2, rsrc, 0));
                    BCHK(flips(take2, mv2, 0));
                    next.insert(next.end(), take2.begin(), take2.end());
                }
            }
        }
        const double t_reheat = now_s() - t0;
        const double share = 1.0 / (double)round_set.size();
        for (int32_t i : round_set) {
            stats[i].seconds_single += t_single * share;
            stats[i].seconds_double += t_double * share;
            stats[i].seconds_seeds += t_seeds * share;
            stats[i].seconds_walk += t_walk * share;
            stats[i].seconds_reheat += t_reheat * share;
        }
        std::sort(next.begin(), next.end());
        std::vector<int32_t> still;
        for (int32_t i : next)
            if (!reached(i)) still.push_back(i);
        active.swap(still);
    }

    // exact final scores (field recomputed from the state, as pqi_score does) and states
    BCHK(upload_list(list));
    launch_b_field(kind, d_inst, d_list, count, max_n, all_sparse, s);
    BCHK(score(list));
    for (int64_t i = 0; i < count; ++i) {
        const pqi_graph* g = graphs[i];
        const double cut = kind == PQI_WEIGHTS_I8 ? (double)((2 * g->w_up_i - (long long)qv[i]) / 4)
                                                  : (2.0 * g->w_up - qv[i]) / 4.0;
        const double en = kind == PQI_WEIGHTS_I8 ? (double)(-(long long)qv[i] / 2) : -qv[i] / 2.0;
        stats[i].cut_value = cut;
        stats[i].min_energy = p->is_spin_glass ? -en : en;
    }
    for (int64_t i = 0; i < count; ++i) {
        stats[i].descent_kernel_seconds = desc_s;  // the call's descent launches (all instances together)
        stats[i].descent_launches = desc_launches;
        stats[i].descent_instance_launches = desc_pairs;
        stats[i].descent_instance_writes = desc_writes;
    }
    BCHK(hipMemcpyAsync(pin_x, d_xall, xbytes, hipMemcpyDeviceToHost, s));
    BCHK(sync());
    for (int64_t i = 0; i < count; ++i) std::memcpy(thetas[i], pin_x + xoff[i], h_inst[i].n);
    return PQI_OK;
}

}  // namespace pqi

extern "C" {

int pqi_batch_local_search(pqi_ctx* c, pqi_graph* const* graphs, int64_t count, uint8_t* const* thetas,
                           const pqi_search_params* p, pqi_search_stats* stats) {
    return run_batched(c, graphs, count, thetas, p, stats, nullptr, nullptr, 0, 0.f);
}

int pqi_learned_search(pqi_ctx* c, pqi_mpnn* net, pqi_graph* const* graphs, int64_t count, uint8_t* const* thetas,
                       const pqi_search_params* p, const double* targets, int64_t passes, float min_score,
                       pqi_search_stats* stats) {
    if (!net) return fail(PQI_ERR_INVALID, "null network");
    if (passes < 0) return fail(PQI_ERR_INVALID, "passes must be >= 0");
    // no learned pass: the plain loop (with targets), the network is not used
    pqi_mpnn* use = passes > 0 || (p && p->max_rounds < 0) ? net : nullptr;
    const bool fallback = p && (p->flags & PQI_FLAG_FALLBACK) && targets && passes > 0 && count > 0 && thetas;
    if (!fallback) return run_batched(c, graphs, count, thetas, p, stats, use, targets, passes, min_score);
    // the starts, for the instances that fall short
    std::vector<std::vector<uint8_t>> starts(count);
    for (int64_t i = 0; i < count; ++i) {
        if (!graphs[i] || !thetas[i]) return fail(PQI_ERR_INVALID, "null graph or state");
        starts[i].assign(thetas[i], thetas[i] + graphs[i]->n);
    }
    if (!c->stream2) {
        if (hipSetDevice(c->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
        if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess)
            return fail(PQI_ERR_HIP, "hipStreamCreate failed");
    }
    // The plain search of an instance depends only on its start and target, so
    // it is started, on the second stream and a thread of its own, for every
    // instance still below its target after the learned search's first
    // descent, while the learned search finishes its round.  An instance that
    // ends below its target was below it then (the value only rises), so the
    // fallback covers it; the others' fallbacks are discarded.
    struct Fallback {
        std::vector<int32_t> list;
        std::vector<pqi_graph*> gs;
        std::vector<uint8_t*> ts;
        std::vector<double> tg;
        std::vector<pqi_search_stats> st;
        bool started = false;
        int rc = PQI_OK;
        std::string err;
    };
    // the fallback is the plain search: max_rounds < 0 (learned passes only)
    // limits the learned search alone
    pqi_search_params pf = *p;
    if (pf.max_rounds < 0) pf.max_rounds = 0;
    if (!c->worker) c->worker = new Worker();
    Worker* wk = c->worker;
    auto start_fallback = [&](Fallback& f, const std::vector<int32_t>& below, int set) {
        const int64_t k = (int64_t)below.size();
        f.list = below;
        f.gs.resize(k);
        f.ts.resize(k);
        f.tg.resize(k);
        f.st.resize(k);
        for (int64_t t = 0; t < k; ++t) {
            f.gs[t] = graphs[below[t]];
            f.ts[t] = starts[below[t]].data();
            f.tg[t] = targets[below[t]];
        }
        f.started = true;
        wk->run([&f, c, &pf, k, set]() {
            f.rc = run_batched(c, f.gs.data(), k, f.ts.data(), &pf, f.st.data(), nullptr, f.tg.data(), 0, 0.f, set);
            if (f.rc != PQI_OK) f.err = pqi_last_error();
        });
    };
    Fallback fb;
    const BelowFn hook = [&](const std::vector<int32_t>& below) {
        if (!below.empty()) start_fallback(fb, below, 1);
    };
    int rc = run_batched(c, graphs, count, thetas, p, stats, use, targets, passes, min_score, 0, &hook);
    if (fb.started) wk->wait();
    if (rc != PQI_OK) return rc;
    if (fb.rc != PQI_OK) return fail(fb.rc, fb.err);
    std::vector<int32_t> shortl;
    for (int64_t i = 0; i < count; ++i) {
        stats[i].fallback = 0;
        if (stats[i].max_energy < targets[i]) shortl.push_back((int32_t)i);
    }
    if (shortl.empty()) return PQI_OK;
    // (a short instance the early fallback did not cover: the learned search
    // never reached its first descent, e.g. max_rounds < 0; searched now)
    std::vector<int32_t> slot(count, -1);
    for (size_t t = 0; t < fb.list.size(); ++t) slot[fb.list[t]] = (int32_t)t;
    std::vector<int32_t> missing;
    for (int32_t i : shortl)
        if (slot[i] < 0) missing.push_back(i);
    Fallback fb2;
    if (!missing.empty()) {
        start_fallback(fb2, missing, 0);
        wk->wait();
        if (fb2.rc != PQI_OK) return fail(fb2.rc, fb2.err);
    }
    // the learned call's descent totals (the same in every record), and the
    // fallbacks' (each run's totals are the same in all of its records)
    double f_s = 0.0;
    int64_t f_launches = 0, f_pairs = 0, f_writes = 0;
    for (const Fallback* f : {&fb, &fb2})
        if (!f->st.empty()) {
            f_s += f->st[0].descent_kernel_seconds;
            f_launches += f->st[0].descent_launches;
            f_pairs += f->st[0].descent_instance_launches;
            f_writes += f->st[0].descent_instance_writes;
        }
    const double l_s = stats[0].descent_kernel_seconds;
    const int64_t l_launches = stats[0].descent_launches, l_pairs = stats[0].descent_instance_launches,
                  l_writes = stats[0].descent_instance_writes;
    for (int32_t i : shortl) {
        const bool early = slot[i] >= 0;
        int32_t t = early ? slot[i] : -1;
        if (!early)
            for (size_t u = 0; u < fb2.list.size(); ++u)
                if (fb2.list[u] == i) t = (int32_t)u;
        const Fallback& f = early ? fb : fb2;
        std::memcpy(thetas[i], starts[i].data(), graphs[i]->n);
        pqi_search_stats r = f.st[t];
        r.learned_accepts = stats[i].learned_accepts;
        r.learned_passes = stats[i].learned_passes;
        r.seconds_learned = stats[i].seconds_learned;
        r.learned_min_gain = stats[i].learned_min_gain;
        r.fallback = 1;
        r.replaced_single_accepts = stats[i].single_accepts;
        stats[i] = r;
    }
    // every record carries the call's descent totals: both searches
    for (int64_t i = 0; i < count; ++i) {
        stats[i].descent_kernel_seconds = l_s + f_s;
        stats[i].descent_launches = l_launches + f_launches;
        stats[i].descent_instance_launches = l_pairs + f_pairs;
        stats[i].descent_instance_writes = l_writes + f_writes;
    }
    return PQI_OK;
}

}  // extern "C"
