// Load driver for the serving loop: Poisson arrivals of alignment chunks from concurrent producers.
//
// A pool of chunks is generated up front (sizes log-uniform in [1k, 64k] jobs; 80% of chunks
// long-tail extension work, 20% mate-rescue work: the kind is drawn per chunk, independently of its
// size, so the expected share of jobs is 80/20 too, and the submitted job counts per kind are
// reported; targets given either as arena bytes or as
// windows of a genome resident in the server).  Each producer thread submits chunks from the pool
// at exponentially distributed intervals so the producers together offer `rate` jobs/s, and a
// waiter thread per producer collects its chunks.  With verification on, every chunk's results in
// the mix are compared with the same chunk run alone through bwg_extend_batch /
// bwg_ksw_align2_batch.  This is benchmark and test infrastructure, not part of libbwg.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../include/bwg.h"

extern "C" {
void *bwgw_generate_on(int kind, uint64_t genome_seed, uint64_t seed, uint64_t n);
uint64_t bwgw_arena_len(const void *h);
uint64_t bwgw_count(const void *h);
void bwgw_copy(const void *h, uint8_t *arena, uint64_t *q_off, uint32_t *q_len, uint64_t *t_off,
               uint32_t *t_len, int32_t *h0, uint8_t *kind);
void bwgw_copy_windows(const void *h, uint64_t *t_win);
void bwgw_free(void *h);
uint64_t bwgw_genome_len();
void bwgw_genome(uint64_t seed, uint8_t *out);
}

namespace {

using Clock = std::chrono::steady_clock;

struct PoolChunk {
    int kind = 0;
    size_t n = 0;
    std::vector<uint8_t> arena;
    std::vector<uint64_t> q_off, t_off;
    std::vector<uint32_t> q_len, t_len;
    std::vector<int32_t> h0;
    std::vector<uint8_t> alone;  // results of the chunk run alone
};

struct Pending {
    uint64_t ticket;
    const PoolChunk *c;
    std::vector<uint8_t> out;
};

}  // namespace

extern "C" {

// out[0..13]: p50 ms, p99 ms, mean ms, max ms, completed jobs/s, chunks, jobs, epochs,
// mismatching chunks, failed submissions, peak admitted device bytes in flight, peak device bytes
// the context's batches held, extension jobs submitted, rescue jobs submitted.
int bwgw_serve_run(bwg_ctx *ctx, bwg_server *srv, double rate_jobs_s, int producers, double duration_s,
                   uint64_t seed, int use_ref, int pool_chunks, int verify, double *out) {
    const uint64_t genome_seed = seed;
    bwg_ext_params ep{};
    bwg_ksw_params kp{};
    int8_t mat[25];
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) mat[i * 5 + j] = (i == 4 || j == 4) ? -1 : (i == j ? 1 : -4);
    ep.m = kp.m = 5;
    std::memcpy(ep.mat, mat, 25);
    std::memcpy(kp.mat, mat, 25);
    ep.o_del = ep.o_ins = kp.o_del = kp.o_ins = 6;
    ep.e_del = ep.e_ins = kp.e_del = kp.e_ins = 1;
    ep.w = 100;
    ep.end_bonus = 5;
    ep.zdrop = 100;
    kp.minsc = 19;
    kp.max_sc = 1;
    if (use_ref) {
        std::vector<uint8_t> g(bwgw_genome_len());
        bwgw_genome(genome_seed, g.data());
        int rc = bwg_server_set_reference(srv, g.data(), g.size());
        if (rc) return rc;
    }
    // ---- chunk pool ----
    std::mt19937_64 rng(seed * 7919 + 17);
    std::vector<PoolChunk> pool((size_t)pool_chunks);
    double total_jobs = 0;
    for (int c = 0; c < pool_chunks; ++c) {
        PoolChunk &pc = pool[(size_t)c];
        const double u = std::uniform_real_distribution<double>(std::log(1000.0), std::log(65536.0))(rng);
        pc.n = (size_t)std::llround(std::exp(u));
        pc.kind = std::uniform_real_distribution<double>(0, 1)(rng) < 0.8 ? BWG_CHUNK_EXTEND : BWG_CHUNK_RESCUE;
        void *h = bwgw_generate_on(pc.kind == BWG_CHUNK_EXTEND ? 3 : 2, genome_seed, seed * 1000003 + (uint64_t)c, pc.n);
        std::vector<uint8_t> full(bwgw_arena_len(h) + 1);
        pc.q_off.resize(pc.n);
        pc.q_len.resize(pc.n);
        pc.t_off.resize(pc.n);
        pc.t_len.resize(pc.n);
        pc.h0.resize(pc.n);
        bwgw_copy(h, full.data(), pc.q_off.data(), pc.q_len.data(), pc.t_off.data(), pc.t_len.data(), pc.h0.data(), nullptr);
        std::vector<uint64_t> win(pc.n);
        bwgw_copy_windows(h, win.data());
        bwgw_free(h);
        if (verify) {
            pc.alone.resize(pc.n * (pc.kind == BWG_CHUNK_EXTEND ? sizeof(bwg_ext_result) : sizeof(bwg_ksw_result)));
            const int rc = pc.kind == BWG_CHUNK_EXTEND
                               ? bwg_extend_batch(ctx, &ep, pc.n, full.data(), full.size(), pc.q_off.data(), pc.q_len.data(),
                                                  pc.t_off.data(), pc.t_len.data(), pc.h0.data(), (bwg_ext_result *)pc.alone.data())
                               : bwg_ksw_align2_batch(ctx, &kp, pc.n, full.data(), full.size(), pc.q_off.data(),
                                                      pc.q_len.data(), pc.t_off.data(), pc.t_len.data(),
                                                      (bwg_ksw_result *)pc.alone.data());
            if (rc) return rc;
        }
        if (use_ref) {
            // queries only in the arena; targets are genome windows
            pc.arena.clear();
            for (size_t k = 0; k < pc.n; ++k) {
                const uint64_t o = pc.arena.size();
                pc.arena.insert(pc.arena.end(), full.begin() + (long)pc.q_off[k], full.begin() + (long)(pc.q_off[k] + pc.q_len[k]));
                pc.q_off[k] = o;
                pc.t_off[k] = win[k];
            }
            if (pc.arena.empty()) pc.arena.push_back(0);
        } else {
            pc.arena.swap(full);
        }
        total_jobs += (double)pc.n;
    }
    const double mean_chunk = total_jobs / pool_chunks;
    bwg_server_counters st;
    bwg_server_stats(srv, &st, 1);
    // ---- producers and waiters ----
    std::atomic<int> mismatches{0}, failures{0};
    std::atomic<uint64_t> ext_jobs{0}, resc_jobs{0};
    std::atomic<bool> stop_submit{false};
    const auto t0 = Clock::now();
    const auto t_end = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(duration_s));
    std::vector<std::thread> th;
    for (int p = 0; p < producers; ++p) {
        th.emplace_back([&, p] {
            std::mt19937_64 r(seed * 104729 + (uint64_t)p);
            std::exponential_distribution<double> gap(rate_jobs_s / mean_chunk / producers);
            std::mutex mu;
            std::condition_variable cv;
            std::deque<Pending *> q;
            bool done = false;
            std::thread waiter([&] {
                for (;;) {
                    Pending *x;
                    {
                        std::unique_lock<std::mutex> g(mu);
                        cv.wait(g, [&] { return !q.empty() || done; });
                        if (q.empty()) return;
                        x = q.front();
                        q.pop_front();
                    }
                    const int rc = bwg_server_wait(srv, x->ticket);
                    if (rc) failures++;
                    else if (verify && std::memcmp(x->out.data(), x->c->alone.data(), x->out.size()) != 0) mismatches++;
                    delete x;
                }
            });
            auto next = t0;
            size_t i = (size_t)p;
            for (;;) {
                next += std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(gap(r)));
                if (next >= t_end) break;
                std::this_thread::sleep_until(next);
                const PoolChunk &pc = pool[i % pool.size()];
                i += (size_t)producers;
                Pending *x = new Pending;
                x->c = &pc;
                x->out.resize(pc.n * (pc.kind == BWG_CHUNK_EXTEND ? sizeof(bwg_ext_result) : sizeof(bwg_ksw_result)));
                bwg_chunk ch{};
                ch.kind = pc.kind;
                ch.ext = &ep;
                ch.ksw = &kp;
                ch.n = pc.n;
                ch.arena = pc.arena.data();
                ch.arena_len = pc.arena.size();
                ch.q_off = pc.q_off.data();
                ch.q_len = pc.q_len.data();
                ch.t_off = pc.t_off.data();
                ch.t_len = pc.t_len.data();
                ch.h0 = pc.h0.data();
                ch.out = x->out.data();
                if (bwg_server_submit(srv, &ch, &x->ticket) != BWG_OK) {
                    failures++;
                    delete x;
                    continue;
                }
                (pc.kind == BWG_CHUNK_EXTEND ? ext_jobs : resc_jobs) += pc.n;
                {
                    std::lock_guard<std::mutex> g(mu);
                    q.push_back(x);
                }
                cv.notify_one();
            }
            {
                std::lock_guard<std::mutex> g(mu);
                done = true;
            }
            cv.notify_one();
            waiter.join();
        });
    }
    for (auto &t : th) t.join();
    const double wall = std::chrono::duration<double>(Clock::now() - t0).count();
    bwg_server_stats(srv, &st, 0);
    out[0] = st.p50_ms;
    out[1] = st.p99_ms;
    out[2] = st.mean_ms;
    out[3] = st.max_ms;
    out[4] = (double)st.jobs / wall;
    out[5] = (double)st.chunks;
    out[6] = (double)st.jobs;
    out[7] = (double)st.epochs;
    out[8] = (double)mismatches.load();
    out[9] = (double)failures.load();
    out[10] = (double)st.peak_device_bytes;
    out[11] = (double)st.peak_pool_bytes;
    out[12] = (double)ext_jobs.load();
    out[13] = (double)resc_jobs.load();
    (void)stop_submit;
    return BWG_OK;
}

}  // extern "C"
