// Deterministic synthetic job generators for the extension and mate-rescue workloads.
//
// These build job batches of the shapes bwa-mem4's aligner hands its Smith-Waterman kernels:
//  * extension jobs as built in crates/bwa-mem/src/across.rs:1123-1186 (a query segment of a read
//    past a seed end and the reference window after it; left extensions have both reversed), with
//    a target window of qlen+100 bases and a seed score h0;
//  * rescue jobs as built in crates/bwa-mem/src/pe.rs:596-679 (a 150 bp mate against a reference
//    window predicted from the insert-size distribution).
// Reads are simulated from a random ACGT reference with substitutions and short indels.  Every job
// draws from its own splitmix64 stream keyed by (seed, job index), so generation is reproducible,
// parallel and independent of thread count.
//
// Layout handed back: one byte arena (query bytes then target bytes per job, codes 0..4) plus
// per-job offsets/lengths and h0, exactly the job description the C ABI in include/bwg.h takes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {  // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    int64_t uniform(int64_t lo, int64_t hi) {  // inclusive
        return lo + (int64_t)(next() % (uint64_t)(hi - lo + 1));
    }
    double normal(double mu, double sd) {  // Box-Muller
        double u1 = unit(), u2 = unit();
        if (u1 < 1e-300) u1 = 1e-300;
        return mu + sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

uint64_t mix(uint64_t a, uint64_t b) {
    Rng r(a * 0x2545F4914F6CDD1Dull ^ (b + 0x632BE59BD9B4E019ull));
    r.next();
    return r.next();
}

struct Genome {
    std::vector<uint8_t> seq;
    Genome(uint64_t seed, size_t len, double n_rate) : seq(len) {
        const size_t block = 1 << 16;
        size_t nb = (len + block - 1) / block;
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                for (size_t b = t; b < nb; b += nt) {
                    Rng r(mix(seed ^ 0x5eedull, b));
                    size_t hi = std::min(len, (b + 1) * block);
                    for (size_t k = b * block; k < hi; ++k)
                        seq[k] = r.unit() < n_rate ? 4 : (uint8_t)(r.next() & 3);
                }
            });
        for (auto &x : th) x.join();
    }
};

// Simulate a read of `len` bases starting at reference position `pos`: each step substitutes
// with probability sub, or starts an indel (insertion/deletion equally likely, length 1..3) with
// probability indel.  ref_of[k] is the reference coordinate the k-th read base sits against (the
// next reference base for inserted bases).
void sim_read(Rng &r, const Genome &g, size_t pos, int len, double sub, double indel,
              std::vector<uint8_t> &read, std::vector<int64_t> &ref_of) {
    read.clear();
    ref_of.clear();
    size_t p = pos;
    while ((int)read.size() < len && p < g.seq.size()) {
        double u = r.unit();
        if (u < indel) {
            int l = (int)r.uniform(1, 3);
            if (r.next() & 1) {  // insertion into the read
                for (int k = 0; k < l && (int)read.size() < len; ++k) {
                    read.push_back((uint8_t)(r.next() & 3));
                    ref_of.push_back((int64_t)p);
                }
            } else {
                p += (size_t)l;  // deletion from the read
            }
            continue;
        }
        uint8_t b = g.seq[p];
        if (r.unit() < sub && b < 4) b = (uint8_t)((b + 1 + (r.next() % 3)) & 3);
        read.push_back(b);
        ref_of.push_back((int64_t)p);
        ++p;
    }
    while ((int)read.size() < len) {  // ran off the genome end (never for the sizes used)
        read.push_back((uint8_t)(r.next() & 3));
        ref_of.push_back((int64_t)g.seq.size() - 1);
    }
}

struct Job {
    std::vector<uint8_t> q, t;
    int32_t h0 = 0;
    uint64_t win = 0;   // target as a genome window: start position
    bool rev = false;   // ... read back to front
};

// One extension job: query = read[s, s+qlen) (right extension) or the reverse of read[s-qlen, s)
// (left extension), target = the reference window of qlen+100 bases on the same side of the seed
// boundary, reversed for a left extension.
void make_ext_job(Rng &r, const Genome &g, int read_len, int qlen, double sub, double indel,
                  int h0_lo, int h0_hi, Job &job) {
    const int tlen = qlen + 100;
    std::vector<uint8_t> read;
    std::vector<int64_t> ref_of;
    size_t margin = 2 * (size_t)read_len + 2 * (size_t)tlen + 64;
    size_t pos = margin + (size_t)(r.next() % (g.seq.size() - 2 * margin));
    sim_read(r, g, pos, read_len, sub, indel, read, ref_of);
    bool left = (r.next() & 1) != 0;
    job.q.resize((size_t)qlen);
    job.t.resize((size_t)tlen);
    if (!left) {
        int s = (int)r.uniform(0, read_len - qlen);
        int64_t rp = ref_of[(size_t)s];
        for (int k = 0; k < qlen; ++k) job.q[(size_t)k] = read[(size_t)(s + k)];
        for (int k = 0; k < tlen; ++k) job.t[(size_t)k] = g.seq[(size_t)(rp + k)];
        job.win = (uint64_t)rp;
        job.rev = false;
    } else {
        int s = (int)r.uniform(qlen, read_len);  // seed starts at read position s
        int64_t rp = s < read_len ? ref_of[(size_t)s] : ref_of[(size_t)read_len - 1] + 1;
        for (int k = 0; k < qlen; ++k) job.q[(size_t)k] = read[(size_t)(s - 1 - k)];
        for (int k = 0; k < tlen; ++k) job.t[(size_t)k] = g.seq[(size_t)(rp - 1 - k)];
        job.win = (uint64_t)(rp - tlen);
        job.rev = true;
    }
    job.h0 = (int32_t)r.uniform(h0_lo, h0_hi);
}

// One rescue job: a 150 bp mate against a window of U[600,1200] bases.  70% of windows hold the
// mate at an offset drawn from the insert-size distribution N(550,150) (clamped into the window);
// 30% are an unrelated stretch of the reference.
void make_resc_job(Rng &r, const Genome &g, Job &job) {
    const int read_len = 150;
    std::vector<uint8_t> read;
    std::vector<int64_t> ref_of;
    int L = (int)r.uniform(600, 1200);
    size_t margin = 4096;
    size_t pos = margin + (size_t)(r.next() % (g.seq.size() - 2 * margin));
    sim_read(r, g, pos, read_len, 0.01, 0.001, read, ref_of);
    job.q = read;
    job.t.resize((size_t)L);
    size_t ws;
    if (r.unit() < 0.30) {
        ws = margin + (size_t)(r.next() % (g.seq.size() - 2 * margin));
    } else {
        double ins = r.normal(550.0, 150.0);
        int64_t off = (int64_t)std::llround(ins) - read_len;
        off = std::max<int64_t>(0, std::min<int64_t>(off, L - read_len));
        ws = pos - (size_t)off;
    }
    for (int k = 0; k < L; ++k) job.t[(size_t)k] = g.seq[ws + (size_t)k];
    job.h0 = 0;
    job.win = ws;
    job.rev = false;
}

int draw_tail_qlen(Rng &r) {  // 70% < 30, 25% 30..150, 5% 150..250
    double u = r.unit();
    if (u < 0.70) return (int)r.uniform(1, 29);
    if (u < 0.95) return (int)r.uniform(30, 150);
    return (int)r.uniform(150, 250);
}
int draw_read_len(Rng &r, int qlen) {  // 100/150/250 at 40/40/20%, at least qlen
    double u = r.unit();
    int rl = u < 0.4 ? 100 : (u < 0.8 ? 150 : 250);
    if (rl < qlen) rl = qlen <= 150 ? 150 : 250;
    return rl;
}

struct Batch {
    std::vector<uint8_t> arena;
    std::vector<uint64_t> t_win;  // target as a genome window: start | 1<<62 (| 1<<63 reversed)
    std::vector<uint64_t> q_off, t_off;
    std::vector<uint32_t> q_len, t_len;
    std::vector<int32_t> h0;
    std::vector<uint8_t> kind;  // 0 extension, 1 rescue (mixed batches only)
};

// kind: 1 = plain extension (U[1,130] from 150 bp reads), 2 = rescue, 3 = long-tail extension,
// 4 = serving mix (80% long-tail extension / 20% rescue).
Batch *generate(int kind, uint64_t genome_seed, uint64_t seed, size_t n) {
    static const size_t kGenome = 8u << 20;
    Genome g(genome_seed, kGenome, 0.001);
    std::vector<Job> jobs(n);
    std::vector<uint8_t> kinds(n, 0);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (size_t k = t; k < n; k += nt) {
                Rng r(mix(seed, k));
                switch (kind) {
                case 1:
                    make_ext_job(r, g, 150, (int)r.uniform(1, 130), 0.01, 0.001, 19, 150, jobs[k]);
                    break;
                case 2:
                    make_resc_job(r, g, jobs[k]);
                    kinds[k] = 1;
                    break;
                case 3: {
                    int q = draw_tail_qlen(r);
                    make_ext_job(r, g, draw_read_len(r, q), q, 0.01, 0.002, 19, 250, jobs[k]);
                    break;
                }
                default:
                    if (r.unit() < 0.8) {
                        int q = draw_tail_qlen(r);
                        make_ext_job(r, g, draw_read_len(r, q), q, 0.01, 0.002, 19, 250, jobs[k]);
                    } else {
                        make_resc_job(r, g, jobs[k]);
                        kinds[k] = 1;
                    }
                }
            }
        });
    for (auto &x : th) x.join();
    Batch *b = new Batch;
    size_t total = 0;
    for (auto &j : jobs) total += j.q.size() + j.t.size();
    b->arena.resize(total);
    b->q_off.resize(n);
    b->t_off.resize(n);
    b->q_len.resize(n);
    b->t_len.resize(n);
    b->h0.resize(n);
    b->t_win.resize(n);
    b->kind = kinds;
    size_t o = 0;
    for (size_t k = 0; k < n; ++k) {
        auto &j = jobs[k];
        b->q_off[k] = o;
        b->q_len[k] = (uint32_t)j.q.size();
        if (!j.q.empty()) std::memcpy(&b->arena[o], j.q.data(), j.q.size());
        o += j.q.size();
        b->t_off[k] = o;
        b->t_len[k] = (uint32_t)j.t.size();
        if (!j.t.empty()) std::memcpy(&b->arena[o], j.t.data(), j.t.size());
        o += j.t.size();
        b->h0[k] = j.h0;
        b->t_win[k] = j.win | (1ull << 62) | (j.rev ? (1ull << 63) : 0);
    }
    return b;
}

}  // namespace

extern "C" {

void *bwgw_generate(int kind, uint64_t seed, uint64_t n) { return generate(kind, seed, seed, (size_t)n); }
// Jobs drawn with `seed` from the genome of `genome_seed` (batches that share one genome).
void *bwgw_generate_on(int kind, uint64_t genome_seed, uint64_t seed, uint64_t n) {
    return generate(kind, genome_seed, seed, (size_t)n);
}
uint64_t bwgw_arena_len(const void *h) { return ((const Batch *)h)->arena.size(); }
uint64_t bwgw_count(const void *h) { return ((const Batch *)h)->q_len.size(); }
void bwgw_copy(const void *h, uint8_t *arena, uint64_t *q_off, uint32_t *q_len, uint64_t *t_off,
               uint32_t *t_len, int32_t *h0, uint8_t *kind) {
    const Batch *b = (const Batch *)h;
    size_t n = b->q_len.size();
    if (arena && !b->arena.empty()) std::memcpy(arena, b->arena.data(), b->arena.size());
    if (q_off) std::memcpy(q_off, b->q_off.data(), n * 8);
    if (q_len) std::memcpy(q_len, b->q_len.data(), n * 4);
    if (t_off) std::memcpy(t_off, b->t_off.data(), n * 8);
    if (t_len) std::memcpy(t_len, b->t_len.data(), n * 4);
    if (h0) std::memcpy(h0, b->h0.data(), n * 4);
    if (kind) std::memcpy(kind, b->kind.data(), n);
}
void bwgw_free(void *h) { delete (Batch *)h; }

// Target windows of the jobs of a generated batch, as genome coordinates with the reference flag
// (and the reverse flag for left extensions) set: the form a server with the genome resident takes.
void bwgw_copy_windows(const void *h, uint64_t *t_win) {
    const Batch *b = (const Batch *)h;
    if (t_win) std::memcpy(t_win, b->t_win.data(), b->t_win.size() * 8);
}

// The genome the generator draws every job of `seed` from.
uint64_t bwgw_genome_len() { return 8u << 20; }
void bwgw_genome(uint64_t seed, uint8_t *out) {
    Genome g(seed, 8u << 20, 0.001);
    std::memcpy(out, g.seq.data(), g.seq.size());
}

}  // extern "C"
