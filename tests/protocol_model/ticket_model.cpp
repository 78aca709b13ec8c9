// Host-thread model of the ticket/done protocol of the persistent transport
// kernel that this library once had and removed (git history: commit f0715e8,
// k_build_chunks and k_sweep_persistent, and its successor f1de03b with
// wall-clock-bounded waits). Test infrastructure only: nothing in the library
// uses it.
//
// The modelled protocol, restated from that code:
//  * chunk lists: directions are dealt to slots round-robin (d % nslots); each
//    slot's list holds, level by level and within a level by increasing
//    direction, ceil(w[d][k] / kChunk) chunks {d, k, start, count} of its
//    directions' level sets, up to a capacity
//    cap = ceil(N * dirs_per_slot / kChunk) + kmax * dirs_per_slot + 16;
//    the list length is min(chunks, cap);
//  * workgroups are dispatched in index order, at most `resident` at a time,
//    and workgroup i works for slot i % nslots (the XCD the hardware deals it
//    to); a workgroup loops: take ticket t = ticket[slot]++; leave when
//    t >= length or abort is set; wait until done[d] >= lvloff[d][k] (the
//    tasks of direction d in levels before k); solve the chunk; done[d] += count;
//  * a wait that exceeds its bound sets abort and the workgroup leaves.
//
// The model checks, for every run: no wait exceeds its bound (no deadlock),
// every task is solved exactly once, every chunk's wait ended only after all
// tasks of the earlier levels of its direction were solved, and every
// direction's done counter ends at its task count.
//
// Input (stdin, text): D kmax N, then D rows of kmax level-set widths.
// Arguments: nslots workgroups resident seed chunk wait_ms [mistake]
// `mistake` plants a plausible error, for the tests that show the model
// catches it: 1 waits for the chunk's own level set too (target lvloff[k+1]),
// 2 lists the level sets of each odd level before those of the level below.
// Output: one line, "ok chunks=<n> max_spins=<n> dropped=<n> short_directions=<n>"
// (exit 0; short directions are those the host would re-run) or
// "FAIL <reason>" (exit 1).

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Chunk {
    uint32_t d, k, start, count;
};

struct Model {
    int D = 0, kmax = 0;
    int64_t N = 0;
    std::vector<std::vector<uint32_t>> width;   // [d][k]
    std::vector<std::vector<uint32_t>> lvloff;  // [d][k], kmax + 1 entries
    int nslots = 1;
    uint32_t chunk = 16;
    std::vector<std::vector<Chunk>> list;  // per slot
    int64_t dropped = 0;
    int mistake = 0;

    // k_build_chunks: per slot, levels in order, directions in increasing order
    void build_lists() {
        const int64_t dps = (D + nslots - 1) / nslots;
        const int64_t cap = (N * dps + chunk - 1) / chunk + (int64_t)kmax * dps + 16;
        list.assign(nslots, {});
        for (int slot = 0; slot < nslots; slot++) {
            int64_t idx = 0;
            for (int kk = 0; kk < kmax; kk++) {
                int k = kk;
                if (mistake == 2 && (kk ^ 1) < kmax) k = kk ^ 1;  // pairs of levels swapped
                for (int d = 0; d < D; d++) {
                    if (d % nslots != slot) continue;
                    const uint32_t beg = lvloff[d][k], w = width[d][k];
                    const uint32_t pieces = (w + chunk - 1) / chunk;
                    for (uint32_t pc = 0; pc < pieces; pc++, idx++) {
                        const uint32_t st = beg + pc * chunk;
                        const uint32_t cnt = std::min(chunk, beg + w - st);
                        if (idx < cap) list[slot].push_back({(uint32_t)d, (uint32_t)k, st, cnt});
                        else dropped++;
                    }
                }
            }
        }
    }
};

struct Run {
    const Model* m;
    int workgroups, resident;
    uint64_t seed;
    std::chrono::milliseconds wait_bound;
    std::vector<std::atomic<uint32_t>> ticket, done;
    std::vector<std::vector<std::atomic<uint32_t>>> solved_in_level;  // [d][k]
    std::vector<std::vector<std::atomic<uint8_t>>> task_solved;       // [d][position]
    std::atomic<uint32_t> abort{0};
    std::atomic<int> next_wg{0};
    std::atomic<int64_t> max_spins{0};
    std::mutex err_mu;
    std::string err;

    Run(const Model& model, int wg, int res, uint64_t s, int wait_ms)
        : m(&model), workgroups(wg), resident(res), seed(s), wait_bound(wait_ms),
          ticket(model.nslots), done(model.D), solved_in_level(model.D), task_solved(model.D) {
        for (int d = 0; d < m->D; d++) {
            solved_in_level[d] = std::vector<std::atomic<uint32_t>>(m->kmax);
            task_solved[d] = std::vector<std::atomic<uint8_t>>(m->lvloff[d][m->kmax]);
        }
    }

    void fail(const std::string& what) {
        std::lock_guard<std::mutex> lock(err_mu);
        if (err.empty()) err = what;
        abort.store(1);
    }

    static uint64_t next(uint64_t& x) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    }

    // one workgroup of k_sweep_persistent
    void workgroup(int wg, uint64_t& rng) {
        const int slot = wg % m->nslots;
        const auto& lst = m->list[slot];
        const uint32_t nch = (uint32_t)lst.size();
        for (;;) {
            const uint32_t t = ticket[slot].fetch_add(1);
            if (t >= nch || abort.load() != 0) return;
            const Chunk c = lst[t];
            const uint32_t target = m->lvloff[c.d][c.k + (m->mistake == 1 ? 1 : 0)];
            // perturb the interleaving between taking the ticket and waiting
            if (next(rng) % 8 == 0) std::this_thread::yield();
            const auto t0 = std::chrono::steady_clock::now();
            int64_t spins = 0;
            while (done[c.d].load(std::memory_order_acquire) < target) {
                if (abort.load() != 0) return;
                if ((++spins & 255) == 0 && std::chrono::steady_clock::now() - t0 > wait_bound) {
                    fail("wait not satisfied: slot " + std::to_string(slot) + " ticket " + std::to_string(t) +
                         " direction " + std::to_string(c.d) + " level " + std::to_string(c.k) + " target " +
                         std::to_string(target) + " done " + std::to_string(done[c.d].load()));
                    return;
                }
                std::this_thread::yield();
            }
            int64_t prev = max_spins.load();
            while (spins > prev && !max_spins.compare_exchange_weak(prev, spins)) {
            }
            for (uint32_t k = 0; k < c.k; k++)
                if (solved_in_level[c.d][k].load() != m->width[c.d][k]) {
                    fail("chunk of direction " + std::to_string(c.d) + " level " + std::to_string(c.k) +
                         " released before level " + std::to_string(k) + " was complete");
                    return;
                }
            if (c.start + c.count > m->lvloff[c.d][c.k + 1] || c.start < m->lvloff[c.d][c.k]) {
                fail("chunk outside its level set");
                return;
            }
            for (uint32_t i = c.start; i < c.start + c.count; i++)
                if (task_solved[c.d][i].fetch_add(1) != 0) {
                    fail("task solved twice");
                    return;
                }
            if (next(rng) % 4 == 0) std::this_thread::yield();
            solved_in_level[c.d][c.k].fetch_add(c.count);
            done[c.d].fetch_add(c.count, std::memory_order_release);
        }
    }

    // a host thread stands for one resident workgroup slot of the device:
    // workgroups start in index order as earlier ones leave
    void resident_slot(int r) {
        uint64_t rng = seed * 0x9E3779B97F4A7C15ull + (uint64_t)r + 1;
        for (;;) {
            const int wg = next_wg.fetch_add(1);
            if (wg >= workgroups) return;
            workgroup(wg, rng);
        }
    }

    int go() {
        std::vector<std::thread> th;
        for (int r = 0; r < resident; r++) th.emplace_back(&Run::resident_slot, this, r);
        for (auto& t : th) t.join();
        // what each direction's counter must end at: the tasks of its listed
        // chunks where a workgroup served its slot (the rest is left to the
        // host's check and level-synchronous re-run)
        int short_dirs = 0;
        if (err.empty()) {
            std::vector<uint64_t> listed(m->D, 0);
            for (int slot = 0; slot < m->nslots && slot < workgroups; slot++)
                for (const Chunk& c : m->list[slot]) listed[c.d] += c.count;
            for (int d = 0; d < m->D; d++) {
                if (done[d].load() != listed[d]) {
                    err = "direction " + std::to_string(d) + " ends at " + std::to_string(done[d].load()) + " of " +
                          std::to_string(listed[d]) + " listed tasks";
                    break;
                }
                if (listed[d] != m->lvloff[d][m->kmax]) short_dirs++;
            }
        }
        if (!err.empty()) {
            std::printf("FAIL %s\n", err.c_str());
            return 1;
        }
        size_t chunks = 0;
        for (auto& l : m->list) chunks += l.size();
        std::printf("ok chunks=%zu max_spins=%lld dropped=%lld short_directions=%d\n", chunks,
                    (long long)max_spins.load(), (long long)m->dropped, short_dirs);
        return 0;
    }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 7 && argc != 8) {
        std::fprintf(stderr, "usage: %s nslots workgroups resident seed chunk wait_ms [mistake] < widths\n", argv[0]);
        return 2;
    }
    Model m;
    m.nslots = std::atoi(argv[1]);
    const int workgroups = std::atoi(argv[2]), resident = std::atoi(argv[3]);
    const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
    m.chunk = (uint32_t)std::atoi(argv[5]);
    const int wait_ms = std::atoi(argv[6]);
    if (argc == 8) m.mistake = std::atoi(argv[7]);
    long long n = 0;
    if (std::scanf("%d %d %lld", &m.D, &m.kmax, &n) != 3 || m.D < 1 || m.kmax < 1 || m.nslots < 1 || m.chunk < 1 ||
        workgroups < 1 || resident < 1) {
        std::printf("FAIL bad input\n");
        return 2;
    }
    m.N = n;
    m.width.assign(m.D, std::vector<uint32_t>(m.kmax));
    m.lvloff.assign(m.D, std::vector<uint32_t>(m.kmax + 1, 0));
    for (int d = 0; d < m.D; d++)
        for (int k = 0; k < m.kmax; k++) {
            unsigned w = 0;
            if (std::scanf("%u", &w) != 1) {
                std::printf("FAIL short input\n");
                return 2;
            }
            m.width[d][k] = w;
            m.lvloff[d][k + 1] = m.lvloff[d][k] + w;
        }
    m.build_lists();
    Run run(m, workgroups, resident, seed, wait_ms);
    return run.go();
}
