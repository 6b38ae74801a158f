// Thread-per-lane execution of emulated kernel launches (see
// tests/emu/include/hip/hip_runtime.h).  Blocks run one after another on a
// pool of worker threads; barriers count live lanes, so lanes that have
// finished the kernel no longer hold others back.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace emu {

thread_local Idx tl_thread, tl_block;
dim3 g_block_dim, g_grid_dim;

namespace {

// Lanes parked at the workgroup barrier count as present for their wave's
// barriers: on hardware they are masked off while the rest of the wave
// runs on, so they never hold the wave back.
struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    int live = 0;
    int arrived = 0;
    int parked = 0;
    unsigned long long gen = 0;

    void reset(int n) {
        std::lock_guard<std::mutex> g(m);
        live = n;
        arrived = 0;
        parked = 0;
    }
    void release_if_complete() {  // m held
        if (arrived > 0 && arrived + parked == live) {
            arrived = 0;
            ++gen;
            cv.notify_all();
        }
    }
    template <typename OnRelease>
    void arrive(OnRelease on_release) {
        std::unique_lock<std::mutex> g(m);
        const unsigned long long my = gen;
        ++arrived;
        if (arrived + parked == live) on_release();
        release_if_complete();
        if (gen != my) return;
        cv.wait(g, [&] { return gen != my; });
    }
    void arrive() {
        arrive([] {});
    }
    void leave() {
        std::lock_guard<std::mutex> g(m);
        --live;
        release_if_complete();
    }
    void park(int d) {
        std::lock_guard<std::mutex> g(m);
        parked += d;
        release_if_complete();
    }
};

constexpr int kMaxThreads = 1024;
Barrier g_block_bar;
Barrier g_wave_bar[kMaxThreads / 64];
std::uint64_t g_slots[kMaxThreads];

// The last lane to reach the workgroup barrier un-parks every wave before
// releasing it (all live lanes are at the barrier at that moment).
static void unpark_all() {
    for (auto& w : g_wave_bar) {
        std::lock_guard<std::mutex> g(w.m);
        w.parked = 0;
    }
}

struct Pool {
    std::vector<std::thread> workers;
    std::mutex m;
    std::condition_variable cv_start, cv_done;
    const std::function<void()>* body = nullptr;
    unsigned block = 0;
    int active = 0;       // threads taking part in the current block
    int remaining = 0;    // threads still running it
    unsigned long long round = 0;
    bool stop = false;

    void ensure(int n) {
        while (static_cast<int>(workers.size()) < n) {
            const int id = static_cast<int>(workers.size());
            workers.emplace_back([this, id] { loop(id); });
        }
    }
    void loop(int id) {
        unsigned long long seen = 0;
        for (;;) {
            const std::function<void()>* fn;
            unsigned blk;
            {
                std::unique_lock<std::mutex> g(m);
                cv_start.wait(g, [&] { return stop || (round != seen && id < active); });
                if (stop) return;
                seen = round;
                fn = body;
                blk = block;
            }
            tl_thread = Idx{static_cast<unsigned>(id), 0, 0};
            tl_block = Idx{blk, 0, 0};
            (*fn)();
            g_wave_bar[id / 64].leave();
            {
                std::lock_guard<std::mutex> g(g_block_bar.m);
                --g_block_bar.live;
                if (g_block_bar.arrived > 0 && g_block_bar.arrived + g_block_bar.parked == g_block_bar.live) {
                    unpark_all();
                    g_block_bar.release_if_complete();
                }
            }
            {
                std::lock_guard<std::mutex> g(m);
                if (--remaining == 0) cv_done.notify_all();
            }
        }
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(m);
            stop = true;
        }
        cv_start.notify_all();
        for (auto& t : workers) t.join();
    }
};

Pool& pool() {
    static Pool p;
    return p;
}

}  // namespace

std::vector<Work>*& capture_target() {
    thread_local std::vector<Work>* target = nullptr;
    return target;
}

// ---- streams
namespace {

struct Item {
    enum Kind { kWork, kRecord, kWaitEvent } kind;
    Work fn;
    Event ev;                                     // kWaitEvent: the record waited for
    unsigned long long seq = 0;                   // place in the device's order of queued items
    const char* label = nullptr;
};

struct Stream {
    std::deque<Item> q;
    unsigned long long done_seq = 0;  // seq of the last item run (stream order)
    bool blocking = false;            // the legacy and per-thread default streams
};

std::recursive_mutex g_dev;  // queues, and the device while it runs an item
// every stream ever created (never destroyed: queued work stays reachable)
std::vector<Stream*>& g_streams = *new std::vector<Stream*>();
unsigned long long g_seq = 0;
Stream* g_legacy = nullptr;

Stream* make_stream(bool blocking) {
    Stream* s = new Stream();
    s->blocking = blocking;
    g_streams.push_back(s);
    return s;
}

Stream* resolve(void* h) {  // g_dev held
    if (h == nullptr) {
        if (!g_legacy) g_legacy = make_stream(true);
        return g_legacy;
    }
    if (h == reinterpret_cast<void*>(2)) {  // hipStreamPerThread
        thread_local Stream* mine = nullptr;
        if (!mine) mine = make_stream(true);
        return mine;
    }
    return static_cast<Stream*>(h);
}

bool done(const Event& e) {
    return e.stream == nullptr || static_cast<Stream*>(e.stream)->done_seq >= e.seq;
}

bool runnable(const Item& it) {
    switch (it.kind) {
        case Item::kWaitEvent: return done(it.ev);
        default: return true;
    }
}

// Run one item: the head of `prefer` when it may run (a query or a
// synchronisation of one stream completes that stream first and leaves the
// others queued, as a fast stream would on the device), else the runnable
// stream head queued last.
bool step(Stream* prefer = nullptr) {
    Stream* best = nullptr;
    if (prefer && !prefer->q.empty() && runnable(prefer->q.front())) best = prefer;
    for (Stream* s : g_streams) {
        if (best == prefer && best) break;
        if (!s->q.empty() && runnable(s->q.front()) && (!best || s->q.front().seq > best->q.front().seq)) best = s;
    }
    if (!best) return false;
    Item it = std::move(best->q.front());
    best->q.pop_front();
    static const bool trace = std::getenv("EMU_TRACE") != nullptr;
    if (trace) {
        int sid = 0;
        while (g_streams[sid] != best) ++sid;
        static const char* kinds[] = {"work", "record", "wait-event"};
        std::fprintf(stderr, "[emu] stream %d item %llu %s %s\n", sid, it.seq, kinds[it.kind],
                     it.label ? it.label : "");
    }
    if (it.kind == Item::kWork) it.fn();
    best->done_seq = it.seq;
    return true;
}

template <typename Pred>
void drain_until(Pred pred, Stream* prefer = nullptr) {
    std::lock_guard<std::recursive_mutex> g(g_dev);
    while (!pred()) {
        if (!step(prefer)) {
            std::fprintf(stderr, "emulated streams: a wait can never be satisfied\n");
            std::abort();
        }
    }
}

void enqueue(void* h, Item it) {
    std::lock_guard<std::recursive_mutex> g(g_dev);
    Stream* s = resolve(h);
    it.seq = ++g_seq;
    s->q.push_back(std::move(it));
}

}  // namespace

void* new_stream() {
    std::lock_guard<std::recursive_mutex> g(g_dev);
    return make_stream(false);
}

void submit_on(void* h, Work w, const char* label) {
    if (std::vector<Work>* r = capture_target()) {
        r->push_back(std::move(w));
        return;
    }
    Item it{Item::kWork, std::move(w)};
    it.label = label;
    enqueue(h, std::move(it));
}

void record_on(void* h, Event* ev) {
    if (capture_target()) return;  // a graph replays its work in the recorded order
    std::lock_guard<std::recursive_mutex> g(g_dev);
    Stream* s = resolve(h);
    Item it{Item::kRecord, nullptr};
    it.seq = ++g_seq;
    ev->stream = s;
    ev->seq = it.seq;
    s->q.push_back(std::move(it));
}

void wait_event_on(void* h, const Event* ev) {
    if (capture_target()) return;
    Item it{Item::kWaitEvent, nullptr};
    {
        std::lock_guard<std::recursive_mutex> g(g_dev);
        it.ev = *ev;  // the record current at the call
    }
    enqueue(h, std::move(it));
}

void drain_stream(void* h) {
    Stream* s;
    {
        std::lock_guard<std::recursive_mutex> g(g_dev);
        s = resolve(h);
    }
    drain_until([s] { return s->q.empty(); }, s);
}

void drain_event(const Event* ev) {
    const Event e = *ev;
    drain_until([e] { return done(e); }, static_cast<Stream*>(e.stream));
}

bool event_done(const Event* ev) {
    std::lock_guard<std::recursive_mutex> g(g_dev);
    return done(*ev);
}

void drain_all() {
    drain_until([] {
        for (Stream* s : g_streams)
            if (!s->q.empty()) return false;
        return true;
    });
}

void drain_blocking() {
    drain_until([] {
        for (Stream* s : g_streams)
            if (s->blocking && !s->q.empty()) return false;
        return true;
    });
}

int lane() { return static_cast<int>(tl_thread.x % 64); }
void block_barrier() {
    g_wave_bar[tl_thread.x / 64].park(1);
    g_block_bar.arrive(unpark_all);
}
// EMU_SLOW_WAVE=w:us makes wave w of workgroup 0 of every launch sleep us
// microseconds at each wave barrier, so that the other waves of the
// workgroup run ahead of it: code that lets a wave reuse LDS another wave
// may still be using without a workgroup barrier then fails.
struct SlowWave {
    int wave = -1;
    int us = 0;
};
static SlowWave slow_wave() {
    SlowWave s;
    if (const char* e = std::getenv("EMU_SLOW_WAVE")) std::sscanf(e, "%d:%d", &s.wave, &s.us);
    return s;
}
void wave_barrier() {
    static const SlowWave slow = slow_wave();
    if (slow.wave >= 0 && tl_block.x == 0 && static_cast<int>(tl_thread.x / 64) == slow.wave)
        std::this_thread::sleep_for(std::chrono::microseconds(slow.us));
    g_wave_bar[tl_thread.x / 64].arrive();
}
void exchange_write(std::uint64_t v) { g_slots[tl_thread.x] = v; }
std::uint64_t exchange_read(int l) { return g_slots[(tl_thread.x / 64) * 64 + static_cast<unsigned>(l)]; }

void launch(dim3 grid, dim3 block, const std::function<void()>& body) {
    // Kernels run one at a time on the emulated device: launches from several
    // host threads (the server's lanes) queue here.
    static std::mutex device;
    std::lock_guard<std::mutex> one_kernel(device);
    const int n = static_cast<int>(block.x);
    if (n < 1 || n > kMaxThreads || block.y != 1 || block.z != 1 || grid.y != 1 || grid.z != 1) std::abort();
    Pool& p = pool();
    p.ensure(n);
    g_block_dim = block;
    g_grid_dim = grid;
    for (unsigned b = 0; b < grid.x; ++b) {
        g_block_bar.reset(n);
        for (int w = 0; w < (n + 63) / 64; ++w) g_wave_bar[w].reset(std::min(64, n - 64 * w));
        {
            std::lock_guard<std::mutex> g(p.m);
            p.body = &body;
            p.block = b;
            p.active = n;
            p.remaining = n;
            ++p.round;
        }
        p.cv_start.notify_all();
        std::unique_lock<std::mutex> g(p.m);
        p.cv_done.wait(g, [&] { return p.remaining == 0; });
    }
}

}  // namespace emu

hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
    std::memset(p, 0, sizeof(*p));
    std::strcpy(p->gcnArchName, "gfx950:emulated");
    p->multiProcessorCount = 8;
    p->sharedMemPerBlock = 163840;
    p->clockRate = 2400000;
    p->totalGlobalMem = std::size_t(1) << 34;
    return hipSuccess;
}
