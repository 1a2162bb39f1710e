// emu_runtime.cpp -- host runtime behind tests/emu/hip/hip_runtime.h (see there).
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

thread_local emu_uint3 threadIdx, blockIdx;
thread_local dim3 blockDim, gridDim;

namespace {

// Barrier over the live threads of a group; a thread that leaves the kernel drops out.
// A wait longer than the limit means a wavefront operation outside uniform control flow
// (or a kernel that does not finish): report it and abort instead of hanging the test.
struct Group {
    std::mutex m;
    std::condition_variable cv;
    int live = 0, arrived = 0;
    uint64_t gen = 0;
    void sync() {
        std::unique_lock<std::mutex> l(m);
        const uint64_t g = gen;
        if (++arrived == live) {
            arrived = 0;
            ++gen;
            cv.notify_all();
            return;
        }
        if (!cv.wait_for(l, std::chrono::seconds(60), [&] { return gen != g; })) {
            std::fprintf(stderr, "emu: barrier timeout (divergent wavefront op or stuck kernel)\n");
            std::abort();
        }
    }
    void drop() {
        std::lock_guard<std::mutex> l(m);
        --live;
        if (live > 0 && arrived == live) {
            arrived = 0;
            ++gen;
            cv.notify_all();
        }
    }
};

struct Wave {
    Group g;
    uint64_t slot[64] = {};
    bool alive[64] = {};
};

thread_local Wave* t_wave = nullptr;
thread_local int t_lane = 0;
thread_local Group* t_block = nullptr;

struct Range {
    uintptr_t lo, hi;
};
std::vector<Range> g_ranges;
std::atomic<long long> g_violations{0};
std::atomic<uintptr_t> g_first_bad{0};
std::atomic<long long> g_accesses{0};
bool g_checking = false;

}  // namespace

hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "no error"; }
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t) {
    std::memcpy(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
    *v = 2;  // compute units of the emulated device
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** dev_ptr, const void* symbol) {
    *dev_ptr = const_cast<void*>(symbol);  // host memory is the emulated device memory
    return hipSuccess;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t) {
    *n = 1;
    return hipSuccess;
}

namespace emu {

void launch(dim3 grid, dim3 block, const std::function<void()>& body) {
    const unsigned nt = block.x * block.y * block.z;
    for (unsigned bz = 0; bz < grid.z; ++bz)
        for (unsigned by = 0; by < grid.y; ++by)
            for (unsigned bx = 0; bx < grid.x; ++bx) {
                Group blk;
                blk.live = (int)nt;
                const unsigned nw = (nt + 63) / 64;
                std::vector<std::unique_ptr<Wave>> waves;
                for (unsigned w = 0; w < nw; ++w) {
                    waves.emplace_back(new Wave());
                    const unsigned n = nt - w * 64 < 64 ? nt - w * 64 : 64;
                    waves.back()->g.live = (int)n;
                    for (unsigned l = 0; l < n; ++l) waves.back()->alive[l] = true;
                }
                std::vector<std::thread> ts;
                ts.reserve(nt);
                for (unsigned t = 0; t < nt; ++t)
                    ts.emplace_back([&, t] {
                        threadIdx = {t % block.x, (t / block.x) % block.y, t / (block.x * block.y)};
                        blockIdx = {bx, by, bz};
                        blockDim = block;
                        gridDim = grid;
                        t_wave = waves[t / 64].get();
                        t_lane = (int)(t % 64);
                        t_block = &blk;
                        body();
                        {
                            std::lock_guard<std::mutex> l(t_wave->g.m);
                            t_wave->alive[t_lane] = false;
                        }
                        t_wave->g.drop();
                        t_block->drop();
                    });
                for (auto& th : ts) th.join();
            }
}

uint64_t exchange(uint64_t v, int mode, int arg) {
    Wave* w = t_wave;
    w->slot[t_lane] = v;
    w->g.sync();
    uint64_t r = v;
    if (mode == 0) {
        if (t_lane + arg < 64) r = w->slot[t_lane + arg];
    } else if (mode == 1) {
        r = w->slot[(t_lane ^ arg) & 63];
    } else if (mode == 2) {
        r = w->slot[arg & 63];
    } else if (mode == 4) {
        if (t_lane - arg >= 0) r = w->slot[t_lane - arg];
    } else {
        for (int l = 0; l < 64; ++l)
            if (w->alive[l]) {
                r = w->slot[l];
                break;
            }
    }
    w->g.sync();
    return r;
}

void syncthreads() { t_block->sync(); }

void access(const void* p, size_t bytes) {
    if (!g_checking) return;
    g_accesses.fetch_add(1, std::memory_order_relaxed);
    const uintptr_t a = (uintptr_t)p;
    for (const Range& r : g_ranges)
        if (a >= r.lo && a + bytes <= r.hi) return;
    if (g_violations.fetch_add(1) == 0) g_first_bad = a;
}

double rcp(double x) { return 1.0 / x; }
double rsq(double x) { return 1.0 / std::sqrt(x); }

}  // namespace emu

extern "C" {
// Accesses through JFA_ACCESS must fall inside one registered [p, p + bytes) range.
void emu_add_range(const void* p, size_t bytes) {
    g_ranges.push_back({(uintptr_t)p, (uintptr_t)p + bytes});
    g_checking = true;
}
void emu_clear_ranges() {
    g_ranges.clear();
    g_checking = false;
    g_violations = 0;
    g_first_bad = 0;
    g_accesses = 0;
}
long long emu_violations() { return g_violations.load(); }
unsigned long long emu_first_violation() { return (unsigned long long)g_first_bad.load(); }
long long emu_accesses() { return g_accesses.load(); }
}
