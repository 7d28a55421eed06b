// engine.hip - host side of the C ABI (include/pintt_gpu.h).
//
// One device context per process: a stream, the twiddle tables (per prime of
// each prime set, built when the set is first needed), a residue arena, the
// carry maps, descriptor buffers and staging buffers for host operands.
//
// Every call is a batch (a single product is a batch of one):
//   1. each product becomes a group: the longer operand is cut into m chunks
//      when that makes the transforms cheaper (unbalanced operands), all
//      sharing the one transform of the shorter operand;
//   2. groups are bucketed by transform length N;
//   3. per bucket, five launches (forward columns, row convolution, inverse
//      columns, CRT + carry, carry fixup) process every group of the bucket,
//      driven by job descriptors uploaded once per batch.
// Calls are serialised by a mutex; work enqueued on a caller's stream is
// ordered behind the previous batch with an event.
//
// Reference counterparts: IntNttEngine::ensure_buffers / TwiddleCache::build
// (src/gpu/int_ntt.cu:110-186,282-299), which rebuilt the tables whenever the
// length changed, and the per-GPU mutex of the wrappers
// (src/arithmetic/int_ntt_multiplier.cpp:137-144).
#include "pintt_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <mutex>
#include <unordered_map>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "bigint_add.h"
#include "ntt_kernels.h"

static_assert(pintt::kBsLeafTerms == PINTT_BS_LEAF_TERMS && pintt::kBsLeafLimbs == PINTT_BS_LEAF_LIMBS,
              "subtree shape");
static_assert(pintt::kLeafMaxDigits == PINTT_DECIMAL_LEAF_MAX_DIGITS &&
                  pintt::kLeafMaxLimbs == PINTT_DECIMAL_LEAF_MAX_LIMBS,
              "decimal leaf limits");

// Device memory of gpu_ntt_mul_device_noalloc, laid out in 32-bit words (each
// region starts on a 256-byte boundary): forward and inverse twiddle tables
// of the narrow prime set for transforms of up to 2^kNoAllocMaxLog2N points
// (three primes, (w, w') pairs), the row factor tables, the power tables the
// builder uses, a residue arena of two operand slots at the longest transform,
// and the carry maps.  72.1 MiB; each device the library runs on holds it
// until the process ends.
namespace pintt {
namespace {
constexpr int kNoAllocMaxLog2N = PINTT_NOALLOC_MAX_LOG2N;
constexpr size_t kNoAllocN = (size_t)1 << kNoAllocMaxLog2N;
constexpr uint32_t kNoAllocMaxChunks = 3;  // every job list of a product fits the kernel arguments
constexpr size_t words_up(size_t w) { return (w + 63) & ~(size_t)63; }
constexpr size_t kNaTwWords = 2 * kNoAllocN;  // one table of one prime
constexpr size_t kNaRowWords = words_up(row_table_words(kSetNarrow));
constexpr size_t kNaFwd = 0;
constexpr size_t kNaInv = kNaFwd + kPrimes * kNaTwWords;
constexpr size_t kNaRow = kNaInv + kPrimes * kNaTwWords;
constexpr size_t kNaPow = kNaRow + kPrimes * kNaRowWords;
constexpr size_t kNaArena = kNaPow + words_up(kPowScratch);
constexpr size_t kNaArenaWords = 2 * kPrimes * kNoAllocN;  // two slots; m <= 3 chunks of half the length fit too
constexpr size_t kNaMaps = kNaArena + kNaArenaWords;
constexpr size_t kNaMapWords = words_up(kNoAllocN / carry_block_words(kCarryMinThreads) + 1);
constexpr size_t kNaWords = kNaMaps + kNaMapWords;
static_assert(PINTT_NOALLOC_MAX_LIMBS == ((size_t)1 << (kNoAllocMaxLog2N - 1)), "limit in the header");
}  // namespace
}  // namespace pintt

static __device__ __align__(256) uint32_t scratch[pintt::kNaWords];

namespace pintt {
namespace {

thread_local std::string t_error;

struct Engine {
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;
    // Streams are ordered lazily: `last` is the stream that received the most
    // recent engine work and `done` is recorded there after it; work for
    // another stream first waits on `done`.  Calls that stay on one stream
    // cost one event record each and no cross-queue wait.
    hipEvent_t done = nullptr;
    hipStream_t last = nullptr;
    bool tw_ready[kSets] = {};
    uint32_t* tw_pow = nullptr;  // powers of the root while tables are built
    uint2* tw_fwd[kSets][kPrimes] = {};
    uint2* tw_inv[kSets][kPrimes] = {};
    uint32_t* tw_row[kSets][kPrimes] = {};  // row factor tables
    size_t arena_words = 0;
    uint32_t* arena = nullptr;
    size_t maps_cap = 0;
    uint32_t* maps = nullptr;  // carry maps, one per carry workgroup
    // Ring of descriptor buffers, so the host can prepare a few batches ahead:
    // slot k is reused only after its previous upload has completed.
    static constexpr int kDescRing = 8;
    int desc_next = 0;
    size_t desc_cap[kDescRing] = {};               // bytes
    unsigned char* desc_host[kDescRing] = {};      // pinned
    unsigned char* desc_dev[kDescRing] = {};
    hipEvent_t desc_done[kDescRing] = {};
    size_t io_cap = 0;  // limbs
    uint64_t* io_in = nullptr;
    uint64_t* io_out = nullptr;
    // Batched add/subtract: job descriptors (pinned + device), per-block maps.
    size_t add_job_cap = 0;
    AddJob* add_jobs_host = nullptr;
    AddJob* add_jobs_dev = nullptr;
    uint32_t* add_status = nullptr;  // per job: 1 when a subtraction went negative
    size_t add_blk_cap = 0;
    uint8_t* add_maps = nullptr;
    uint8_t* add_carry = nullptr;
    uint32_t* add_error = nullptr;
    // Binary-splitting subtrees: lengths, signs and the error word (device).
    size_t bs_cap = 0;
    uint32_t* bs_len = nullptr;
    int32_t* bs_sign = nullptr;
    uint32_t* bs_err = nullptr;
    // Decimal leaves: job descriptors (pinned + device).
    size_t leaf_job_cap = 0;
    LeafJob* leaf_jobs_host = nullptr;
    LeafJob* leaf_jobs_dev = nullptr;
    bool timing = false;
    bool school = true;  // small products to the schoolbook kernel (off while tuning)
    hipEvent_t ev[5] = {};
    bool have_times = false;
    // Tile plan per transform length: tuned on the device at first use
    // (unless disabled), or set explicitly.
    bool autotune = true;
    bool have_plan[kMaxLog2N + 1] = {};
    Plan plan[kMaxLog2N + 1] = {};
    // Plans for large buckets (kBatchPlanMinPoints points of forward
    // transforms or more), tuned on a tree level; the plan above otherwise.
    bool have_bplan[kMaxLog2N + 1] = {};
    Plan bplan[kMaxLog2N + 1] = {};
    float plan_ms[kMaxLog2N + 1] = {};
    int split_mode = 1;  // 0: never cut operands into chunks, 1: when cheaper, m >= 2: m chunks when valid
    bool graph_branches = true;  // graphs capture each length bucket as a branch of its own
    bool pair_rows = true;  // first two chunks of a group in one row launch (PINTT_PAIR=0: off)
    size_t tune_cap = 0;      // limbs per tuning operand
    size_t tune_out_cap = 0;  // limbs of tuning outputs
    uint64_t* tune_in = nullptr;
    uint64_t* tune_out = nullptr;
    // This device's copy of `scratch` (its device address, from
    // hipGetSymbolAddress) and the tables built in it at initialisation.
    uint32_t* na_base = nullptr;
    Twiddles na_tw{};
    // Caller's costs of chunk counts (gpu_ntt_set_chunk_cost), or null.
    pintt_chunk_cost_fn chunk_cost = nullptr;
    void* chunk_ctx = nullptr;
};

Engine g;
std::mutex g_mu;

// Pinned host staging for host operands: gathered by host threads, one DMA
// each way per call (pageable copies go through the runtime's own small
// staging buffers, one call per operand).  A caller holds a slot from its
// gather to its scatter but the engine lock only while it enqueues device
// work, so one caller's host copies overlap another caller's device work.
// The device-side staging (io_in, io_out) is shared: every use is ordered on
// the library stream.
struct Staging {
    size_t cap = 0;  // limbs
    uint64_t* in = nullptr;
    uint64_t* out = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
};
constexpr int kStagingSlots = 2;
Staging g_stage[kStagingSlots];
std::mutex g_stage_mu;
std::condition_variable g_stage_cv;

int fail(int code, const std::string& what) {
    t_error = what;
    return code;
}

int hip_fail(hipError_t e, const char* where) {
    return fail(e == hipErrorOutOfMemory ? PINTT_ERR_NO_MEMORY : PINTT_ERR_HIP,
                std::string(where) + ": " + hipGetErrorString(e));
}

#define PINTT_HIP(call)                                   \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

#define PINTT_TRY(expr)      \
    do {                     \
        int rc_ = (expr);    \
        if (rc_) return rc_; \
    } while (0)

uint32_t powmod(uint32_t b, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = b % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}
uint32_t invmod(uint32_t a, uint32_t p) { return powmod(a % p, p - 2, p); }
uint32_t mulmod(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)((uint64_t)a * b % p); }
uint2 shoup(uint32_t w, uint32_t p) { return make_uint2(w, (uint32_t)(((uint64_t)w << 32) / p)); }

CrtConst crt_constants(int log2n) {
    const int set = set_for(log2n);
    const uint32_t p0 = kPrimeSet[set][0], p1 = kPrimeSet[set][1], p2 = kPrimeSet[set][2];
    auto scale = [&](uint32_t p) {  // 2^32 / N mod p
        return mulmod((uint32_t)((1ull << 32) % p), invmod((uint32_t)((1ull << log2n) % p), p), p);
    };
    CrtConst c;
    c.k0 = shoup(scale(p0), p0);
    uint32_t ip0_1 = invmod(p0 % p1, p1);
    c.k1 = shoup(mulmod(scale(p1), ip0_1, p1), p1);
    c.c1 = shoup(ip0_1, p1);
    uint32_t ip01_2 = invmod(mulmod(p0 % p2, p1 % p2, p2), p2);
    c.k2 = shoup(mulmod(scale(p2), ip01_2, p2), p2);
    c.c20 = shoup(ip01_2, p2);
    c.c21 = shoup(invmod(p1 % p2, p2), p2);
    const unsigned __int128 m = (unsigned __int128)p0 * p1 * p2;
    const unsigned __int128 negm = ((unsigned __int128)1 << 96) - m;
    const uint32_t ps[kPrimes] = {p0, p1, p2};
    for (int i = 0; i < kPrimes; ++i) {
        const unsigned __int128 mi = m / ps[i];
        c.e[i] = shoup(mulmod(scale(ps[i]), invmod((uint32_t)(mi % ps[i]), ps[i]), ps[i]), ps[i]);
        c.mlo[i] = (uint32_t)mi;
        c.mhi[i] = (uint32_t)(mi >> 32);
        c.inv[i] = (float)(1.0 / (double)ps[i]);
        c.negm[i] = (uint32_t)(negm >> (32 * i));
    }
    return c;
}

// Once per device, with it current: the device address of `scratch` and the
// twiddle tables of gpu_ntt_mul_device_noalloc, built in it and waited for.
int setup_noalloc_scratch() {
    uint32_t* base = nullptr;
    PINTT_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&base), HIP_SYMBOL(scratch)));
    if (!base) return fail(PINTT_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
    Twiddles tw{};
    for (int i = 0; i < kPrimes; ++i) {
        uint2* fwd = reinterpret_cast<uint2*>(base + kNaFwd + i * kNaTwWords);
        uint2* inv = reinterpret_cast<uint2*>(base + kNaInv + i * kNaTwWords);
        uint32_t* row = base + kNaRow + i * kNaRowWords;
        PINTT_HIP(launch_build_twiddles(fwd, inv, row, base + kNaPow, kSetNarrow, i, kNoAllocMaxLog2N, g.stream));
        tw.fwd[i] = fwd;
        tw.inv[i] = inv;
        tw.row[i] = row;
    }
    PINTT_HIP(hipStreamSynchronize(g.stream));
    g.na_base = base;
    g.na_tw = tw;
    return PINTT_OK;
}

int ensure_init(int device) {
    if (g.ready) {
        if (device < 0 || device == g.device) return PINTT_OK;
        return fail(PINTT_ERR_ARGS, "engine already initialised on another device");
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(PINTT_ERR_NO_DEVICE, "no HIP device available");
    int dev = device < 0 ? 0 : device;
    if (dev >= count) return fail(PINTT_ERR_NO_DEVICE, "device index out of range");
    PINTT_HIP(hipSetDevice(dev));
    hipDeviceProp_t prop;
    PINTT_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(PINTT_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                             ", the kernels are built for gfx950 only");
    PINTT_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    PINTT_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    for (auto& e : g.desc_done) PINTT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : g.ev) PINTT_HIP(hipEventCreate(&e));
    const char* at = std::getenv("PINTT_AUTOTUNE");
    g.autotune = !(at && at[0] == '0');
    const char* sp = std::getenv("PINTT_SPLIT");
    g.split_mode = (sp && sp[0] == '0') ? 0 : 1;
    const char* pr = std::getenv("PINTT_PAIR");
    g.pair_rows = !(pr && pr[0] == '0');
    g.device = dev;
    PINTT_TRY(setup_noalloc_scratch());
    g.ready = true;
    return PINTT_OK;
}

// Engine work about to be enqueued on `s`: ordered after all earlier engine
// work, on whichever stream it went.
int begin_on(hipStream_t s) {
    if (g.last && g.last != s) PINTT_HIP(hipStreamWaitEvent(s, g.done, 0));
    return PINTT_OK;
}

// Engine work just enqueued on `s` (ends the ordering chain there).
int end_on(hipStream_t s) {
    PINTT_HIP(hipEventRecord(g.done, s));
    g.last = s;
    return PINTT_OK;
}

// Until all engine work on every stream has completed.
int drain() {
    if (g.last) PINTT_HIP(hipEventSynchronize(g.done));
    PINTT_HIP(hipStreamSynchronize(g.stream));
    return PINTT_OK;
}

// Replace a device allocation by a larger one (contents are not kept).
template <class T>
int grow(T*& ptr, size_t count) {
    if (ptr) {
        PINTT_TRY(drain());
        PINTT_HIP(hipFree(ptr));
        ptr = nullptr;
    }
    PINTT_HIP(hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T)));
    return PINTT_OK;
}

// The tables of a prime set are built once, at the set's longest transform
// (narrow set: 2^22 entries, 192 MiB for the three primes forward and inverse;
// wide set: 2^25 entries, 1.5 GiB), when a length it serves is first seen, so
// that captured graphs never see them move.
int ensure_twiddles(int log2n) {
    const int set = set_for(log2n);
    if (g.tw_ready[set]) return PINTT_OK;
    size_t cnt = (size_t)1 << kSetMaxLog2N[set];
    if (!g.tw_pow) PINTT_HIP(hipMalloc(reinterpret_cast<void**>(&g.tw_pow), kPowScratch * sizeof(uint32_t)));
    for (int i = 0; i < kPrimes; ++i) {
        PINTT_TRY(grow(g.tw_fwd[set][i], cnt));
        PINTT_TRY(grow(g.tw_inv[set][i], cnt));
        PINTT_TRY(grow(g.tw_row[set][i], row_table_words(set)));
    }
    PINTT_TRY(begin_on(g.stream));
    for (int i = 0; i < kPrimes; ++i)
        PINTT_HIP(launch_build_twiddles(g.tw_fwd[set][i], g.tw_inv[set][i], g.tw_row[set][i], g.tw_pow, set, i,
                                        kSetMaxLog2N[set], g.stream));
    PINTT_TRY(end_on(g.stream));  // products on any stream wait for the tables
    g.tw_ready[set] = true;
    return PINTT_OK;
}

int ensure_arena(size_t words) {
    if (words <= g.arena_words) return PINTT_OK;
    PINTT_TRY(grow(g.arena, words));
    g.arena_words = words;
    return PINTT_OK;
}

// Carry maps: every map a fixup reads is written earlier in the same launch
// pair, so the buffer needs no clearing.
int ensure_carry_state(size_t blocks) {
    if (blocks > g.maps_cap) {
        PINTT_TRY(grow(g.maps, blocks));
        g.maps_cap = blocks;
    }
    return PINTT_OK;
}

// Ring slot `k` with room for `bytes`, its previous upload finished.  A slot
// that must grow makes every smaller slot grow with it, with half again as much
// room: the stream drains once, not once per slot as batches get larger.
int ensure_descriptors(int k, size_t bytes) {
    PINTT_HIP(hipEventSynchronize(g.desc_done[k]));
    if (bytes <= g.desc_cap[k]) return PINTT_OK;
    PINTT_TRY(drain());
    const size_t cap = std::max(bytes + bytes / 2, (size_t)1 << 16);
    for (int j = 0; j < Engine::kDescRing; ++j) {
        if (g.desc_cap[j] >= cap) continue;
        if (g.desc_host[j]) PINTT_HIP(hipHostFree(g.desc_host[j]));
        g.desc_host[j] = nullptr;
        g.desc_cap[j] = 0;
        PINTT_HIP(hipHostMalloc(reinterpret_cast<void**>(&g.desc_host[j]), cap, hipHostMallocDefault));
        PINTT_TRY(grow(g.desc_dev[j], cap));
        g.desc_cap[j] = cap;
    }
    return PINTT_OK;
}

int log2_for(size_t an, size_t bn) {
    size_t need = 2 * an + 2 * bn - 1;
    int k = kMinLog2N;
    while (k <= kMaxLog2N && ((size_t)1 << k) < need) ++k;
    return k > kMaxLog2N ? -1 : k;
}

Twiddles twiddles(int set) {
    Twiddles t;
    for (int i = 0; i < kPrimes; ++i) {
        t.fwd[i] = g.tw_fwd[set][i];
        t.inv[i] = g.tw_inv[set][i];
        t.row[i] = g.tw_row[set][i];
    }
    return t;
}

std::vector<Plan> candidate_plans(int log2n) {
    std::vector<Plan> out;
    for (int r = 5; r <= 11; ++r) {
        int c = log2n - r;
        if (c < 5 || c > 14) continue;
        for (int w = 3; w <= 6; ++w)
            for (int rows = 0; rows <= 8; ++rows) {
                Plan pl{log2n, r, c, w, rows};
                if (plan_valid(pl)) out.push_back(pl);
            }
    }
    return out;
}

// ---------------------------------------------------------------------------
// Batch planning.

struct Request {  // one product, device pointers
    const uint64_t* a;
    size_t an;
    const uint64_t* b;
    size_t bn;
    uint64_t* out;
    bool square;
};

struct Group {
    Request r;    // a is the longer operand (swapped if needed)
    int log2n;
    uint32_t m;   // chunks of a
    size_t h;     // limbs per chunk (multiple of 16 when m > 1)
    size_t nblk;  // carry blocks
};

int plan_for(int log2n, Plan* out);

// How a batch may be planned: whether tile plans may be tuned on the device
// first (which allocates and waits), and the most chunks a product may be cut
// into.
struct PlanMode {
    bool tune = true;
    uint32_t max_chunks = kMaxChunks;
    bool throughput = false;  // the batch is large enough to be bound by throughput, not latency
    bool disjoint = false;  // every bucket in its own part of the arena, so buckets may run at once
};

// Buckets of at least this many points of forward transforms (2^24: some
// hundreds of transforms of 2^15 .. 2^17 points) are throughput-bound and take
// a length's batch plan; smaller ones, single products among them, are bound
// by latency and take the plan tuned for the length alone.
constexpr size_t kBatchPlanMinPoints = (size_t)1 << 24;

// The plan of a length without tuning: the tuned or loaded one, else the
// heuristic.
Plan plan_lookup(int log2n) { return g.have_plan[log2n] ? g.plan[log2n] : make_plan(log2n); }

// Estimated cost of `transforms` transforms of length 2^k: a third of the
// tuned time of one full product (three transforms) when known, else N log N.
double transform_cost(int k, int transforms) {
    double base = g.have_plan[k] && g.plan_ms[k] > 0 ? g.plan_ms[k] / 3.0
                                                     : 1e-7 * (double)((size_t)1 << k) * k;
    return base * transforms;
}

// The cut of an an x bn product (an >= bn) into m chunks: chunk limbs *h (a
// multiple of 16 when m > 1, so chunk offsets stay multiples of 32
// coefficients) and transform length 2^*k; false when the rules forbid it (a
// chunk shorter than b, an empty last chunk, no shorter transform).
bool chunk_shape(size_t an, size_t bn, uint32_t m, size_t* h, int* k) {
    const int full = log2_for(an, bn);
    if (full < 0 || m < 1 || m > (uint32_t)kMaxChunks) return false;
    if (m == 1) {
        *h = an;
        *k = full;
        return true;
    }
    size_t c = (an + m - 1) / m;
    c = (c + 15) & ~(size_t)15;
    if (c < bn || (m - 1) * c >= an) return false;  // at most two chunk products per word
    const int kk = log2_for(c, bn);
    if (kk < 0 || kk >= full) return false;
    *h = c;
    *k = kk;
    return true;
}

// Unbalanced products alone (or in small batches) are cut into chunks only
// from this transform length on.
constexpr int kChunkAloneMinLog2N = 21;

int plan_group(const Request& req, Group* gr, const PlanMode& mode = PlanMode{}) {
    Request r = req;
    if (!r.square && r.an < r.bn) {
        std::swap(r.a, r.b);
        std::swap(r.an, r.bn);
    }
    gr->r = r;
    gr->m = 1;
    gr->h = r.an;
    gr->log2n = log2_for(r.an, r.bn);
    if (gr->log2n < 0) return fail(PINTT_ERR_TOO_LARGE, "product exceeds the 2^25 transform limit");
    // Only products that some cut applies to are priced: most products of a
    // tree level are near-balanced or too short for a shorter transform, and
    // the caller's pricing runs on the host once per product.
    auto cuttable = [&]() {
        for (uint32_t m = 2; m <= mode.max_chunks && m <= (uint32_t)kMaxChunks; ++m) {
            size_t h;
            int k;
            if (chunk_shape(r.an, r.bn, m, &h, &k)) return true;
        }
        return false;
    };
    if (!r.square && g.split_mode == 1 && g.chunk_cost && cuttable()) {
        // The caller's prices, cheapest allowed cut first.
        double cost[kMaxChunks];
        if (g.chunk_cost(r.an, r.bn, cost, g.chunk_ctx)) {
            double best = 0;
            bool found = false;
            for (uint32_t m = 1; m <= mode.max_chunks && m <= (uint32_t)kMaxChunks; ++m) {
                size_t h;
                int k;
                if (!chunk_shape(r.an, r.bn, m, &h, &k) || !(cost[m - 1] == cost[m - 1])) continue;
                if (!found || cost[m - 1] < best) {
                    found = true;
                    best = cost[m - 1];
                    gr->m = m;
                    gr->h = h;
                    gr->log2n = k;
                }
            }
            if (found) {
                gr->nblk = (2 * (r.an + r.bn) + carry_block_words(carry_threads()) - 1) /
                           carry_block_words(carry_threads());
                return PINTT_OK;
            }
        }
    }
    // The tuned times price throughput: a product or a small batch below
    // 2^kChunkAloneMinLog2N points is bound by latency, where the shorter
    // transforms of a cut do not pay for its extra transform (a single 2^19-
    // point product cut in two took 13% longer).  Forced cuts still apply.
    const bool latency_bound = !mode.throughput && gr->log2n < kChunkAloneMinLog2N;
    if (!r.square && g.split_mode && !(latency_bound && g.split_mode == 1)) {
        if (mode.tune && g.autotune && g.split_mode == 1 && !g.have_plan[gr->log2n]) {
            Plan tmp;
            PINTT_TRY(plan_for(gr->log2n, &tmp));
        }
        double best = transform_cost(gr->log2n, 3);
        for (uint32_t m = 2; m <= mode.max_chunks; ++m) {
            size_t h;
            int k;
            if (!chunk_shape(r.an, r.bn, m, &h, &k)) continue;
            if (mode.tune && g.autotune && g.split_mode == 1 && !g.have_plan[k]) {
                Plan tmp;
                PINTT_TRY(plan_for(k, &tmp));
            }
            double c = transform_cost(k, 2 * (int)m + 1);
            if (g.split_mode >= 2) c = (m == (uint32_t)g.split_mode) ? -1.0 : 1e300;  // forced
            if (c < best) {
                best = c;
                gr->m = m;
                gr->h = h;
                gr->log2n = k;
            }
        }
    }
    const size_t wb = carry_block_words(carry_threads());
    gr->nblk = (2 * (r.an + r.bn) + wb - 1) / wb;
    return PINTT_OK;
}

struct RowLaunch {
    RowKind kind;
    size_t off;  // byte offset of its RowJob array in the descriptor blob
    uint32_t count;
    uint32_t chunk;  // chunk index it serves (kRowPair: chunks 0 and 1)
};

struct BucketLayout {
    int log2n;
    std::vector<size_t> groups;                   // indices into the group list
    size_t fwd_off, inv_off, carry_off;           // byte offsets in the descriptor blob
    uint32_t nfwd, ninv, nblocks;
    uint32_t max_m;
    std::vector<RowLaunch> rows;                  // in launch order
    bool pair;                                    // chunks 0 and 1 of a group share one row launch
    size_t arena_base = 0;                        // first arena word of its slots
};

template <class T>
size_t align_up(size_t x) {
    return (x + alignof(T) - 1) / alignof(T) * alignof(T);
}

struct BatchPlan {
    std::vector<Group> groups;
    std::vector<BucketLayout> buckets;
    std::vector<Plan> plans;
    std::vector<size_t> school;  // requests small enough for the schoolbook kernel
    size_t school_off = 0;
    size_t bytes = 0, arena_words = 0, total_groups = 0, total_blocks = 0;
    int min_log2n = kMaxLog2N + 1, max_log2n = 0;
};

// Products whose operands both fit the schoolbook kernel run there (unless
// the tuner is timing transforms).
size_t school_max_limbs() {
    static const size_t m = [] {
        const char* e = std::getenv("PINTT_SCHOOL_MAX");
        const long v = e ? std::atol(e) : kSchoolMaxLimbs;
        return (size_t)std::max(0L, std::min(v, (long)kSchoolMaxLimbs));
    }();
    return m;
}
bool school_eligible(const Request& r) {
    return g.school && r.an <= school_max_limbs() && r.bn <= school_max_limbs();
}

// Groups, buckets, descriptor layout and resource sizes of a batch.  May tune
// tile plans for lengths seen for the first time (which waits on the device).
int plan_batch(const std::vector<Request>& reqs, BatchPlan* bp, const PlanMode& mode_in = PlanMode{}) {
    bp->groups.resize(reqs.size());
    // Points of the batch's transforms, uncut: whether it is bound by
    // throughput (the regime the tuned times measure) or by latency.
    PlanMode mode = mode_in;
    {
        size_t points = 0;
        for (const Request& r : reqs) {
            if (school_eligible(r)) continue;
            const int k = log2_for(r.an, r.bn);
            if (k >= 0) points += (r.square ? 2u : 3u) * ((size_t)1 << k);
        }
        mode.throughput = points >= kBatchPlanMinPoints;
    }
    // Schoolbook products need no transform plan: only their request is kept
    // (a tree level's first batches hold tens of thousands of them).
    for (size_t i = 0; i < reqs.size(); ++i) {
        if (school_eligible(reqs[i])) {
            bp->groups[i].r = reqs[i];
            bp->school.push_back(i);
        } else {
            PINTT_TRY(plan_group(reqs[i], &bp->groups[i], mode));
        }
    }
    const std::vector<Group>& groups = bp->groups;
    std::vector<BucketLayout>& buckets = bp->buckets;
    for (size_t i = 0; i < groups.size(); ++i) {
        if (school_eligible(reqs[i])) continue;
        auto it = std::find_if(buckets.begin(), buckets.end(),
                               [&](const BucketLayout& b) { return b.log2n == groups[i].log2n; });
        if (it == buckets.end()) {
            buckets.push_back(BucketLayout{});
            it = buckets.end() - 1;
            it->log2n = groups[i].log2n;
        }
        it->groups.push_back(i);
    }
    // Tile plans first: the row launches depend on them.  A bucket with
    // kBatchPlanMinPoints points of forward transforms or more takes the
    // length's batch plan when there is one.
    bp->plans.resize(buckets.size());
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
        const int k = buckets[bi].log2n;
        size_t nfwd = 0;
        for (size_t gi : buckets[bi].groups) nfwd += groups[gi].m + (groups[gi].r.square ? 0 : 1);
        if (g.have_bplan[k] && (nfwd << k) >= kBatchPlanMinPoints) {
            bp->plans[bi] = g.bplan[k];
        } else if (mode.tune) {
            PINTT_TRY(plan_for(k, &bp->plans[bi]));
        } else {
            bp->plans[bi] = plan_lookup(k);
        }
    }
    // Buckets run one after another on one stream and reuse the arena.
    size_t bytes = 0;
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
        BucketLayout& b = buckets[bi];
        const Plan& pl = bp->plans[bi];
        b.nfwd = b.ninv = b.nblocks = b.max_m = 0;
        size_t slots = 0, nsq = 0;
        for (size_t gi : b.groups) {
            const Group& gr = groups[gi];
            b.nfwd += gr.m + (gr.r.square ? 0 : 1);
            b.ninv += gr.m;
            b.nblocks += (uint32_t)gr.nblk;
            slots += gr.m + (gr.r.square ? 0 : 1);
            b.max_m = std::max(b.max_m, gr.r.square ? 0u : gr.m);
            nsq += gr.r.square ? 1 : 0;
        }
        b.pair = g.pair_rows && pair_case_available(set_for(pl.log2n), pl.log2c, pl.log2rows);
        bytes = align_up<FwdJob>(bytes);
        b.fwd_off = bytes;
        bytes += b.nfwd * sizeof(FwdJob);
        // Row launches: squares; first chunks (transform b), or first pairs of
        // chunks; later chunks (reuse b's transformed rows).
        b.rows.clear();
        auto add_rows = [&](RowKind kind, uint32_t count, uint32_t chunk) {
            if (!count) return;
            bytes = align_up<RowJob>(bytes);
            b.rows.push_back(RowLaunch{kind, bytes, count, chunk});
            bytes += count * sizeof(RowJob);
        };
        auto count_groups = [&](auto pred) {
            uint32_t c = 0;
            for (size_t gi : b.groups)
                if (!groups[gi].r.square && pred(groups[gi].m)) ++c;
            return c;
        };
        add_rows(kRowSquare, (uint32_t)nsq, 0);
        if (b.pair) {
            add_rows(kRowTransformB, count_groups([](uint32_t m) { return m == 1; }), 0);
            add_rows(kRowPair, count_groups([](uint32_t m) { return m >= 2; }), 0);
            for (uint32_t i = 2; i < b.max_m; ++i)
                add_rows(kRowReuseB, count_groups([i](uint32_t m) { return m > i; }), i);
        } else {
            add_rows(kRowTransformB, count_groups([](uint32_t m) { return m >= 1; }), 0);
            for (uint32_t i = 1; i < b.max_m; ++i)
                add_rows(kRowReuseB, count_groups([i](uint32_t m) { return m > i; }), i);
        }
        bytes = align_up<InvJob>(bytes);
        b.inv_off = bytes;
        bytes += b.ninv * sizeof(InvJob);
        bytes = align_up<CarryJob>(bytes);
        b.carry_off = bytes;
        bytes += b.groups.size() * sizeof(CarryJob);
        const size_t words = slots * kPrimes * ((size_t)1 << b.log2n);
        if (mode.disjoint) {
            b.arena_base = bp->arena_words;
            bp->arena_words += words;
        } else {
            bp->arena_words = std::max(bp->arena_words, words);
        }
        bp->total_groups += b.groups.size();
        bp->total_blocks += b.nblocks;
        bp->min_log2n = std::min(bp->min_log2n, b.log2n);
        bp->max_log2n = std::max(bp->max_log2n, b.log2n);
    }
    bytes = align_up<SchoolJob>(bytes);
    bp->school_off = bytes;
    bytes += bp->school.size() * sizeof(SchoolJob);
    bp->bytes = bytes;
    return PINTT_OK;
}

// Write the descriptors of `bp` into H, pointing at the given arena and carry
// maps.
void fill_descriptors(const BatchPlan& bp, unsigned char* H, uint32_t* arena, uint32_t* maps) {
    SchoolJob* sj = reinterpret_cast<SchoolJob*>(H + bp.school_off);
    for (size_t k = 0; k < bp.school.size(); ++k) {
        const Request& r = bp.groups[bp.school[k]].r;
        sj[k] = SchoolJob{r.a, r.square ? r.a : r.b, r.out, (uint32_t)r.an, (uint32_t)(r.square ? r.an : r.bn)};
    }
    size_t maps_next = 0;
    for (const auto& b : bp.buckets) {
        const size_t n = (size_t)1 << b.log2n;
        const size_t slot = kPrimes * n;
        FwdJob* fj = reinterpret_cast<FwdJob*>(H + b.fwd_off);
        std::vector<RowJob*> rj;
        std::vector<uint32_t> rfill(b.rows.size(), 0);
        for (const RowLaunch& rl : b.rows) rj.push_back(reinterpret_cast<RowJob*>(H + rl.off));
        auto launch_of = [&](RowKind kind, uint32_t chunk) {
            for (size_t li = 0; li < b.rows.size(); ++li)
                if (b.rows[li].kind == kind && b.rows[li].chunk == chunk) return li;
            return b.rows.size();  // not reached: plan_batch made every launch a job needs
        };
        InvJob* ij = reinterpret_cast<InvJob*>(H + b.inv_off);
        CarryJob* cj = reinterpret_cast<CarryJob*>(H + b.carry_off);
        size_t next_slot = 0;
        uint32_t f = 0, inv = 0, blk0 = 0;
        for (size_t k = 0; k < b.groups.size(); ++k) {
            const Group& gr = bp.groups[b.groups[k]];
            const Request& r = gr.r;
            CarryJob car{};
            uint32_t* const base = arena + b.arena_base;
            uint32_t* bslot = r.square ? nullptr : base + slot * (next_slot + gr.m);
            const size_t first = next_slot;
            const bool pair = b.pair && !r.square && gr.m >= 2;
            for (uint32_t i = 0; i < gr.m; ++i) {
                uint32_t* dst = base + slot * next_slot++;
                size_t lo = i * gr.h;
                size_t limbs = std::min(gr.h, r.an - lo);
                fj[f++] = FwdJob{reinterpret_cast<const uint32_t*>(r.a + lo), (uint32_t)(2 * limbs), dst};
                if (r.square) {
                    const size_t li = launch_of(kRowSquare, 0);
                    rj[li][rfill[li]++] = RowJob{dst, nullptr, nullptr, 0};
                } else if (pair && i < 2) {
                    if (i == 0) {
                        const size_t li = launch_of(kRowPair, 0);
                        rj[li][rfill[li]++] = RowJob{dst, bslot, base + slot * (first + 1), gr.m > 2 ? 1u : 0u};
                    }
                } else if (i == 0) {
                    const size_t li = launch_of(kRowTransformB, 0);
                    rj[li][rfill[li]++] = RowJob{dst, bslot, nullptr, gr.m > 1 ? 1u : 0u};
                } else {
                    const size_t li = launch_of(kRowReuseB, i);
                    rj[li][rfill[li]++] = RowJob{dst, bslot, nullptr, 0};
                }
                ij[inv++] = InvJob{dst};
                car.res[i] = dst;
                car.off[i] = (uint32_t)(2 * lo);
                car.len[i] = (uint32_t)std::min(n, 2 * limbs + 2 * r.bn - 1);
            }
            if (!r.square) {
                next_slot++;  // bslot
                fj[f++] = FwdJob{reinterpret_cast<const uint32_t*>(r.b), (uint32_t)(2 * r.bn), bslot};
            }
            car.m = gr.m;
            car.blk0 = blk0;
            car.out = reinterpret_cast<uint32_t*>(r.out);
            car.nwords = 2 * (r.an + r.bn);
            car.maps = maps + maps_next;
            cj[k] = car;
            blk0 += (uint32_t)gr.nblk;
            maps_next += gr.nblk;
        }
    }
}

// The kernel launches of a batch whose descriptors sit at device address D.
// True when every launch of the plan can carry its jobs in kernel arguments.
bool fits_inline(const BatchPlan& bp) {
    if (bp.school.size() > (size_t)kInlineJobs) return false;
    for (const BucketLayout& b : bp.buckets) {
        if (b.nfwd > (uint32_t)kInlineJobs || b.ninv > (uint32_t)kInlineJobs || b.groups.size() > (size_t)kInlineJobs)
            return false;
        for (const RowLaunch& rl : b.rows)
            if (rl.count > (uint32_t)kInlineJobs) return false;
    }
    return true;
}

// D: device copy of the descriptors (may be null when H is given and every
// launch fits inline); H: host copy (null inside captures, which always read
// device descriptors).  tw: twiddle tables to use instead of the library's.
struct LaunchSource {
    const BatchPlan& bp;
    const unsigned char* D;
    const unsigned char* H;
    template <class T>
    JobList<T> list(size_t off, uint32_t n) const {
        if (H && n <= (uint32_t)kInlineJobs) return inline_jobs<T>(reinterpret_cast<const T*>(H + off), n);
        return device_jobs<T>(reinterpret_cast<const T*>(D + off));
    }
};

// The schoolbook products of a batch (one launch), if any.
int enqueue_school(const LaunchSource& src, hipStream_t s) {
    const BatchPlan& bp = src.bp;
    if (bp.school.empty()) return PINTT_OK;
    const uint32_t ns = (uint32_t)bp.school.size();
    PINTT_HIP(launch_schoolbook(src.list<SchoolJob>(bp.school_off, ns), ns, s));
    return PINTT_OK;
}

// The passes of bucket bi, in order on `s`.
int enqueue_bucket(const LaunchSource& src, size_t bi, hipStream_t s, bool timed, const Twiddles* tw_tables) {
    // Grid z carries the job index: launches of more than 65535 jobs go in pieces.
    constexpr uint32_t kMaxJobs = 65535;
    const BucketLayout& b = src.bp.buckets[bi];
    const Plan& pl = src.bp.plans[bi];
    const Twiddles tw = tw_tables ? *tw_tables : twiddles(set_for(b.log2n));
    if (timed) PINTT_HIP(hipEventRecord(g.ev[0], s));
    for (uint32_t k = 0; k < b.nfwd; k += kMaxJobs) {
        const uint32_t c = std::min(kMaxJobs, b.nfwd - k);
        PINTT_HIP(launch_forward_columns(pl, src.list<FwdJob>(b.fwd_off + k * sizeof(FwdJob), c), c, tw, s));
    }
    if (timed) PINTT_HIP(hipEventRecord(g.ev[1], s));
    for (const RowLaunch& rl : b.rows)
        for (uint32_t k = 0; k < rl.count; k += kMaxJobs) {
            const uint32_t c = std::min(kMaxJobs, rl.count - k);
            PINTT_HIP(launch_row_convolution(pl, src.list<RowJob>(rl.off + k * sizeof(RowJob), c), c, rl.kind, tw,
                                             s));
        }
    if (timed) PINTT_HIP(hipEventRecord(g.ev[2], s));
    for (uint32_t k = 0; k < b.ninv; k += kMaxJobs) {
        const uint32_t c = std::min(kMaxJobs, b.ninv - k);
        PINTT_HIP(launch_inverse_columns(pl, src.list<InvJob>(b.inv_off + k * sizeof(InvJob), c), c, tw, s));
    }
    if (timed) PINTT_HIP(hipEventRecord(g.ev[3], s));
    const uint32_t ng = (uint32_t)b.groups.size();
    PINTT_HIP(launch_crt_carry(src.list<CarryJob>(b.carry_off, ng), ng, b.nblocks, b.log2n, crt_constants(b.log2n),
                               s));
    if (timed) PINTT_HIP(hipEventRecord(g.ev[4], s));
    return PINTT_OK;
}

// Every launch of a batch, in order on `s` (buckets one after another).
int enqueue_launches(const BatchPlan& bp, const unsigned char* D, const unsigned char* H, hipStream_t s,
                     bool timed, const Twiddles* tw_tables = nullptr) {
    const LaunchSource src{bp, D, H};
    PINTT_TRY(enqueue_school(src, s));
    for (size_t bi = 0; bi < bp.buckets.size(); ++bi) PINTT_TRY(enqueue_bucket(src, bi, s, timed, tw_tables));
    return PINTT_OK;
}

// Enqueue a batch on `s`.  Every pointer in `reqs` is a device pointer.
int enqueue_batch(const std::vector<Request>& reqs, hipStream_t s) {
    if (reqs.empty()) return PINTT_OK;
    BatchPlan bp;
    PINTT_TRY(plan_batch(reqs, &bp));
    if (!bp.buckets.empty()) {
        PINTT_TRY(ensure_twiddles(bp.min_log2n));
        PINTT_TRY(ensure_twiddles(bp.max_log2n));
    }
    PINTT_TRY(ensure_arena(bp.arena_words));
    PINTT_TRY(ensure_carry_state(bp.total_blocks));
    if (fits_inline(bp)) {
        // Every job list rides in kernel arguments: nothing to upload.
        std::vector<unsigned char> H(bp.bytes);
        fill_descriptors(bp, H.data(), g.arena, g.maps);
        // Ordered after all earlier engine work (the arena and carry maps are
        // shared; tables are built at first use).
        PINTT_TRY(begin_on(s));
        PINTT_TRY(enqueue_launches(bp, nullptr, H.data(), s, g.timing));
    } else {
        const int ring = g.desc_next;
        g.desc_next = (g.desc_next + 1) % Engine::kDescRing;
        PINTT_TRY(ensure_descriptors(ring, bp.bytes));
        fill_descriptors(bp, g.desc_host[ring], g.arena, g.maps);
        // The upload follows all earlier engine work (the slot's previous
        // readers among it); the launches follow the upload.
        PINTT_TRY(begin_on(g.stream));
        PINTT_HIP(hipMemcpyAsync(g.desc_dev[ring], g.desc_host[ring], bp.bytes, hipMemcpyHostToDevice, g.stream));
        PINTT_HIP(hipEventRecord(g.desc_done[ring], g.stream));
        PINTT_TRY(end_on(g.stream));
        PINTT_TRY(begin_on(s));
        PINTT_TRY(enqueue_launches(bp, g.desc_dev[ring], g.desc_host[ring], s, g.timing));
    }
    if (g.timing) g.have_times = true;
    return end_on(s);
}

// A batch captured once into a HIP graph, with its own arena, carry state and
// descriptors, replayed on demand (the operand and output pointers are fixed).
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t* arena = nullptr;
    uint32_t* maps = nullptr;
    unsigned char* desc = nullptr;
    size_t nodes = 0;
};

void free_graph(BatchGraph* bg) {
    if (bg->exec) (void)hipGraphExecDestroy(bg->exec);
    if (bg->graph) (void)hipGraphDestroy(bg->graph);
    (void)hipFree(bg->arena);
    (void)hipFree(bg->maps);
    (void)hipFree(bg->desc);
    delete bg;
}

// Streams and events of a branched capture; released after the capture has
// ended (a stream forked from the capture stays in it until then).
struct Branches {
    std::vector<hipStream_t> st;
    std::vector<hipEvent_t> ev;
    ~Branches() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t x : st)
            if (x) (void)hipStreamDestroy(x);
    }
};

// Capture the launches of `bp` into a graph on the capturing stream cs.  With
// branches, the schoolbook launch and every length bucket (which then own
// disjoint parts of the arena and of the carry maps) hang off one fork node
// as independent chains and join at the end, so the graph may run them at
// the same time: small buckets and the tails of large launches fill the
// machine beside the largest bucket instead of after it.
int capture_launches(const BatchPlan& bp, const unsigned char* D, hipStream_t cs, Branches* br) {
    if (!br || bp.buckets.size() + (bp.school.empty() ? 0 : 1) < 2) return enqueue_launches(bp, D, nullptr, cs, false);
    const LaunchSource src{bp, D, nullptr};
    const size_t nb = bp.buckets.size() + 1;  // the last branch: schoolbook products
    br->st.assign(nb, nullptr);
    br->ev.assign(nb + 1, nullptr);
    for (size_t i = 0; i < nb; ++i)
        if (hipStreamCreateWithFlags(&br->st[i], hipStreamNonBlocking) != hipSuccess)
            return fail(PINTT_ERR_HIP, "branch stream");
    for (size_t i = 0; i <= nb; ++i)
        if (hipEventCreateWithFlags(&br->ev[i], hipEventDisableTiming) != hipSuccess)
            return fail(PINTT_ERR_HIP, "branch event");
    // ev[nb] forks; ev[i] joins branch i.  The buckets with the most
    // transform work are captured first.
    std::vector<size_t> order(bp.buckets.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    auto work = [&bp](size_t i) { return (size_t)bp.buckets[i].nfwd << bp.buckets[i].log2n; };
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work(x) > work(y); });
    if (hipEventRecord(br->ev[nb], cs) != hipSuccess) return fail(PINTT_ERR_HIP, "fork");
    for (size_t i = 0; i < nb; ++i) {
        if (hipStreamWaitEvent(br->st[i], br->ev[nb], 0) != hipSuccess) return fail(PINTT_ERR_HIP, "branch wait");
        PINTT_TRY(i == nb - 1 ? enqueue_school(src, br->st[i])
                              : enqueue_bucket(src, order[i], br->st[i], false, nullptr));
        if (hipEventRecord(br->ev[i], br->st[i]) != hipSuccess || hipStreamWaitEvent(cs, br->ev[i], 0) != hipSuccess)
            return fail(PINTT_ERR_HIP, "join");
    }
    return PINTT_OK;
}

int create_graph(const std::vector<Request>& reqs, BatchGraph** out) {
    BatchPlan bp;
    PlanMode mode;
    mode.disjoint = g.graph_branches;
    PINTT_TRY(plan_batch(reqs, &bp, mode));
    if (!bp.buckets.empty()) {
        PINTT_TRY(ensure_twiddles(bp.min_log2n));
        PINTT_TRY(ensure_twiddles(bp.max_log2n));
    }
    BatchGraph* bg = new BatchGraph;
    auto cleanup = [&](int rc) {
        free_graph(bg);
        return rc;
    };
    if (hipMalloc(reinterpret_cast<void**>(&bg->arena), std::max<size_t>(bp.arena_words, 1) * 4) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&bg->maps), std::max<size_t>(bp.total_blocks, 1) * 4) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&bg->desc), bp.bytes) != hipSuccess)
        return cleanup(fail(PINTT_ERR_NO_MEMORY, "graph buffers"));
    std::vector<unsigned char> H(bp.bytes);
    fill_descriptors(bp, H.data(), bg->arena, bg->maps);
    if (hipMemcpy(bg->desc, H.data(), bp.bytes, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(PINTT_ERR_HIP, "graph descriptor upload"));
    // Capture the launches on a private stream (nothing else runs on it).
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(PINTT_ERR_HIP, "capture stream"));
    int rc = PINTT_OK;
    {
        Branches br;  // outlives the capture
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = fail(PINTT_ERR_HIP, "hipStreamBeginCapture");
        } else {
            rc = capture_launches(bp, bg->desc, cs, g.graph_branches ? &br : nullptr);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(cs, &graph);
            if (!rc && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
            bg->graph = graph;
        }
    }
    (void)hipStreamDestroy(cs);
    if (rc) return cleanup(rc);
    if (hipGraphGetNodes(bg->graph, nullptr, &bg->nodes) != hipSuccess) bg->nodes = 0;
    hipError_t e = hipGraphInstantiate(&bg->exec, bg->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipGraphInstantiate"));
    *out = bg;
    return PINTT_OK;
}

// Time every compiled tile plan for length 2^log2n on synthetic operands
// (an = bn = N/4 limbs, no chunking) and keep the fastest.  Below 2^22
// points a single product is launch-bound and every plan measures the same,
// so a plan is timed on a batch of 2^(22 - log2n) products (2^22 points of
// transforms in all, the regime of tree levels and batch calls), all reading
// the same operands and writing their own outputs.  Runs on the library
// stream and waits for it, so it must not happen inside a capture.
int tune(int log2n) {
    const size_t n = (size_t)1 << log2n;
    const size_t an = n / 4, bn = n / 4;
    const size_t count = log2n >= 22 ? 1 : ((size_t)1 << (22 - log2n));
    if (g.tune_cap < an) {
        PINTT_TRY(grow(g.tune_in, 2 * an));
        PINTT_HIP(hipMemset(g.tune_in, 0xA7, 2 * an * sizeof(uint64_t)));
        g.tune_cap = an;
    }
    if (g.tune_out_cap < count * 2 * an) {
        PINTT_TRY(grow(g.tune_out, count * 2 * an));
        g.tune_out_cap = count * 2 * an;
    }
    const int saved_split = g.split_mode;
    const bool saved_timing = g.timing;
    g.split_mode = 0;
    g.timing = false;
    g.school = false;
    Plan best = make_plan(log2n);
    float best_ms = 1e30f;
    std::vector<Request> one;
    for (size_t i = 0; i < count; ++i)
        one.push_back(Request{g.tune_in, an, g.tune_in + an, bn, g.tune_out + i * 2 * an, false});
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int rc = hipEventCreate(&t0) == hipSuccess && hipEventCreate(&t1) == hipSuccess
                 ? PINTT_OK
                 : fail(PINTT_ERR_HIP, "hipEventCreate");
    for (const Plan& pl : candidate_plans(log2n)) {
        if (rc) break;
        g.plan[log2n] = pl;
        g.have_plan[log2n] = true;
        const int reps = log2n >= 22 ? 3 : 5;
        if ((rc = enqueue_batch(one, g.stream))) break;  // warm-up
        if (hipEventRecord(t0, g.stream) != hipSuccess) rc = fail(PINTT_ERR_HIP, "hipEventRecord");
        for (int i = 0; i < reps && !rc; ++i) rc = enqueue_batch(one, g.stream);
        float ms = 0.f;
        if (!rc && (hipEventRecord(t1, g.stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                    hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
            rc = fail(PINTT_ERR_HIP, "timing a tile plan");
        if (rc) break;
        ms /= reps;
        if (ms < best_ms) {
            best_ms = ms;
            best = pl;
        }
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    g.split_mode = saved_split;
    g.school = true;
    g.timing = saved_timing;
    g.plan[log2n] = best;
    // Kept per product: the chunk rule prices lengths against each other.
    g.plan_ms[log2n] = best_ms < 1e29f ? best_ms / (float)count : 0.f;
    g.have_plan[log2n] = true;
    return rc;
}

int plan_for(int log2n, Plan* out) {
    if (!g.have_plan[log2n]) {
        if (g.autotune) {
            PINTT_TRY(tune(log2n));
        } else {
            g.plan[log2n] = make_plan(log2n);
            g.have_plan[log2n] = true;
        }
    }
    *out = g.plan[log2n];
    return PINTT_OK;
}

int ensure_io(size_t limbs) {
    if (limbs <= g.io_cap) return PINTT_OK;
    PINTT_TRY(grow(g.io_in, limbs));
    PINTT_TRY(grow(g.io_out, limbs));
    g.io_cap = limbs;
    return PINTT_OK;
}

// A staging slot for one caller, returned (after its last transfer is
// complete) when the holder goes out of scope.  Needs an initialised engine.
class StagingSlot {
public:
    StagingSlot() = default;
    StagingSlot(const StagingSlot&) = delete;
    StagingSlot& operator=(const StagingSlot&) = delete;
    ~StagingSlot() {
        if (!s_) return;
        if (pending_) (void)hipEventSynchronize(s_->done);
        {
            std::lock_guard<std::mutex> lock(g_stage_mu);
            s_->busy = false;
        }
        g_stage_cv.notify_one();
    }
    int acquire(size_t limbs) {
        {
            std::unique_lock<std::mutex> lock(g_stage_mu);
            g_stage_cv.wait(lock, [] {
                for (const Staging& x : g_stage)
                    if (!x.busy) return true;
                return false;
            });
            for (Staging& x : g_stage)
                if (!x.busy) {
                    s_ = &x;
                    break;
                }
            s_->busy = true;
        }
        PINTT_HIP(hipSetDevice(g.device));
        if (!s_->done) PINTT_HIP(hipEventCreateWithFlags(&s_->done, hipEventDisableTiming));
        if (limbs <= s_->cap) return PINTT_OK;
        if (s_->in) (void)hipHostFree(s_->in);
        if (s_->out) (void)hipHostFree(s_->out);
        s_->in = s_->out = nullptr;
        s_->cap = 0;
        const size_t cap = std::max(limbs, (size_t)1 << 20) + (limbs >> 3);
        if (hipHostMalloc((void**)&s_->in, cap * sizeof(uint64_t)) != hipSuccess ||
            hipHostMalloc((void**)&s_->out, cap * sizeof(uint64_t)) != hipSuccess)
            return fail(PINTT_ERR_NO_MEMORY, "hipHostMalloc of the staging buffers failed");
        s_->cap = cap;
        return PINTT_OK;
    }
    uint64_t* in() const { return s_->in; }
    uint64_t* out() const { return s_->out; }
    // On the library stream after the slot's transfers (engine lock held).
    int record() {
        PINTT_HIP(hipEventRecord(s_->done, g.stream));
        pending_ = true;
        return PINTT_OK;
    }
    // Until those transfers are complete (engine lock not held).
    int wait() {
        pending_ = false;
        PINTT_HIP(hipEventSynchronize(s_->done));
        return PINTT_OK;
    }

private:
    Staging* s_ = nullptr;
    bool pending_ = false;
};

// Copies pieces (dst, src, limbs) with up to `threads` host threads, each
// taking an equal share of the total bytes (pieces are cut where needed).
struct Piece {
    uint64_t* dst;
    const uint64_t* src;
    size_t n;
};

void parallel_copy(const std::vector<Piece>& pieces, size_t total) {
    const size_t kPerThread = (size_t)1 << 17;  // 1 MiB of limbs at least per thread
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    size_t nt = std::min<size_t>({(size_t)hw, (size_t)16, std::max<size_t>(1, total / kPerThread)});
    auto run = [&](size_t lo, size_t hi) {  // limbs [lo, hi) of the concatenation
        size_t pos = 0;
        for (const Piece& p : pieces) {
            size_t a = std::max(lo, pos), b = std::min(hi, pos + p.n);
            if (a < b) std::memcpy(p.dst + (a - pos), p.src + (a - pos), (b - a) * sizeof(uint64_t));
            pos += p.n;
            if (pos >= hi) break;
        }
    };
    if (nt <= 1) {
        run(0, total);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t) th.emplace_back(run, total * t / nt, total * (t + 1) / nt);
    for (auto& t : th) t.join();
}

int wait_stream() { return drain(); }

bool overlaps(const uint64_t* x, size_t xn, const uint64_t* y, size_t yn) { return x < y + yn && y < x + xn; }

// Host operands: staged into device memory, products computed as one batch,
// results copied back.
int host_batch(size_t count, const uint64_t* const* a, const size_t* an, const uint64_t* const* b,
               const size_t* bn, uint64_t* const* out) {
    if (!a || !an || !b || !bn || !out) return fail(PINTT_ERR_ARGS, "null pointer");
    size_t total_out = 0;
    for (size_t i = 0; i < count; ++i) {
        if (!a[i] || !b[i] || !out[i] || an[i] == 0 || bn[i] == 0)
            return fail(PINTT_ERR_ARGS, "null pointer or zero length");
        if (overlaps(out[i], an[i] + bn[i], a[i], an[i]) || overlaps(out[i], an[i] + bn[i], b[i], bn[i]))
            return fail(PINTT_ERR_ARGS, "output overlaps an operand");
        if (an[i] + bn[i] > ((size_t)1 << (kMaxLog2N - 1)))
            return fail(PINTT_ERR_TOO_LARGE, "product exceeds the 2^25 transform limit");
        total_out += an[i] + bn[i];
    }
    // Operand offsets inside the staging (host and device alike).  An operand
    // that several products share (same address and length, e.g. one power
    // of ten times many numbers) is staged and copied once.
    std::vector<size_t> a_off(count), b_off(count), o_off(count);
    std::vector<std::pair<const uint64_t*, size_t>> staged_src;  // (address, limbs) in staging order
    std::unordered_map<const uint64_t*, std::vector<size_t>> seen;  // address -> indices into staged
    std::vector<size_t> staged_off;
    size_t total_in = 0;
    auto stage = [&](const uint64_t* x, size_t n) {
        auto& v = seen[x];
        for (size_t k : v)
            if (staged_src[k].second == n) return staged_off[k];
        v.push_back(staged_src.size());
        staged_src.emplace_back(x, n);
        staged_off.push_back(total_in);
        total_in += n;
        return staged_off.back();
    };
    for (size_t i = 0; i < count; ++i) {
        a_off[i] = stage(a[i], an[i]);
        b_off[i] = stage(b[i], bn[i]);
    }
    {
        std::lock_guard<std::mutex> lock(g_mu);
        PINTT_TRY(ensure_init(-1));
    }
    StagingSlot st;
    PINTT_TRY(st.acquire(std::max(total_in, total_out)));
    std::vector<Piece> gather, scatter;
    gather.reserve(staged_src.size());
    scatter.reserve(count);
    for (size_t k = 0; k < staged_src.size(); ++k)
        gather.push_back(Piece{st.in() + staged_off[k], staged_src[k].first, staged_src[k].second});
    size_t out_off = 0;
    for (size_t i = 0; i < count; ++i) {
        o_off[i] = out_off;
        scatter.push_back(Piece{out[i], st.out() + out_off, an[i] + bn[i]});
        out_off += an[i] + bn[i];
    }
    parallel_copy(gather, total_in);
    {
        std::lock_guard<std::mutex> lock(g_mu);
        PINTT_HIP(hipSetDevice(g.device));
        PINTT_TRY(ensure_io(std::max(total_in, total_out)));
        std::vector<Request> reqs(count);
        for (size_t i = 0; i < count; ++i)
            reqs[i] = Request{g.io_in + a_off[i], an[i], g.io_in + b_off[i], bn[i], g.io_out + o_off[i],
                              a[i] == b[i] && an[i] == bn[i]};
        PINTT_TRY(begin_on(g.stream));
        PINTT_HIP(hipMemcpyAsync(g.io_in, st.in(), total_in * sizeof(uint64_t), hipMemcpyHostToDevice, g.stream));
        PINTT_TRY(enqueue_batch(reqs, g.stream));
        PINTT_HIP(hipMemcpyAsync(st.out(), g.io_out, total_out * sizeof(uint64_t), hipMemcpyDeviceToHost, g.stream));
        PINTT_TRY(end_on(g.stream));
        PINTT_TRY(st.record());
    }
    PINTT_TRY(st.wait());
    parallel_copy(scatter, total_out);
    return PINTT_OK;
}

}  // namespace
}  // namespace pintt

using namespace pintt;

extern "C" {

int gpu_ntt_init(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    return ensure_init(device);
}

void gpu_ntt_shutdown(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g.ready) return;
    (void)hipSetDevice(g.device);
    (void)drain();
    for (int set = 0; set < kSets; ++set)
        for (int i = 0; i < kPrimes; ++i) {
            (void)hipFree(g.tw_fwd[set][i]);
            (void)hipFree(g.tw_inv[set][i]);
            (void)hipFree(g.tw_row[set][i]);
        }
    (void)hipFree(g.arena);
    (void)hipFree(g.maps);
    (void)hipFree(g.tw_pow);
    for (int k = 0; k < Engine::kDescRing; ++k) {
        (void)hipFree(g.desc_dev[k]);
        if (g.desc_host[k]) (void)hipHostFree(g.desc_host[k]);
        (void)hipEventDestroy(g.desc_done[k]);
    }
    (void)hipFree(g.io_in);
    (void)hipFree(g.io_out);
    for (Staging& x : g_stage) {
        if (x.in) (void)hipHostFree(x.in);
        if (x.out) (void)hipHostFree(x.out);
        if (x.done) (void)hipEventDestroy(x.done);
        x = Staging{};
    }
    if (g.add_jobs_host) (void)hipHostFree(g.add_jobs_host);
    (void)hipFree(g.add_jobs_dev);
    (void)hipFree(g.add_status);
    (void)hipFree(g.add_maps);
    (void)hipFree(g.add_carry);
    (void)hipFree(g.add_error);
    (void)hipFree(g.bs_len);
    (void)hipFree(g.bs_sign);
    (void)hipFree(g.bs_err);
    if (g.leaf_jobs_host) (void)hipHostFree(g.leaf_jobs_host);
    (void)hipFree(g.leaf_jobs_dev);
    (void)hipFree(g.tune_in);
    (void)hipFree(g.tune_out);
    for (auto& e : g.ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(g.done);
    (void)hipStreamDestroy(g.stream);
    const pintt_chunk_cost_fn cf = g.chunk_cost;  // a caller's registration outlives the device state
    void* const cc = g.chunk_ctx;
    g = Engine{};
    g.chunk_cost = cf;
    g.chunk_ctx = cc;
}

int gpu_ntt_mul(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, uint64_t* out) {
    return host_batch(1, &a, &an, &b, &bn, &out);
}

int gpu_ntt_sqr(const uint64_t* a, size_t an, uint64_t* out) {
    return host_batch(1, &a, &an, &a, &an, &out);
}

int gpu_ntt_mul_batch(size_t count, const uint64_t* const* a, const size_t* an, const uint64_t* const* b,
                      const size_t* bn, uint64_t* const* out) {
    if (count == 0) return PINTT_OK;
    return host_batch(count, a, an, b, bn, out);
}

int gpu_ntt_mul_batch_device(size_t count, const uint64_t* const* d_a, const size_t* an,
                             const uint64_t* const* d_b, const size_t* bn, uint64_t* const* d_out,
                             void* stream) {
    if (count == 0) return PINTT_OK;
    if (!d_a || !an || !d_b || !bn || !d_out) return fail(PINTT_ERR_ARGS, "null pointer");
    std::vector<Request> reqs(count);
    for (size_t i = 0; i < count; ++i) {
        if (!d_a[i] || !d_b[i] || !d_out[i] || an[i] == 0 || bn[i] == 0)
            return fail(PINTT_ERR_ARGS, "null pointer or zero length");
        if (an[i] + bn[i] > ((size_t)1 << (kMaxLog2N - 1)))
            return fail(PINTT_ERR_TOO_LARGE, "product exceeds the 2^25 transform limit");
        reqs[i] = Request{d_a[i], an[i], d_b[i], bn[i], d_out[i], d_a[i] == d_b[i] && an[i] == bn[i]};
    }
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_HIP(hipSetDevice(g.device));
    return enqueue_batch(reqs, stream ? static_cast<hipStream_t>(stream) : g.stream);
}

int gpu_ntt_mul_device(const uint64_t* d_a, size_t an, const uint64_t* d_b, size_t bn, uint64_t* d_out,
                       void* stream) {
    return gpu_ntt_mul_batch_device(1, &d_a, &an, &d_b, &bn, &d_out, stream);
}

int gpu_ntt_mul_device_noalloc(const uint64_t* d_a, size_t an, const uint64_t* d_b, size_t bn, uint64_t* d_out,
                               void* stream) {
    if (!d_a || !d_b || !d_out || an == 0 || bn == 0) return fail(PINTT_ERR_ARGS, "null pointer or zero length");
    if (an + bn > (size_t)PINTT_NOALLOC_MAX_LIMBS)
        return fail(PINTT_ERR_TOO_LARGE, "gpu_ntt_mul_device_noalloc: an + bn = " + std::to_string(an + bn) +
                                             " limbs exceeds the " + std::to_string(PINTT_NOALLOC_MAX_LIMBS) +
                                             "-limb limit of its static scratch (use gpu_ntt_mul_device)");
    const std::vector<Request> one{Request{d_a, an, d_b, bn, d_out, d_a == d_b && an == bn}};
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_HIP(hipSetDevice(g.device));
    // Plans as loaded or tuned before, else the heuristic: never tuned here.
    BatchPlan bp;
    PlanMode mode;
    mode.tune = false;
    mode.max_chunks = kNoAllocMaxChunks;
    PINTT_TRY(plan_batch(one, &bp, mode));
    if (!fits_inline(bp) || bp.arena_words > kNaArenaWords || bp.total_blocks > kNaMapWords ||
        (!bp.buckets.empty() && bp.max_log2n > kNoAllocMaxLog2N))
        return fail(PINTT_ERR_TOO_LARGE, "gpu_ntt_mul_device_noalloc: the product's plan does not fit the scratch");
    std::vector<unsigned char> H(bp.bytes);
    fill_descriptors(bp, H.data(), g.na_base + kNaArena, g.na_base + kNaMaps);
    // Every job list rides in the kernel arguments: no descriptor upload, no
    // ordering with the library's own products (whose buffers are separate).
    return enqueue_launches(bp, nullptr, H.data(), stream ? static_cast<hipStream_t>(stream) : g.stream, false,
                            &g.na_tw);
}

int gpu_ntt_graph_create(size_t count, const uint64_t* const* d_a, const size_t* an, const uint64_t* const* d_b,
                         const size_t* bn, uint64_t* const* d_out, void** graph) {
    if (!graph || count == 0 || !d_a || !an || !d_b || !bn || !d_out) return fail(PINTT_ERR_ARGS, "bad arguments");
    std::vector<Request> reqs(count);
    for (size_t i = 0; i < count; ++i) {
        if (!d_a[i] || !d_b[i] || !d_out[i] || an[i] == 0 || bn[i] == 0)
            return fail(PINTT_ERR_ARGS, "null pointer or zero length");
        if (an[i] + bn[i] > ((size_t)1 << (kMaxLog2N - 1)))
            return fail(PINTT_ERR_TOO_LARGE, "product exceeds the 2^25 transform limit");
        reqs[i] = Request{d_a[i], an[i], d_b[i], bn[i], d_out[i], d_a[i] == d_b[i] && an[i] == bn[i]};
    }
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_HIP(hipSetDevice(g.device));
    PINTT_TRY(drain());
    BatchGraph* bg = nullptr;
    PINTT_TRY(create_graph(reqs, &bg));
    *graph = bg;
    return PINTT_OK;
}

int gpu_ntt_graph_launch(void* graph, void* stream) {
    if (!graph) return fail(PINTT_ERR_ARGS, "null graph");
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    BatchGraph* bg = static_cast<BatchGraph*>(graph);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g.stream;
    // Serialised with every other batch on the device, after the tables.
    PINTT_TRY(begin_on(s));
    PINTT_HIP(hipGraphLaunch(bg->exec, s));
    return end_on(s);
}

// End of the device allocation holding p (p itself when unknown: no merging).
static const uint64_t* allocation_end(const uint64_t* p) {
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<uint64_t*>(p)) != hipSuccess || !base) {
        (void)hipGetLastError();
        return p;
    }
    return reinterpret_cast<const uint64_t*>(static_cast<const char*>(base) + size);
}

int gpu_ntt_dev_alloc(size_t limbs, uint64_t** d_ptr) {
    if (!d_ptr || limbs == 0) return fail(PINTT_ERR_ARGS, "bad arguments");
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_HIP(hipSetDevice(g.device));
    if (hipMalloc(reinterpret_cast<void**>(d_ptr), limbs * sizeof(uint64_t)) != hipSuccess) {
        *d_ptr = nullptr;
        return fail(PINTT_ERR_NO_MEMORY, "hipMalloc failed");
    }
    return PINTT_OK;
}

int gpu_ntt_dev_free(uint64_t* d_ptr) {
    if (!d_ptr) return PINTT_OK;
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_TRY(drain());
    PINTT_HIP(hipFree(d_ptr));
    return PINTT_OK;
}

// Host <-> device copies of many arrays through the pinned staging buffers
// (gathered / scattered by host threads, one DMA each way).
int gpu_ntt_upload(size_t count, uint64_t* const* d_dst, const uint64_t* const* h_src, const size_t* limbs) {
    if (count == 0) return PINTT_OK;
    if (!d_dst || !h_src || !limbs) return fail(PINTT_ERR_ARGS, "null pointer");
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) total += limbs[i];
    {
        std::lock_guard<std::mutex> lock(g_mu);
        PINTT_TRY(ensure_init(-1));
    }
    StagingSlot st;
    PINTT_TRY(st.acquire(total));
    std::vector<Piece> gather;
    gather.reserve(count);
    size_t off = 0;
    for (size_t i = 0; i < count; ++i) {
        gather.push_back(Piece{st.in() + off, h_src[i], limbs[i]});
        off += limbs[i];
    }
    parallel_copy(gather, total);
    {
        std::lock_guard<std::mutex> lock(g_mu);
        PINTT_HIP(hipSetDevice(g.device));
        PINTT_TRY(begin_on(g.stream));
        // One DMA per run of destinations that follow each other inside one allocation.
        off = 0;
        for (size_t i = 0; i < count;) {
            const uint64_t* end = allocation_end(d_dst[i]);
            size_t j = i + 1, run = limbs[i];
            while (j < count && d_dst[j] == d_dst[j - 1] + limbs[j - 1] && d_dst[j] + limbs[j] <= end)
                run += limbs[j++];
            if (run)
                PINTT_HIP(hipMemcpyAsync(d_dst[i], st.in() + off, run * sizeof(uint64_t), hipMemcpyHostToDevice,
                                         g.stream));
            off += run;
            i = j;
        }
        PINTT_TRY(end_on(g.stream));
        PINTT_TRY(st.record());
    }
    return st.wait();
}

int gpu_ntt_download(size_t count, uint64_t* const* h_dst, const uint64_t* const* d_src, const size_t* limbs) {
    if (count == 0) return PINTT_OK;
    if (!h_dst || !d_src || !limbs) return fail(PINTT_ERR_ARGS, "null pointer");
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) total += limbs[i];
    {
        std::lock_guard<std::mutex> lock(g_mu);
        PINTT_TRY(ensure_init(-1));
    }
    StagingSlot st;
    PINTT_TRY(st.acquire(total));
    std::vector<Piece> scatter;
    scatter.reserve(count);
    {
        std::lock_guard<std::mutex> lock(g_mu);
        PINTT_HIP(hipSetDevice(g.device));
        PINTT_TRY(begin_on(g.stream));
        size_t off = 0;
        for (size_t i = 0; i < count;) {
            const uint64_t* end = allocation_end(d_src[i]);
            size_t j = i + 1, run = limbs[i];
            while (j < count && d_src[j] == d_src[j - 1] + limbs[j - 1] && d_src[j] + limbs[j] <= end)
                run += limbs[j++];
            if (run)
                PINTT_HIP(hipMemcpyAsync(st.out() + off, d_src[i], run * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                         g.stream));
            for (size_t k = i; k < j; ++k) {
                scatter.push_back(Piece{h_dst[k], st.out() + off, limbs[k]});
                off += limbs[k];
            }
            i = j;
        }
        PINTT_TRY(end_on(g.stream));
        PINTT_TRY(st.record());
    }
    PINTT_TRY(st.wait());
    parallel_copy(scatter, total);
    return PINTT_OK;
}

int gpu_ntt_dev_copy(uint64_t* d_dst, const uint64_t* d_src, size_t limbs) {
    if (!d_dst || !d_src) return fail(PINTT_ERR_ARGS, "null pointer");
    if (limbs == 0) return PINTT_OK;
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_TRY(begin_on(g.stream));
    PINTT_HIP(hipMemcpyAsync(d_dst, d_src, limbs * sizeof(uint64_t), hipMemcpyDeviceToDevice, g.stream));
    return end_on(g.stream);
}

namespace {
// negative == nullptr: a negative subtraction is an error; else negative[i]
// receives its flag.
int add_batch(size_t count, const uint64_t* const* d_x, const size_t* xn, const uint64_t* const* d_y,
              const size_t* yn, const int* subtract, uint64_t* const* d_out, int* negative) {
    if (count == 0) return PINTT_OK;
    if (!d_x || !xn || !d_y || !yn || !subtract || !d_out) return fail(PINTT_ERR_ARGS, "null pointer");
    std::vector<AddJob> jobs(count);
    size_t blocks = 0;
    for (size_t i = 0; i < count; ++i) {
        if (!d_x[i] || !d_y[i] || !d_out[i] || xn[i] == 0 || yn[i] == 0)
            return fail(PINTT_ERR_ARGS, "null pointer or zero length");
        const size_t w = std::max(xn[i], yn[i]);
        const size_t nb = (w + kAddLimbsPerBlock - 1) / kAddLimbsPerBlock;
        if (blocks + nb > 0xFFFFFFFFull) return fail(PINTT_ERR_TOO_LARGE, "too many add blocks");
        jobs[i] = AddJob{d_x[i], d_y[i], d_out[i], xn[i], yn[i], subtract[i] ? 1u : 0u, (uint32_t)blocks,
                         (uint32_t)nb, 0u};
        blocks += nb;
    }
    // The error flag and the per-job signs come back through a staging slot,
    // waited for without the engine lock: other threads keep enqueueing
    // while this batch and the work before it finish.
    StagingSlot st;
    PINTT_TRY(st.acquire((count + 1) / 2 + 1));
    uint32_t* const host = reinterpret_cast<uint32_t*>(st.out());  // [0]: error flag, [1..count]: signs
    {
        std::lock_guard<std::mutex> lock(g_mu);
        PINTT_TRY(ensure_init(-1));
        PINTT_HIP(hipSetDevice(g.device));
        // Device scratch that earlier batches may still be using grows only
        // after the stream drains.
        if (count > g.add_job_cap || blocks > g.add_blk_cap) {
            PINTT_TRY(drain());
            if (count > g.add_job_cap) {
                (void)hipFree(g.add_status);
                g.add_status = nullptr;
                g.add_job_cap = 0;
                const size_t cap = std::max<size_t>(count, 1024);
                if (hipMalloc(reinterpret_cast<void**>(&g.add_status), cap * sizeof(uint32_t)) != hipSuccess)
                    return fail(PINTT_ERR_NO_MEMORY, "add status");
                if (g.add_jobs_host) (void)hipHostFree(g.add_jobs_host);
                (void)hipFree(g.add_jobs_dev);
                g.add_jobs_host = nullptr;
                g.add_jobs_dev = nullptr;
                if (hipHostMalloc(reinterpret_cast<void**>(&g.add_jobs_host), cap * sizeof(AddJob)) != hipSuccess ||
                    hipMalloc(reinterpret_cast<void**>(&g.add_jobs_dev), cap * sizeof(AddJob)) != hipSuccess)
                    return fail(PINTT_ERR_NO_MEMORY, "add descriptors");
                g.add_job_cap = cap;
            }
            if (blocks > g.add_blk_cap) {
                (void)hipFree(g.add_maps);
                (void)hipFree(g.add_carry);
                g.add_maps = g.add_carry = nullptr;
                g.add_blk_cap = 0;
                const size_t cap = std::max<size_t>(blocks, 1 << 16);
                if (hipMalloc(reinterpret_cast<void**>(&g.add_maps), cap) != hipSuccess ||
                    hipMalloc(reinterpret_cast<void**>(&g.add_carry), cap) != hipSuccess)
                    return fail(PINTT_ERR_NO_MEMORY, "add block state");
                g.add_blk_cap = cap;
            }
        }
        if (!g.add_error && hipMalloc(reinterpret_cast<void**>(&g.add_error), sizeof(uint32_t)) != hipSuccess)
            return fail(PINTT_ERR_NO_MEMORY, "add error flag");
        // Descriptors through the ring that product batches use (its slot
        // waits only for its own previous upload).
        const int ring = g.desc_next;
        g.desc_next = (g.desc_next + 1) % Engine::kDescRing;
        const size_t bytes = count * sizeof(AddJob);
        PINTT_TRY(ensure_descriptors(ring, bytes));
        std::memcpy(g.desc_host[ring], jobs.data(), bytes);
        PINTT_TRY(begin_on(g.stream));
        PINTT_HIP(hipMemcpyAsync(g.desc_dev[ring], g.desc_host[ring], bytes, hipMemcpyHostToDevice, g.stream));
        PINTT_HIP(hipEventRecord(g.desc_done[ring], g.stream));
        // A fill kernel, not hipMemsetAsync: no memset node if a caller ever captures this.
        PINTT_HIP(launch_fill_u32(g.add_error, 0u, 1, g.stream));
        PINTT_HIP(launch_add_batch(reinterpret_cast<const AddJob*>(g.desc_dev[ring]), (uint32_t)count,
                                   (uint32_t)blocks, g.add_maps, g.add_carry, g.add_error, g.add_status, g.stream));
        PINTT_HIP(hipMemcpyAsync(host, g.add_error, sizeof(uint32_t), hipMemcpyDeviceToHost, g.stream));
        if (negative)
            PINTT_HIP(hipMemcpyAsync(host + 1, g.add_status, count * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                     g.stream));
        PINTT_TRY(end_on(g.stream));
        PINTT_TRY(st.record());
    }
    PINTT_TRY(st.wait());
    if (negative) {
        for (size_t i = 0; i < count; ++i) negative[i] = host[1 + i] ? 1 : 0;
    } else if (host[0]) {
        return fail(PINTT_ERR_ARGS, "subtraction with |x| < |y|");
    }
    return PINTT_OK;
}
}  // namespace

int gpu_ntt_add_batch_device(size_t count, const uint64_t* const* d_x, const size_t* xn, const uint64_t* const* d_y,
                             const size_t* yn, const int* subtract, uint64_t* const* d_out) {
    return add_batch(count, d_x, xn, d_y, yn, subtract, d_out, nullptr);
}

int gpu_ntt_add_batch_device_async(size_t count, const uint64_t* const* d_x, const size_t* xn,
                                   const uint64_t* const* d_y, const size_t* yn, const int* subtract,
                                   uint64_t* const* d_out, uint32_t* d_negative) {
    if (count == 0) return PINTT_OK;
    if (!d_x || !xn || !d_y || !yn || !subtract || !d_out || !d_negative) return fail(PINTT_ERR_ARGS, "null pointer");
    std::vector<AddJob> jobs(count);
    size_t blocks = 0;
    for (size_t i = 0; i < count; ++i) {
        if (!d_x[i] || !d_y[i] || !d_out[i] || xn[i] == 0 || yn[i] == 0)
            return fail(PINTT_ERR_ARGS, "null pointer or zero length");
        const size_t w = std::max(xn[i], yn[i]);
        const size_t nb = (w + kAddLimbsPerBlock - 1) / kAddLimbsPerBlock;
        if (blocks + nb > 0xFFFFFFFFull) return fail(PINTT_ERR_TOO_LARGE, "too many add blocks");
        jobs[i] = AddJob{d_x[i], d_y[i], d_out[i], xn[i], yn[i], subtract[i] ? 1u : 0u, (uint32_t)blocks,
                         (uint32_t)nb, 0u};
        blocks += nb;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_HIP(hipSetDevice(g.device));
    // Device scratch that earlier batches may still be using grows only after
    // the stream drains.
    if (count > g.add_job_cap || blocks > g.add_blk_cap) {
        PINTT_TRY(drain());
        if (count > g.add_job_cap) {
            (void)hipFree(g.add_status);
            g.add_status = nullptr;
            const size_t cap = std::max<size_t>(count, 1024);
            if (hipMalloc(reinterpret_cast<void**>(&g.add_status), cap * sizeof(uint32_t)) != hipSuccess)
                return fail(PINTT_ERR_NO_MEMORY, "add status");
            if (g.add_jobs_host) (void)hipHostFree(g.add_jobs_host);
            (void)hipFree(g.add_jobs_dev);
            g.add_jobs_host = nullptr;
            g.add_jobs_dev = nullptr;
            if (hipHostMalloc(reinterpret_cast<void**>(&g.add_jobs_host), cap * sizeof(AddJob)) != hipSuccess ||
                hipMalloc(reinterpret_cast<void**>(&g.add_jobs_dev), cap * sizeof(AddJob)) != hipSuccess)
                return fail(PINTT_ERR_NO_MEMORY, "add descriptors");
            g.add_job_cap = cap;
        }
        if (blocks > g.add_blk_cap) {
            (void)hipFree(g.add_maps);
            (void)hipFree(g.add_carry);
            g.add_maps = g.add_carry = nullptr;
            g.add_blk_cap = 0;
            const size_t cap = std::max<size_t>(blocks, 1 << 16);
            if (hipMalloc(reinterpret_cast<void**>(&g.add_maps), cap) != hipSuccess ||
                hipMalloc(reinterpret_cast<void**>(&g.add_carry), cap) != hipSuccess)
                return fail(PINTT_ERR_NO_MEMORY, "add block state");
            g.add_blk_cap = cap;
        }
    }
    // Descriptors through the ring that product batches use (its slot waits
    // only for its own previous upload).
    const int ring = g.desc_next;
    g.desc_next = (g.desc_next + 1) % Engine::kDescRing;
    const size_t bytes = count * sizeof(AddJob);
    PINTT_TRY(ensure_descriptors(ring, bytes));
    std::memcpy(g.desc_host[ring], jobs.data(), bytes);
    PINTT_TRY(begin_on(g.stream));
    PINTT_HIP(hipMemcpyAsync(g.desc_dev[ring], g.desc_host[ring], bytes, hipMemcpyHostToDevice, g.stream));
    PINTT_HIP(hipEventRecord(g.desc_done[ring], g.stream));
    PINTT_HIP(launch_add_batch(reinterpret_cast<const AddJob*>(g.desc_dev[ring]), (uint32_t)count, (uint32_t)blocks,
                               g.add_maps, g.add_carry, d_negative, g.add_status, g.stream));
    return end_on(g.stream);
}

int gpu_ntt_add_batch_device_flags(size_t count, const uint64_t* const* d_x, const size_t* xn,
                                   const uint64_t* const* d_y, const size_t* yn, const int* subtract,
                                   uint64_t* const* d_out, int* negative) {
    if (count && !negative) return fail(PINTT_ERR_ARGS, "null pointer");
    return add_batch(count, d_x, xn, d_y, yn, subtract, d_out, negative);
}

int gpu_ntt_bs_leaves(uint64_t a, uint64_t b, size_t count, uint64_t* d_out, uint32_t* lengths, int32_t* signs) {
    if (count == 0) return PINTT_OK;
    if (!d_out || !lengths || !signs) return fail(PINTT_ERR_ARGS, "null pointer");
    if (b > (1ull << 24) || a >= b) return fail(PINTT_ERR_ARGS, "terms must satisfy a < b <= 2^24");
    if (count > 0x7FFFFFFFull) return fail(PINTT_ERR_TOO_LARGE, "too many subtrees");
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_HIP(hipSetDevice(g.device));
    if (count > g.bs_cap) {
        PINTT_TRY(drain());
        (void)hipFree(g.bs_len);
        (void)hipFree(g.bs_sign);
        g.bs_len = nullptr;
        g.bs_sign = nullptr;
        g.bs_cap = 0;
        if (hipMalloc(reinterpret_cast<void**>(&g.bs_len), count * 3 * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&g.bs_sign), count * 2 * sizeof(int32_t)) != hipSuccess)
            return fail(PINTT_ERR_NO_MEMORY, "subtree lengths");
        g.bs_cap = count;
    }
    if (!g.bs_err && hipMalloc(reinterpret_cast<void**>(&g.bs_err), sizeof(uint32_t)) != hipSuccess)
        return fail(PINTT_ERR_NO_MEMORY, "subtree error word");
    PINTT_TRY(begin_on(g.stream));
    PINTT_HIP(launch_fill_u32(g.bs_err, 0u, 1, g.stream));
    PINTT_HIP(launch_bs_leaves(a, b, (uint32_t)count, d_out, g.bs_len, g.bs_sign, g.bs_err, g.stream));
    uint32_t err = 0;
    PINTT_HIP(hipMemcpyAsync(lengths, g.bs_len, count * 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, g.stream));
    PINTT_HIP(hipMemcpyAsync(signs, g.bs_sign, count * 2 * sizeof(int32_t), hipMemcpyDeviceToHost, g.stream));
    PINTT_HIP(hipMemcpyAsync(&err, g.bs_err, sizeof(err), hipMemcpyDeviceToHost, g.stream));
    PINTT_TRY(end_on(g.stream));
    PINTT_HIP(hipStreamSynchronize(g.stream));
    if (err) return fail(PINTT_ERR_ARGS, "subtree: a subtraction for R went negative");
    return PINTT_OK;
}

int gpu_ntt_decimal_leaves_device(size_t count, const uint64_t* const* d_x, const size_t* xn, const size_t* width,
                                  const size_t* offset, uint8_t* d_out, size_t out_len) {
    if (count == 0) return PINTT_OK;
    if (!d_x || !xn || !width || !offset || !d_out) return fail(PINTT_ERR_ARGS, "null pointer");
    if (count > 0xFFFFFFFFull) return fail(PINTT_ERR_TOO_LARGE, "too many leaves");
    std::vector<LeafJob> jobs(count);
    for (size_t i = 0; i < count; ++i) {
        if (!d_x[i] || xn[i] == 0 || xn[i] > (size_t)kLeafMaxLimbs || width[i] == 0 ||
            width[i] > (size_t)kLeafMaxDigits || offset[i] > out_len || width[i] > out_len - offset[i])
            return fail(PINTT_ERR_ARGS, "leaf: null pointer, length or width out of range");
        jobs[i] = LeafJob{d_x[i], xn[i], width[i], offset[i]};
    }
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_HIP(hipSetDevice(g.device));
    PINTT_TRY(drain());  // the descriptor buffers below may be reused
    if (count > g.leaf_job_cap) {
        if (g.leaf_jobs_host) (void)hipHostFree(g.leaf_jobs_host);
        (void)hipFree(g.leaf_jobs_dev);
        g.leaf_jobs_host = nullptr;
        g.leaf_jobs_dev = nullptr;
        g.leaf_job_cap = 0;
        const size_t cap = std::max<size_t>(count, 1024);
        if (hipHostMalloc(reinterpret_cast<void**>(&g.leaf_jobs_host), cap * sizeof(LeafJob)) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&g.leaf_jobs_dev), cap * sizeof(LeafJob)) != hipSuccess)
            return fail(PINTT_ERR_NO_MEMORY, "leaf descriptors");
        g.leaf_job_cap = cap;
    }
    std::memcpy(g.leaf_jobs_host, jobs.data(), count * sizeof(LeafJob));
    PINTT_TRY(begin_on(g.stream));
    PINTT_HIP(hipMemcpyAsync(g.leaf_jobs_dev, g.leaf_jobs_host, count * sizeof(LeafJob), hipMemcpyHostToDevice,
                             g.stream));
    PINTT_HIP(launch_decimal_leaves(g.leaf_jobs_dev, (uint32_t)count, d_out, g.stream));
    PINTT_TRY(end_on(g.stream));
    PINTT_HIP(hipStreamSynchronize(g.stream));
    return PINTT_OK;
}

size_t gpu_ntt_graph_nodes(void* graph) { return graph ? static_cast<BatchGraph*>(graph)->nodes : 0; }

void gpu_ntt_graph_destroy(void* graph) {
    if (!graph) return;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g.ready) (void)drain();
    (void)hipDeviceSynchronize();
    free_graph(static_cast<BatchGraph*>(graph));
}

int gpu_ntt_synchronize(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g.ready) return PINTT_OK;
    return wait_stream();
}

int gpu_ntt_transform_log2(size_t an, size_t bn) {
    if (an == 0 || bn == 0) return -1;
    return log2_for(an, bn);
}

int gpu_ntt_product_plan(size_t an, size_t bn, int* log2n, int* chunks, size_t* chunk_limbs) {
    if (an == 0 || bn == 0 || !log2n || !chunks) return fail(PINTT_ERR_ARGS, "bad arguments");
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    PINTT_HIP(hipSetDevice(g.device));
    Group gr;
    static const uint64_t dummy = 0;
    PINTT_TRY(plan_group(Request{&dummy, an, &dummy + (an == bn ? 1 : 0), bn, nullptr, false}, &gr));
    *log2n = gr.log2n;
    *chunks = (int)gr.m;
    if (chunk_limbs) *chunk_limbs = gr.h;
    return PINTT_OK;
}

size_t gpu_ntt_max_limbs(void) { return ((size_t)1 << (kMaxLog2N - 1)); }

void gpu_ntt_set_chunk_cost(pintt_chunk_cost_fn fn, void* ctx) {
    std::lock_guard<std::mutex> lock(g_mu);
    g.chunk_cost = fn;
    g.chunk_ctx = ctx;
}

int gpu_ntt_chunk_plan(size_t an, size_t bn, int chunks, size_t* chunk_limbs, int* log2n) {
    if (an < bn) std::swap(an, bn);
    size_t h;
    int k;
    if (an == 0 || bn == 0 || chunks < 1 || !chunk_shape(an, bn, (uint32_t)chunks, &h, &k)) return 0;
    if (chunk_limbs) *chunk_limbs = h;
    if (log2n) *log2n = k;
    return 1;
}

const char* gpu_ntt_last_error(void) { return t_error.c_str(); }

int gpu_ntt_autotune(int log2n_min, int log2n_max) {
    std::lock_guard<std::mutex> lock(g_mu);
    PINTT_TRY(ensure_init(-1));
    if (log2n_min < kMinLog2N) log2n_min = kMinLog2N;
    if (log2n_max > kMaxLog2N) log2n_max = kMaxLog2N;
    PINTT_HIP(hipSetDevice(g.device));
    for (int k = log2n_min; k <= log2n_max; ++k) PINTT_TRY(tune(k));
    return PINTT_OK;
}

int gpu_ntt_get_plan(int log2n, int* log2r, int* log2w, int* log2rows, float* ms) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (log2n < kMinLog2N || log2n > kMaxLog2N) return fail(PINTT_ERR_ARGS, "length out of range");
    Plan pl = g.have_plan[log2n] ? g.plan[log2n] : make_plan(log2n);
    if (log2r) *log2r = pl.log2r;
    if (log2w) *log2w = pl.log2w;
    if (log2rows) *log2rows = pl.log2rows;
    if (ms) *ms = g.have_plan[log2n] ? g.plan_ms[log2n] : 0.f;
    return PINTT_OK;
}

int gpu_ntt_set_plan(int log2n, int log2r, int log2w, int log2rows) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (log2n < kMinLog2N || log2n > kMaxLog2N) return fail(PINTT_ERR_ARGS, "length out of range");
    Plan pl{log2n, log2r, log2n - log2r, log2w, log2rows};
    if (log2r <= 0) pl = make_plan(log2n);
    if (!plan_valid(pl)) return fail(PINTT_ERR_ARGS, "tile plan not compiled into this library");
    g.plan[log2n] = pl;
    g.have_plan[log2n] = true;
    return PINTT_OK;
}

int gpu_ntt_set_batch_plan(int log2n, int log2r, int log2w, int log2rows) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (log2n < kMinLog2N || log2n > kMaxLog2N) return fail(PINTT_ERR_ARGS, "length out of range");
    if (log2r <= 0) {
        g.have_bplan[log2n] = false;
        return PINTT_OK;
    }
    Plan pl{log2n, log2r, log2n - log2r, log2w, log2rows};
    if (!plan_valid(pl)) return fail(PINTT_ERR_ARGS, "tile plan not compiled into this library");
    g.bplan[log2n] = pl;
    g.have_bplan[log2n] = true;
    return PINTT_OK;
}

int gpu_ntt_get_batch_plan(int log2n, int* log2r, int* log2w, int* log2rows) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (log2n < kMinLog2N || log2n > kMaxLog2N) return fail(PINTT_ERR_ARGS, "length out of range");
    if (!g.have_bplan[log2n]) return fail(PINTT_ERR_ARGS, "no batch plan for this length");
    if (log2r) *log2r = g.bplan[log2n].log2r;
    if (log2w) *log2w = g.bplan[log2n].log2w;
    if (log2rows) *log2rows = g.bplan[log2n].log2rows;
    return PINTT_OK;
}

int gpu_ntt_save_plans(const char* path) {
    if (!path) return fail(PINTT_ERR_ARGS, "null path");
    std::lock_guard<std::mutex> lock(g_mu);
    std::FILE* f = std::fopen(path, "w");
    if (!f) return fail(PINTT_ERR_ARGS, std::string("cannot write ") + path);
    std::fprintf(f, "# log2n log2r log2w log2rows tuned_ms (per product)\n");
    for (int n = kMinLog2N; n <= kMaxLog2N; ++n)
        if (g.have_plan[n])
            std::fprintf(f, "%d %d %d %d %.6g\n", n, g.plan[n].log2r, g.plan[n].log2w, g.plan[n].log2rows,
                         (double)g.plan_ms[n]);
    std::fprintf(f, "# batch log2n log2r log2w log2rows: plans for large buckets\n");
    for (int n = kMinLog2N; n <= kMaxLog2N; ++n)
        if (g.have_bplan[n])
            std::fprintf(f, "batch %d %d %d %d\n", n, g.bplan[n].log2r, g.bplan[n].log2w, g.bplan[n].log2rows);
    return std::fclose(f) == 0 ? PINTT_OK : fail(PINTT_ERR_ARGS, "write failed");
}

int gpu_ntt_load_plans(const char* path) {
    if (!path) return fail(PINTT_ERR_ARGS, "null path");
    std::lock_guard<std::mutex> lock(g_mu);
    std::FILE* f = std::fopen(path, "r");
    if (!f) return fail(PINTT_ERR_ARGS, std::string("cannot read ") + path);
    char line[256];
    int rc = PINTT_OK;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int n, r, w, rows;
        double ms;
        if (std::sscanf(line, "batch %d %d %d %d", &n, &r, &w, &rows) == 4) {
            if (n < kMinLog2N || n > kMaxLog2N) continue;
            Plan pl{n, r, n - r, w, rows};
            if (!plan_valid(pl)) {
                rc = fail(PINTT_ERR_ARGS, "plan file names a tile plan not compiled into this library");
                continue;
            }
            g.bplan[n] = pl;
            g.have_bplan[n] = true;
            continue;
        }
        if (std::sscanf(line, "%d %d %d %d %lf", &n, &r, &w, &rows, &ms) != 5) continue;
        if (n < kMinLog2N || n > kMaxLog2N) continue;
        Plan pl{n, r, n - r, w, rows};
        if (!plan_valid(pl)) {
            rc = fail(PINTT_ERR_ARGS, "plan file names a tile plan not compiled into this library");
            continue;
        }
        g.plan[n] = pl;
        g.plan_ms[n] = (float)ms;
        g.have_plan[n] = true;
    }
    std::fclose(f);
    return rc;
}

int gpu_ntt_list_plans(int log2n, int* out, int max_plans) {
    if (log2n < kMinLog2N || log2n > kMaxLog2N || !out) return 0;
    std::vector<Plan> c = candidate_plans(log2n);
    int k = 0;
    for (const Plan& pl : c) {
        if (k >= max_plans) break;
        out[3 * k] = pl.log2r;
        out[3 * k + 1] = pl.log2w;
        out[3 * k + 2] = pl.log2rows;
        ++k;
    }
    return k;
}

void gpu_ntt_set_split(int mode) {
    std::lock_guard<std::mutex> lock(g_mu);
    g.split_mode = mode < 0 ? 0 : (mode > kMaxChunks ? kMaxChunks : mode);
}

void gpu_ntt_set_graph_branches(int enable) {
    std::lock_guard<std::mutex> lock(g_mu);
    g.graph_branches = enable != 0;
}

void gpu_ntt_set_timing(int enable) {
    std::lock_guard<std::mutex> lock(g_mu);
    g.timing = enable != 0;
    g.have_times = false;
}

int gpu_ntt_pass_times(double* ms, int max) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g.ready || !g.have_times || !ms) return 0;
    if (hipEventSynchronize(g.ev[4]) != hipSuccess) return 0;
    int k = 0;
    for (int i = 0; i < 4 && k < max; ++i) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, g.ev[i], g.ev[i + 1]) != hipSuccess) return k;
        ms[k++] = t;
    }
    return k;
}

}  // extern "C"
