// multiplier.cpp - default batch path and the host thread helper.
#include "multiplier.h"

#include <algorithm>
#include <atomic>
#include <numeric>
#include <thread>
#include <vector>

namespace pi {

void parallel_for(size_t count, unsigned threads, void (*fn)(void*, size_t), void* ctx, const double* cost) {
    if (count == 0) return;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::vector<size_t> order(count);
    std::iota(order.begin(), order.end(), 0);
    if (cost) std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    const unsigned nt = (unsigned)std::min<size_t>(threads, count);
    if (nt <= 1) {
        for (size_t i : order) fn(ctx, i);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t k; (k = next.fetch_add(1)) < count;) fn(ctx, order[k]);
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

namespace {
struct BatchCtx {
    Multiplier* m;
    const MulJob* jobs;
};
void run_job(void* p, size_t i) {
    auto* c = static_cast<BatchCtx*>(p);
    const MulJob& j = c->jobs[i];
    if (j.a == j.b) c->m->square(j.result, j.a);
    else c->m->multiply(j.result, j.a, j.b);
}
}  // namespace

void Multiplier::multiply_batch(const MulJob* jobs, size_t count, unsigned threads) {
    std::vector<double> cost(count);
    for (size_t i = 0; i < count; ++i) cost[i] = (double)mpz_size(jobs[i].a) + (double)mpz_size(jobs[i].b);
    BatchCtx ctx{this, jobs};
    parallel_for(count, threads, run_job, &ctx, cost.data());
}

}  // namespace pi
