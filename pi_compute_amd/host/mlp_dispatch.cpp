// mlp_dispatch.cpp - see mlp_dispatch.h.
#include "mlp_dispatch.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "pintt_gpu.h"
#include "pintt_mlp.h"

namespace pi {
namespace {
constexpr int IN = PINTT_MLP_INPUTS, H = PINTT_MLP_HIDDEN;
constexpr int oB1 = H * IN, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H;
}  // namespace

MlpFeatures MlpModel::features(const ProductShape& s) {
    const size_t big = std::max(s.an, s.bn), small = std::min(s.an, s.bn);
    MlpFeatures f;
    f[0] = (float)std::log2((double)big);
    f[1] = (float)std::log2((double)small);
    f[2] = (float)std::log2((double)(s.batch ? s.batch : 1));
    f[3] = s.square ? 1.f : 0.f;
    for (int m = 1; m <= 4; ++m) {
        size_t h = 0;
        int k = 0;
        const bool ok = (m == 1 || !s.square) && gpu_ntt_chunk_plan(big, small, m, &h, &k);
        f[3 + m] = ok ? (float)((double)k - std::log2((double)(2 * h + 2 * small - 1))) : -1.f;
    }
    return f;
}

// Operand sizes must lie in the trained range; a batch larger than any
// trained one is priced as the largest (a product's share of a batch stops
// shrinking once the batch fills the device).
bool MlpModel::in_range(const MlpFeatures& f) const {
    for (int j = 0; j < 2; ++j)
        if (f[j] < lo[j] || f[j] > hi[j]) return false;
    return f[2] >= lo[2];
}

MlpFeatures MlpModel::standardise(const MlpFeatures& f) const {
    MlpFeatures x;
    for (int j = 0; j < IN; ++j) x[j] = ((j == 2 ? std::min(f[j], hi[j]) : f[j]) - mean[j]) / scale[j];
    return x;
}

void MlpModel::forward(const float* x, float* y) const {
    const float* w = params.data();
    float h1[H], h2[H];
    for (int i = 0; i < H; ++i) {
        float s = w[oB1 + i];
        for (int j = 0; j < IN; ++j) s += w[i * IN + j] * x[j];
        h1[i] = s > 0.f ? s : 0.f;
    }
    for (int i = 0; i < H; ++i) {
        float s = w[oB2 + i];
        for (int j = 0; j < H; ++j) s += w[oW2 + i * H + j] * h1[j];
        h2[i] = s > 0.f ? s : 0.f;
    }
    for (int k = 0; k < outputs; ++k) {
        float s = w[oW3 + H * outputs + k];
        for (int j = 0; j < H; ++j) s += w[oW3 + k * H + j] * h2[j];
        y[k] = s;
    }
}

// Text format (version 2: PINTT_MLP_INPUTS = 8 inputs; version 1 files had
// four and are refused):
//   pintt-mlp 2
//   outputs K
//   mean m0 .. m7 / scale ... / lo ... / hi ...
//   params N
//   N floats
bool MlpModel::load(const std::string& path, std::string* err) {
    std::ifstream f(path);
    auto bad = [&](const char* what) {
        if (err) *err = path + ": " + what;
        return false;
    };
    if (!f) return bad("cannot open");
    std::string tag, key;
    int version = 0;
    if (!(f >> tag >> version) || tag != "pintt-mlp") return bad("not a pintt-mlp file");
    if (version != 2) return bad("pintt-mlp version 1 (four inputs): retrain with pi_train_mlp");
    if (!(f >> key >> outputs) || key != "outputs" || outputs < 2 || outputs > PINTT_MLP_MAX_OUTPUTS)
        return bad("bad outputs");
    for (auto* arr : {&mean, &scale, &lo, &hi}) {
        if (!(f >> key)) return bad("truncated");
        for (float& v : *arr)
            if (!(f >> v)) return bad("truncated");
    }
    size_t n = 0;
    if (!(f >> key >> n) || key != "params" || n != gpu_mlp_param_count(outputs)) return bad("bad parameter count");
    params.resize(n);
    for (float& v : params)
        if (!(f >> v)) return bad("truncated parameters");
    for (float s : scale)
        if (!(s > 0.f)) return bad("non-positive scale");
    return true;
}

bool MlpModel::save(const std::string& path) const {
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "pintt-mlp 2\noutputs %d\n", outputs);
    const char* names[4] = {"mean", "scale", "lo", "hi"};
    const MlpFeatures* arrs[4] = {&mean, &scale, &lo, &hi};
    for (int a = 0; a < 4; ++a) {
        std::fprintf(f, "%s", names[a]);
        for (float v : *arrs[a]) std::fprintf(f, " %.9g", v);
        std::fprintf(f, "\n");
    }
    std::fprintf(f, "params %zu\n", params.size());
    for (size_t i = 0; i < params.size(); ++i) std::fprintf(f, "%.9g%c", params[i], (i % 8 == 7) ? '\n' : ' ');
    std::fprintf(f, "\n");
    return std::fclose(f) == 0;
}

MlpDispatcher::MlpDispatcher(MlpModel model, const Dispatcher* fallback, unsigned cpu_threads)
    : model_(std::move(model)), fallback_(fallback) {
    if (!fallback_) throw std::invalid_argument("MlpDispatcher needs a fallback dispatcher");
    if (model_.params.size() != gpu_mlp_param_count(model_.outputs))
        throw std::invalid_argument("MlpDispatcher: parameter count does not match the layout");
    unsigned t = cpu_threads ? cpu_threads : std::max(1u, std::thread::hardware_concurrency());
    log_threads_ = std::log((float)t);
}

static MlpModel load_or_throw(const std::string& path) {
    MlpModel m;
    std::string err;
    if (!m.load(path, &err)) throw std::runtime_error("MlpDispatcher: " + err);
    return m;
}

MlpDispatcher::MlpDispatcher(const std::string& model_path, const Dispatcher* fallback, unsigned cpu_threads)
    : MlpDispatcher(load_or_throw(model_path), fallback, cpu_threads) {}

// Predicted GPU output for the shape: resident operands use output 2 when present.
static inline int gpu_output(const MlpModel& m, const ProductShape& s) { return (s.resident && m.outputs > 2) ? 2 : 1; }

Path MlpDispatcher::decide(const ProductShape& s, const float* y) const {
    const float cpu = y[0] - (s.batch > 1 ? log_threads_ : 0.f);
    return y[gpu_output(model_, s)] < cpu ? Path::Gpu : Path::Cpu;
}

// Output of a device-resident product cut into m chunks (models with seven outputs).
constexpr int kChunkOutput0 = 3;

bool MlpDispatcher::chunk_costs(size_t an, size_t bn, double* cost) const {
    if (model_.outputs < kChunkOutput0 + 4) return false;
    ProductShape s{an, bn, 1, false, true};
    const auto f = MlpModel::features(s);
    if (!model_.in_range(f)) return false;
    const auto x = model_.standardise(f);
    float y[PINTT_MLP_MAX_OUTPUTS];
    model_.forward(x.data(), y);
    for (int m = 0; m < 4; ++m) cost[m] = y[kChunkOutput0 + m];
    return true;
}

Path MlpDispatcher::choose(const ProductShape& s) const {
    const auto f = MlpModel::features(s);
    if (!model_.in_range(f)) return fallback_->choose(s);
    const auto x = model_.standardise(f);
    float y[PINTT_MLP_MAX_OUTPUTS];
    model_.forward(x.data(), y);
    return decide(s, y);
}

// A batch runs its device products and its host products at the same time,
// so the batch takes max(sum of device times, sum of host times / threads).
// Products are ranked by predicted host/device time ratio and the device takes
// the prefix of that ranking that minimises the batch time, taken as
// max(device sum, host sum / threads, longest host product).  With products
// divisible this ranking is optimal (exchange argument); whole products make
// the split at most one product's time above the best subset.
void MlpDispatcher::choose_batch(const ProductShape* s, size_t n, Path* out) const {
    if (n == 0) return;
    if (n == 1) {
        out[0] = choose(s[0]);
        return;
    }
    const int K = model_.outputs;
    std::vector<float> x(n * IN), y(n * K);
    std::vector<char> ok(n);
    for (size_t i = 0; i < n; ++i) {
        const auto f = MlpModel::features(s[i]);
        ok[i] = model_.in_range(f);
        const auto z = model_.standardise(f);
        for (int j = 0; j < IN; ++j) x[i * IN + j] = z[j];
    }
    bool scored = false;
    if (n >= device_batch_)
        scored = gpu_mlp_forward(model_.params.data(), K, x.data(), n, y.data()) == PINTT_OK;
    if (!scored)
        for (size_t i = 0; i < n; ++i) model_.forward(&x[i * IN], &y[i * K]);
    const double threads = std::exp((double)log_threads_);
    std::vector<size_t> free_idx;
    for (size_t i = 0; i < n; ++i) {
        if (ok[i]) {
            free_idx.push_back(i);
            continue;
        }
        out[i] = fallback_->choose(s[i]);  // outside the trained range: the rule decides, cost unknown
    }
    auto g = [&](size_t i) { return y[i * K + gpu_output(model_, s[i])]; };
    std::sort(free_idx.begin(), free_idx.end(),
              [&](size_t a, size_t b) { return (y[a * K] - g(a)) > (y[b * K] - g(b)); });
    // A host product runs on one thread: the host share takes at least its
    // longest product (suffix maxima of the ranking).
    const size_t nf = free_idx.size();
    std::vector<double> longest(nf + 1, 0.0);
    for (size_t k = nf; k-- > 0;) longest[k] = std::max(longest[k + 1], std::exp((double)y[free_idx[k] * K]));
    double host = 0, dev = 0;
    for (size_t i : free_idx) host += std::exp((double)y[i * K]);
    double best = std::max({dev, host / threads, longest[0]});
    size_t best_k = 0;
    for (size_t k = 0; k < nf; ++k) {
        const size_t i = free_idx[k];
        dev += std::exp((double)g(i));
        host -= std::exp((double)y[i * K]);
        const double t = std::max({dev, host / threads, longest[k + 1]});
        if (t < best) {
            best = t;
            best_k = k + 1;
        }
    }
    for (size_t k = 0; k < free_idx.size(); ++k) out[free_idx[k]] = k < best_k ? Path::Gpu : Path::Cpu;
}

}  // namespace pi
