// train_mlp.cpp - pi_train_mlp: trains the dispatch cost model on the GPU from
// HipNttMultiplier timing logs and writes the model file.
//
//   pi_train_mlp --out MODEL --log FILE [--log FILE ...] [--steps 4000]
//                [--lr 0.003] [--seed 1] [--holdout 0.2] [--threads T]
//                [--holdout-out FILE]
//
// Lines: {"an":..,"bn":..,"batch":..,"square":0|1,"path":P,"ms":..}, P one of
// "cpu", "gpu", "gpu_dev", "gpu_dev_m1" .. "gpu_dev_m4".
// Target: ln(ms) of the path the line was measured on (output 0 cpu, 1 gpu,
// 2 gpu with device-resident operands, 3 .. 6 one device-resident product
// cut into 1 .. 4 chunks; the model has as many outputs as the highest path
// in the logs needs).
// Prints one JSON summary: sample counts, train/holdout loss; on holdout
// shapes measured on both paths the summed time of the model's choices, of the
// threshold rule's choices and of the best choice; on holdout shapes measured
// with every allowed cut, the summed time of the model's cut, the engine's own
// cut, the uncut product and the best cut.  --holdout-out writes the holdout
// lines (the test slice).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "dispatch.h"
#include "mlp_dispatch.h"
#include "operands.h"
#include "pintt_gpu.h"
#include "pintt_mlp.h"

using namespace pi;

namespace {

struct Sample {
    ProductShape shape;
    int path;
    double ms;
    std::string line;
};

bool parse(const std::string& line, Sample& s) {
    unsigned long long an, bn, batch;
    int sq;
    char path[16];
    double ms;
    if (std::sscanf(line.c_str(), "{\"an\":%llu,\"bn\":%llu,\"batch\":%llu,\"square\":%d,\"path\":\"%15[a-z_0-9]\",\"ms\":%lf}",
                    &an, &bn, &batch, &sq, path, &ms) != 6)
        return false;
    if (an == 0 || bn == 0 || !(ms > 0)) return false;
    s.shape = ProductShape{(size_t)an, (size_t)bn, (size_t)std::max(1ull, batch), sq != 0};
    if (std::strcmp(path, "cpu") == 0) s.path = 0;
    else if (std::strcmp(path, "gpu") == 0) s.path = 1;
    else if (std::strcmp(path, "gpu_dev") == 0) s.path = 2;
    else if (std::strncmp(path, "gpu_dev_m", 9) == 0 && path[9] >= '1' && path[9] <= '4' && !path[10])
        s.path = 2 + (path[9] - '0');
    else return false;
    s.ms = ms;
    s.line = line;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    std::string out, holdout_out;
    std::vector<std::string> logs;
    int steps = 4000;
    float lr = 3e-3f;
    uint64_t seed = 1;
    double holdout = 0.2;
    unsigned threads = 16;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() { return std::string(i + 1 < argc ? argv[++i] : ""); };
        if (a == "--out") out = val();
        else if (a == "--log") logs.push_back(val());
        else if (a == "--steps") steps = std::stoi(val());
        else if (a == "--lr") lr = std::stof(val());
        else if (a == "--seed") seed = std::stoull(val());
        else if (a == "--holdout") holdout = std::stod(val());
        else if (a == "--threads") threads = std::stoul(val());
        else if (a == "--holdout-out") holdout_out = val();
        else {
            std::cerr << "unknown option " << a << std::endl;
            return 2;
        }
    }
    if (out.empty() || logs.empty()) {
        std::cerr << "usage: pi_train_mlp --out MODEL --log FILE [--log FILE ...]" << std::endl;
        return 2;
    }
    std::vector<Sample> all;
    for (const auto& path : logs) {
        std::ifstream f(path);
        if (!f) {
            std::cerr << "cannot open " << path << std::endl;
            return 1;
        }
        std::string line;
        Sample s;
        while (std::getline(f, line))
            if (parse(line, s)) all.push_back(s);
    }
    if (all.size() < 16) {
        std::cerr << "too few samples: " << all.size() << std::endl;
        return 1;
    }
    // Holdout by shape key, so that both paths of one shape land on the same side.
    SplitMix64 g(seed);
    std::map<std::tuple<size_t, size_t, size_t, bool>, bool> side;
    std::vector<Sample> train, test;
    for (const auto& s : all) {
        auto key = std::make_tuple(s.shape.an, s.shape.bn, s.shape.batch, s.shape.square);
        auto it = side.find(key);
        if (it == side.end()) it = side.emplace(key, g.uniform() < holdout).first;
        (it->second ? test : train).push_back(s);
    }

    MlpModel model;
    int K = 2;
    for (const auto& smp : all) K = std::max(K, smp.path + 1);
    model.outputs = K;
    const size_t n = train.size();
    constexpr int IN = PINTT_MLP_INPUTS;
    std::vector<MlpFeatures> feats(n);
    for (size_t i = 0; i < n; ++i) feats[i] = MlpModel::features(train[i].shape);
    for (int j = 0; j < IN; ++j) {
        double m = 0, v = 0, lo = 1e30, hi = -1e30;
        for (auto& f : feats) {
            m += f[j];
            lo = std::min(lo, (double)f[j]);
            hi = std::max(hi, (double)f[j]);
        }
        m /= n;
        for (auto& f : feats) v += (f[j] - m) * (f[j] - m);
        double sd = std::sqrt(v / n);
        model.mean[j] = (float)m;
        model.scale[j] = sd > 1e-6 ? (float)sd : 1.f;
        model.lo[j] = (float)lo - 0.25f;
        model.hi[j] = (float)hi + 0.25f;
    }
    std::vector<float> x(n * IN), t(n);
    std::vector<int> p(n);
    double tmean[PINTT_MLP_MAX_OUTPUTS] = {};
    size_t tcount[PINTT_MLP_MAX_OUTPUTS] = {};
    for (size_t i = 0; i < n; ++i) {
        auto z = model.standardise(feats[i]);
        std::copy(z.begin(), z.end(), &x[i * IN]);
        t[i] = (float)std::log(train[i].ms);
        p[i] = train[i].path;
        tmean[p[i]] += t[i];
        tcount[p[i]]++;
    }
    // He-uniform initial weights from the seeded stream, output biases at each
    // path's mean target.
    model.params.assign(gpu_mlp_param_count(K), 0.f);
    const int H = PINTT_MLP_HIDDEN;
    auto fill = [&](size_t off, size_t count, int fan_in) {
        float lim = std::sqrt(6.f / fan_in);
        for (size_t i = 0; i < count; ++i) model.params[off + i] = (float)((2 * g.uniform() - 1) * lim);
    };
    const size_t oW1 = 0, oW2 = H * IN + H, oW3 = oW2 + H * H + H, oB3 = oW3 + (size_t)K * H;
    fill(oW1, H * IN, IN);
    fill(oW2, H * H, H);
    fill(oW3, (size_t)K * H, H);
    for (int k = 0; k < K; ++k) model.params[oB3 + k] = tcount[k] ? (float)(tmean[k] / tcount[k]) : 0.f;

    std::vector<float> hist(steps);
    int rc = gpu_mlp_train(model.params.data(), K, x.data(), p.data(), t.data(), n, steps, lr, hist.data());
    if (rc != PINTT_OK) {
        std::cerr << "gpu_mlp_train failed: " << gpu_mlp_last_error() << std::endl;
        return 1;
    }
    float final_loss = 0;
    std::vector<float> grad(model.params.size());
    gpu_mlp_gradient(model.params.data(), K, x.data(), p.data(), t.data(), n, grad.data(), &final_loss);

    // Holdout: loss, and decision quality on shapes measured on both paths.
    double test_loss = 0;
    ThresholdDispatcher heur(10000);
    MlpDispatcher mlp(model, &heur, threads);
    std::map<std::tuple<size_t, size_t, size_t, bool>, std::array<double, PINTT_MLP_MAX_OUTPUTS>> pairs;
    std::map<std::tuple<size_t, size_t, size_t, bool>, std::array<int, PINTT_MLP_MAX_OUTPUTS>> pcount;
    for (const auto& s : test) {
        auto z = model.standardise(MlpModel::features(s.shape));
        float y[PINTT_MLP_MAX_OUTPUTS];
        model.forward(z.data(), y);
        double e = y[s.path] - std::log(s.ms);
        test_loss += 0.5 * e * e;
        auto key = std::make_tuple(s.shape.an, s.shape.bn, s.shape.batch, s.shape.square);
        auto& pr = pairs[key];
        auto& pc = pcount[key];
        pr[s.path] += s.ms;
        pc[s.path]++;
    }
    test_loss /= std::max<size_t>(1, test.size());
    // Host staging (path 1) and device-resident operands (path 2) are scored
    // separately, each on the shapes measured on the host and on that path.
    double t_mlp[3] = {0, 0, 0}, t_heur[3] = {0, 0, 0}, t_best[3] = {0, 0, 0};
    size_t npairs[3] = {0, 0, 0}, agree[3] = {0, 0, 0};
    for (auto& kv : pairs) {
        auto& pc = pcount[kv.first];
        for (int gp = 1; gp <= 2; ++gp) {
            if (!pc[0] || !pc[gp]) continue;
            ProductShape sh{std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first),
                            gp == 2};
            // Host time per product as seen by a batch: divided by the threads sharing it.
            double cost[2] = {kv.second[0] / pc[0] / (sh.batch > 1 ? threads : 1), kv.second[gp] / pc[gp]};
            int cm = mlp.choose(sh) == Path::Gpu, ch = heur.choose(sh) == Path::Gpu;
            t_mlp[gp] += cost[cm];
            t_heur[gp] += cost[ch];
            t_best[gp] += std::min(cost[0], cost[1]);
            agree[gp] += cm == ch;
            ++npairs[gp];
        }
    }
    // Cuts: single device-resident products measured with every allowed cut.
    double c_mlp = 0, c_engine = 0, c_uncut = 0, c_best = 0;
    size_t c_shapes = 0, c_split_chosen = 0;
    if (K >= 7)
        for (auto& kv : pairs) {
            auto& pc = pcount[kv.first];
            const size_t an = std::get<0>(kv.first), bn = std::get<1>(kv.first);
            if (std::get<2>(kv.first) != 1 || std::get<3>(kv.first) || !pc[2] || !pc[3]) continue;
            bool complete = true;
            for (int c = 1; c <= 4; ++c)
                if (gpu_ntt_chunk_plan(an, bn, c, nullptr, nullptr) && !pc[2 + c]) complete = false;
            if (!complete) continue;
            double cost[4];
            if (!mlp.chunk_costs(std::max(an, bn), std::min(an, bn), cost)) continue;
            int pick = 1;
            double best = 1e300;
            for (int c = 1; c <= 4; ++c) {
                if (!pc[2 + c]) continue;
                const double t = kv.second[2 + c] / pc[2 + c];
                best = std::min(best, t);
                if (cost[c - 1] < cost[pick - 1] || !pc[2 + pick]) pick = c;
            }
            c_mlp += kv.second[2 + pick] / pc[2 + pick];
            c_engine += kv.second[2] / pc[2];
            c_uncut += kv.second[3] / pc[3];
            c_best += best;
            c_split_chosen += pick > 1;
            ++c_shapes;
        }
    if (!holdout_out.empty()) {
        std::ofstream h(holdout_out);
        for (const auto& smp : test) h << smp.line << "\n";
        if (!h) {
            std::cerr << "cannot write " << holdout_out << std::endl;
            return 1;
        }
    }
    if (!model.save(out)) {
        std::cerr << "cannot write " << out << std::endl;
        return 1;
    }
    std::printf("{\"samples\":%zu,\"train\":%zu,\"holdout\":%zu,\"outputs\":%d,\"steps\":%d,\"loss_first\":%.5f,"
                "\"loss_last\":%.5f,\"holdout_loss\":%.5f,"
                "\"staged\":{\"pairs\":%zu,\"agree_with_threshold\":%zu,\"ms_mlp\":%.3f,\"ms_threshold\":%.3f,\"ms_best\":%.3f},"
                "\"resident\":{\"pairs\":%zu,\"agree_with_threshold\":%zu,\"ms_mlp\":%.3f,\"ms_threshold\":%.3f,\"ms_best\":%.3f},"
                "\"cuts\":{\"shapes\":%zu,\"split_chosen\":%zu,\"ms_mlp\":%.3f,\"ms_engine\":%.3f,\"ms_uncut\":%.3f,"
                "\"ms_best\":%.3f}}\n",
                all.size(), train.size(), test.size(), K, steps, steps ? hist[0] : 0.f, final_loss, test_loss,
                npairs[1], agree[1], t_mlp[1], t_heur[1], t_best[1], npairs[2], agree[2], t_mlp[2], t_heur[2], t_best[2],
                c_shapes, c_split_chosen, c_mlp, c_engine, c_uncut, c_best);
    return 0;
}
