// main.cpp - pi_compute_hip: Chudnovsky pi with the gfx950 multiplier.
// Options mirror the reference CLI (src/main.cpp:28-55) where they overlap.
//
//   pi_compute_hip --digits N [--output FILE] [--gpu] [--gpu-threshold L]
//                  [--threads T] [--tree levels|recursive] [--per-call]
//                  [--dispatch threshold|mlp] [--model FILE] [--cuts dispatcher|engine] [--log FILE]
//                  [--verbose]
//
//   --gpu            HipNttMultiplier (default: GmpMultiplier, the GMP-only run)
//   --tree levels    level-synchronous tree, one multiply_batch per level (default)
//   --tree recursive depth-first tree of the reference, one multiply per product
//   --per-call       with --tree levels: device products one engine call each
//   --dispatch mlp   learned cost model from --model (threshold rule otherwise)
//   --cuts engine    chunk counts of unbalanced products from the engine's tuned
//                    costs (default "dispatcher": the learned model prices them)
//   --log FILE       per-product timing lines (training data for the cost model)
//   --final mpf|newton        final quotient and square root: GMP mpf (the
//                             reference's sequence) or Newton through the
//                             multiplier (default: newton with --gpu)
//   --convert host|multiplier decimal conversion (default: multiplier with --gpu)
//   --plans FILE              tile plans: loaded if FILE exists, saved after the run
//   --convert-leaf D          multiplier conversion: host conversion below D digits (32768;
//                             device-resident conversion levels: device leaves of at most 4096)
//   --bs-segments S           device levels: S segments, host subtrees of the next one computed
//                             while the device reduces the current one (8; 1 = one tree)
//   --bs-leaf T               terms per host subtree at the bottom of the tree (128 with
//                             device levels, else 256)
//   --no-overlap              Newton final step: compute sqrt(10005) and 10^D after binary
//                             splitting instead of beside it
//   --host-subtrees           with --device-levels: the bottom subtrees of binary splitting on
//                             host threads (GMP) instead of the device
//   --host-final-step         with --device-levels: the Newton final step on host numbers
//                             (products staged through PCIe) instead of device-resident ones
//   --host-conversion-levels  with --device-levels: the multiplier conversion's levels on
//                             host numbers (products staged through PCIe) instead of on
//                             device-resident ones
//   --device-levels [L|auto]  with --gpu --tree levels: P, Q, R stay on the GPU from
//                             the first level whose operands have >= L limbs (default 0),
//                             or (auto) whose products the dispatcher sends to the GPU
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>

#include "gmp_multiplier.h"
#include "hip_ntt_multiplier.h"
#include "mlp_dispatch.h"
#include "pi_engine.h"
#include "pintt_gpu.h"

namespace {

// pi_compute_amd/data, found from this executable's path (pi_compute_amd/bin/).
std::string data_dir() {
    char buf[4096];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
    if (n <= 0) return "data";
    std::string exe(buf, (size_t)n);
    const size_t slash = exe.rfind('/');
    const std::string bin = slash == std::string::npos ? "." : exe.substr(0, slash);
    return bin + "/../data";
}

void usage() {
    std::cout << "usage: pi_compute_hip --digits N [--output FILE] [--gpu] [--gpu-threshold L]\n"
                 "                      [--threads T] [--tree levels|recursive] [--per-call]\n"
                 "                      [--dispatch threshold|mlp] [--model FILE] [--cuts dispatcher|engine]\n"
                 "                      [--log FILE]\n"
                 "                      [--final mpf|newton] [--convert host|multiplier] [--convert-leaf D]\n"
                 "                      [--device-levels [LIMBS|auto]] [--bs-segments S] [--bs-leaf T] [--no-overlap]\n"
                 "                      [--host-subtrees] [--host-final-step] [--host-conversion-levels]\n"
                 "                      [--plans FILE]\n"
                 "                      [--verbose]\n";
}

}  // namespace

int main(int argc, char** argv) {
    pi::PiConfig cfg;
    cfg.output_file = "pi_digits.txt";
    bool gpu = false, per_call = false, device_auto = false;
    size_t threshold = 10000;
    std::string dispatch = "threshold", model, log, final_step, convert, plans, cuts = "dispatcher";
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::cerr << "missing value for " << a << std::endl;
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--digits") cfg.digits = std::stoul(next());
        else if (a == "--output") cfg.output_file = next();
        else if (a == "--gpu") gpu = true;
        else if (a == "--gpu-threshold") threshold = std::stoul(next());
        else if (a == "--threads") cfg.num_threads = std::stoul(next());
        else if (a == "--tree") {
            std::string t = next();
            if (t == "levels") cfg.level_batches = true;
            else if (t == "recursive") cfg.level_batches = false;
            else { usage(); return 2; }
        } else if (a == "--per-call") per_call = true;
        else if (a == "--dispatch") dispatch = next();
        else if (a == "--model") model = next();
        else if (a == "--cuts") cuts = next();
        else if (a == "--log") log = next();
        else if (a == "--final") final_step = next();
        else if (a == "--convert") convert = next();
        else if (a == "--plans") plans = next();
        else if (a == "--convert-leaf") cfg.conversion_leaf_digits = std::stoul(next());
        else if (a == "--bs-segments") cfg.bs_segments = (unsigned)std::stoul(next());
        else if (a == "--no-overlap") cfg.overlap_constants = false;
        else if (a == "--host-conversion-levels") cfg.device_conversion = false;
        else if (a == "--host-final-step") cfg.device_final = false;
        else if (a == "--host-subtrees") cfg.device_leaves = false;
        else if (a == "--bs-leaf") cfg.bs_leaf_span = std::stoul(next());
        else if (a == "--device-levels") {
            cfg.device_levels = true;
            if (i + 1 < argc && argv[i + 1][0] != '-') {
                std::string v = argv[++i];
                if (v == "auto") device_auto = true;
                else cfg.device_from_limbs = std::stoul(v);
            }
        }
        else if (a == "--verbose" || a == "-v") cfg.verbose = true;
        else if (a == "--help" || a == "-h") { usage(); return 0; }
        else { std::cerr << "unknown option " << a << std::endl; usage(); return 2; }
    }
    if (final_step.empty()) final_step = gpu ? "newton" : "mpf";
    if (convert.empty()) convert = gpu ? "multiplier" : "host";
    if ((final_step != "mpf" && final_step != "newton") || (convert != "host" && convert != "multiplier")) {
        usage();
        return 2;
    }
    cfg.final_step = final_step == "newton" ? pi::PiConfig::FinalStep::Newton : pi::PiConfig::FinalStep::Mpf;
    cfg.conversion = convert == "multiplier" ? pi::PiConfig::Conversion::Multiplier : pi::PiConfig::Conversion::Host;
    if (cfg.device_levels && !gpu) {
        std::cerr << "--device-levels needs --gpu" << std::endl;
        return 2;
    }
    if (cfg.digits == 0) {
        std::cerr << "--digits must be > 0" << std::endl;
        return 2;
    }
    try {
        std::unique_ptr<pi::Multiplier> m;
        pi::HipNttMultiplier* hip = nullptr;
        std::unique_ptr<pi::MlpDispatcher> mlp;
        std::string mode = "GMP";
        if (gpu) {
            auto h = std::make_unique<pi::HipNttMultiplier>(threshold);
            hip = h.get();
            // Tile plans tuned on this GPU model ship in data/; a run tunes only lengths they miss.
            const std::string shipped = data_dir() + "/plans_gfx950.txt";
            if (plans.empty()) {
                std::FILE* sf = std::fopen(shipped.c_str(), "r");
                if (sf) {
                    std::fclose(sf);
                    (void)gpu_ntt_load_plans(shipped.c_str());
                }
            } else {
                std::FILE* pf = std::fopen(plans.c_str(), "r");
                if (pf) {
                    std::fclose(pf);
                    if (gpu_ntt_load_plans(plans.c_str()) != PINTT_OK)
                        std::cerr << "plan file: " << gpu_ntt_last_error() << std::endl;
                }
            }
            hip->set_batching(!per_call);
            if (dispatch == "mlp") {
                if (model.empty()) throw std::runtime_error("--dispatch mlp needs --model FILE");
                mlp = std::make_unique<pi::MlpDispatcher>(model, &hip->dispatcher(), cfg.num_threads);
                hip->set_dispatcher(mlp.get());
            } else if (dispatch != "threshold") {
                throw std::runtime_error("unknown dispatcher " + dispatch);
            }
            if (cuts != "dispatcher" && cuts != "engine") throw std::runtime_error("unknown --cuts " + cuts);
            hip->set_dispatcher_cuts(cuts == "dispatcher");
            if (!log.empty() && !hip->open_log(log)) throw std::runtime_error("cannot open log " + log);
            if (device_auto) cfg.device_switch = &hip->dispatcher();
            mode = std::string("HIP NTT, ") + hip->device_name() + ", " + hip->dispatcher().name() +
                   " dispatch, " + (cfg.level_batches ? (per_call ? "levels/per-call" : "levels/batched") : "recursive") +
                   (cfg.device_levels ? (device_auto ? ", device-resident levels (dispatcher switch)" : ", device-resident levels")
                                      : "") +
                   ", final " + final_step + ", conversion " + convert;
            m = std::move(h);
        } else {
            m = std::make_unique<pi::GmpMultiplier>();
        }
        pi::PiEngine engine(*m);
        pi::PiResult r = engine.compute(cfg);
        if (hip) {
            if (!plans.empty() && gpu_ntt_save_plans(plans.c_str()) != PINTT_OK)
                std::cerr << "plan file: " << gpu_ntt_last_error() << std::endl;
            hip->close_log();
            if (cfg.verbose) hip->print_stats();
        }
        std::cout << "Computed " << cfg.digits << " digits of pi in " << r.elapsed_seconds << " seconds ("
                  << r.terms_used << " terms, " << mode << ")." << std::endl;
        std::cout << "{\"digits\":" << cfg.digits << ",\"seconds\":" << r.elapsed_seconds
                  << ",\"bs_seconds\":" << r.bs_seconds << ",\"final_seconds\":" << r.final_seconds
                  << ",\"conv_seconds\":" << r.conv_seconds;
        if (hip)
            std::cout << ",\"gpu_calls\":" << hip->stats().gpu_calls.load()
                      << ",\"gpu_batches\":" << hip->stats().gpu_batches.load()
                      << ",\"cpu_calls\":" << hip->stats().cpu_calls.load()
                      << ",\"fallback_calls\":" << hip->stats().fallback_calls.load();
        std::cout << "}" << std::endl;
        std::cout << "Output written to: " << cfg.output_file << std::endl;
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
}
