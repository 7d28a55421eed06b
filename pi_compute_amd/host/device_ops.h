// device_ops.h - small helpers for host code that keeps numbers in device
// memory between GPU engine calls (include/pintt_gpu.h): status checks,
// allocations freed together, and a device number's pointer and length.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pintt_gpu.h"

namespace pi {
namespace dev {

inline void check(int rc, const char* what) {
    if (rc != PINTT_OK) throw std::runtime_error(std::string(what) + ": " + gpu_ntt_last_error());
}

// Device allocations freed together when the owner goes out of scope.
class Buffers {
public:
    Buffers() = default;
    ~Buffers() {
        for (uint64_t* p : blocks_) (void)gpu_ntt_dev_free(p);
    }
    Buffers(const Buffers&) = delete;
    Buffers& operator=(const Buffers&) = delete;
    uint64_t* alloc(size_t limbs) {
        uint64_t* p = nullptr;
        check(gpu_ntt_dev_alloc(std::max<size_t>(limbs, 1), &p), "gpu_ntt_dev_alloc");
        blocks_.push_back(p);
        return p;
    }

private:
    std::vector<uint64_t*> blocks_;
};

// A device number: n >= 1 limbs at d, the top ones possibly zero.
struct Num {
    uint64_t* d = nullptr;
    size_t n = 0;
};

}  // namespace dev
}  // namespace pi
