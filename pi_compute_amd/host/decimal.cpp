// decimal.cpp - see decimal.h.
#include "decimal.h"

#include <cstring>
#include <future>
#include <thread>
#include <vector>

namespace pi {
namespace {

constexpr size_t kLeafDigits = 4096;

struct Powers {
    std::vector<__mpz_struct> p;  // p[i] = 10^(2^i)
    explicit Powers(size_t max_digits) {
        size_t d = 1;
        while (d < max_digits) {
            p.emplace_back();
            mpz_init(&p.back());
            if (p.size() == 1)
                mpz_set_ui(&p.back(), 10);
            else
                mpz_mul(&p.back(), &p[p.size() - 2], &p[p.size() - 2]);
            d <<= 1;
        }
    }
    ~Powers() {
        for (auto& x : p) mpz_clear(&x);
    }
};

void leaf(mpz_srcptr n, char* slot, size_t width) {
    size_t len = mpz_sizeinbase(n, 10) + 2;
    std::vector<char> buf(len);
    mpz_get_str(buf.data(), 10, n);
    len = strlen(buf.data());
    if (len == 1 && buf[0] == '0') len = 0;
    memset(slot, '0', width - len);
    memcpy(slot + width - len, buf.data(), len);
}

// n < 10^width; writes exactly `width` characters.
void convert(mpz_srcptr n, char* slot, size_t width, const Powers& pw, int par_depth) {
    if (width <= kLeafDigits) {
        leaf(n, slot, width);
        return;
    }
    int lvl = 0;
    while (lvl + 1 < (int)pw.p.size() && ((size_t)2 << lvl) < width) ++lvl;
    const size_t low = (size_t)1 << lvl;  // low < width
    mpz_t q, r;
    mpz_init(q);
    mpz_init(r);
    mpz_tdiv_qr(q, r, n, &pw.p[lvl]);
    if (par_depth > 0) {
        auto fut = std::async(std::launch::async, [&]() { convert(r, slot + width - low, low, pw, par_depth - 1); });
        convert(q, slot, width - low, pw, par_depth - 1);
        fut.get();
    } else {
        convert(q, slot, width - low, pw, 0);
        convert(r, slot + width - low, low, pw, 0);
    }
    mpz_clear(q);
    mpz_clear(r);
}

}  // namespace

std::string to_decimal(mpz_srcptr n, unsigned threads) {
    if (mpz_sgn(n) == 0) return "0";
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    // sizeinbase may overshoot by one; convert at that width and strip.
    const size_t width = mpz_sizeinbase(n, 10);
    std::string out(width, '0');
    mpz_t a;
    mpz_init(a);
    mpz_abs(a, n);
    if (width <= kLeafDigits) {
        leaf(a, &out[0], width);
    } else {
        Powers pw(width);
        int depth = 0;
        for (unsigned t = threads; t > 1; t >>= 1) ++depth;
        convert(a, &out[0], width, pw, depth + 1);
    }
    mpz_clear(a);
    size_t z = out.find_first_not_of('0');
    if (z != 0) out.erase(0, z);
    return out;
}

}  // namespace pi
