// halton.cpp -- host construction of the Halton tables (Faure digit
// permutations) and of the per-pixel enumeration constants.
#include "halton.h"

#include <stdexcept>

namespace aur {

// Faure's permutation of {0..b-1}: identity up to 3; even b: the permutation
// of b/2 doubled, then doubled plus one; odd b: the permutation of b-1 with
// values >= (b-1)/2 moved up by one and (b-1)/2 inserted at the centre.
static const std::vector<uint16_t> &faure(std::vector<std::vector<uint16_t>> &memo, uint32_t b) {
    if (!memo[b].empty()) return memo[b];
    std::vector<uint16_t> p(b);
    if (b <= 3) {
        for (uint32_t i = 0; i < b; ++i) p[i] = (uint16_t)i;
    } else if (b % 2 == 0) {
        const std::vector<uint16_t> h = faure(memo, b / 2);
        for (uint32_t i = 0; i < b / 2; ++i) {
            p[i] = (uint16_t)(2 * h[i]);
            p[b / 2 + i] = (uint16_t)(2 * h[i] + 1);
        }
    } else {
        const uint32_t c = b / 2;
        const std::vector<uint16_t> q = faure(memo, b - 1);
        for (uint32_t i = 0; i + 1 < b; ++i) p[i < c ? i : i + 1] = (uint16_t)(q[i] + (q[i] >= c ? 1 : 0));
        p[c] = (uint16_t)c;
    }
    memo[b] = p;
    return memo[b];
}

void halton_tables(std::vector<DHaltonDim> *dims, std::vector<uint16_t> *tables) {
    dims->assign(AUR_HALTON_DIMS, DHaltonDim());
    tables->clear();
    std::vector<std::vector<uint16_t>> memo(2048);
    uint32_t b = 1;
    for (int d = 0; d < AUR_HALTON_DIMS; ++d) {
        for (;;) {  // next prime
            ++b;
            bool prime = true;
            for (uint32_t k = 2; k * k <= b; ++k)
                if (b % k == 0) { prime = false; break; }
            if (prime) break;
        }
        DHaltonDim &D = (*dims)[d];
        D = DHaltonDim();
        if (d == 0) continue;  // base 2 is a bit reversal
        uint32_t g = 1, block = b;
        while (block * b <= 500) { block *= b; ++g; }
        uint64_t maxp = block;
        uint32_t n = 1;
        while (maxp * block < (1ull << 32)) { maxp *= block; ++n; }
        D.block = block;
        D.n_blocks = n;
        D.top = (uint32_t)(maxp / block);
        D.off = (uint32_t)tables->size();
        D.scale = (float)(0x1.fffffcp-1 / (double)maxp);
        const std::vector<uint16_t> &perm = faure(memo, b);
        for (uint32_t i = 0; i < block; ++i) {
            uint32_t v = 0, r = i;
            for (uint32_t k = 0; k < g; ++k) {
                v = v * b + perm[r % b];
                r /= b;
            }
            tables->push_back((uint16_t)v);
        }
    }
}

void halton_enum(int width, int height, HaltonView *v) {
    if (width <= 0 || height <= 0) throw std::invalid_argument("bad resolution");
    uint32_t p2 = 0, pw = 1, p3 = 0, ph = 1;
    while (pw < (uint32_t)width) { pw *= 2; ++p2; }
    while (ph < (uint32_t)height) { ph *= 3; ++p3; }
    if ((uint64_t)pw * ph > 0xffffffffull) throw std::invalid_argument("resolution too large for the Halton enumeration");
    // multiplicative inverses by the extended Euclidean algorithm
    auto inverse = [](int64_t a, int64_t m) -> uint32_t {
        if (m == 1) return 0;
        int64_t r0 = a % m, r1 = m, s0 = 1, s1 = 0;
        while (r1) {
            int64_t q = r0 / r1, t = r0 - q * r1;
            r0 = r1; r1 = t;
            t = s0 - q * s1; s0 = s1; s1 = t;
        }
        return (uint32_t)((s0 % m + m) % m);
    };
    v->width = width;
    v->p2 = p2;
    v->p3 = p3;
    v->inc = pw * ph;
    v->mx = ph * inverse(ph, pw);
    v->my = pw * inverse(pw, ph);
    v->sx = (float)pw;
    v->sy = (float)ph;
}

}  // namespace aur
