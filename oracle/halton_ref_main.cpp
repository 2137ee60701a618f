// halton_ref_main.cpp -- TEST INFRASTRUCTURE: a driver that prints values of
// the reference's vendored Halton library (Aurora/external/halton, header
// only) so the oracle's restatement can be checked against it.  Built by
// `make -C oracle ref` into oracle/_ref/; never part of the product.
//
// Output (text, one record per line):
//   S <dim> <index> <float bits>        Halton_sampler::sample after init_faure
//   R <w> <h> <increment> <scale_x bits> <scale_y bits>
//   E <w> <h> <i> <x> <y> <index>       Halton_enum::get_index
#include <cstdio>
#include <cstring>

#include "halton_enum.h"
#include "halton_sampler.h"

static unsigned bits(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}

int main() {
    static Halton_sampler hs;  // large: keep off the stack
    hs.init_faure();
    const unsigned idx[] = {0u, 1u, 2u, 3u, 4u, 5u, 6u, 7u, 8u, 9u, 10u, 11u, 12u, 13u, 17u, 100u, 242u, 243u, 244u,
                            500u, 1000u, 4095u, 65535u, 65536u, 123457u, 1048579u, 14348906u, 14348907u, 244140624u,
                            244140625u, 123456789u, 1000000007u, 2147483647u, 2147483648u, 3404825446u, 3404825447u,
                            3486784400u, 3486784401u, 4000000000u, 4294967295u};
    for (unsigned d = 0; d < Halton_sampler::get_num_dimensions(); ++d)
        for (unsigned k = 0; k < sizeof(idx) / sizeof(idx[0]); ++k)
            std::printf("S %u %u %u\n", d, idx[k], bits(hs.sample(d, idx[k])));
    const unsigned res[][2] = {{1, 1}, {7, 5}, {64, 36}, {48, 48}, {256, 256}, {1920, 1080}, {1000, 1}};
    for (const auto &r : res) {
        Halton_enum e(r[0], r[1]);
        const unsigned inc = (unsigned)(e.scale_x(1.f) * e.scale_y(1.f));  // 2^p2 * 3^p3, exact here
        std::printf("R %u %u %u %u %u\n", r[0], r[1], inc, bits(e.scale_x(1.f)), bits(e.scale_y(1.f)));
        const unsigned xs[] = {0u, 1u, r[0] / 3, r[0] / 2, r[0] - 1};
        const unsigned ys[] = {0u, r[1] / 2, r[1] - 1};
        for (unsigned i = 0; i < 5; ++i)
            for (unsigned x : xs)
                for (unsigned y : ys) std::printf("E %u %u %u %u %u %u\n", r[0], r[1], i, x, y, e.get_index(i, x, y));
    }
    return 0;
}
