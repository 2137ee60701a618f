/*
 * oracle_halton.c -- CPU restatement of the Halton QMC sampler the reference
 * vendors (Aurora/external/halton, L. Gruenschloss, MIT licence; used by
 * camera.cpp:39-45, lambert.h:45-52, cookTorrance.cpp:50-57 and
 * squareLight.h:58-65 when USE_HALTON is defined, constants.h:31-41).
 *
 * TEST INFRASTRUCTURE ONLY (the checker).
 *
 * The published construction:
 *   - dimension d uses the d-th prime b as its base; base 2 is a plain bit
 *     reversal written into a float mantissa;
 *   - other bases use Faure's digit permutations (identity for b <= 3; for
 *     even b the permutation of b/2 doubled, then doubled plus one; for odd b
 *     the permutation of b-1 with every value >= (b-1)/2 shifted up by one and
 *     (b-1)/2 inserted in the middle);
 *   - digits are processed in groups of g, the largest group with b^g <= 500:
 *     a table maps every g-digit block to its permuted, reversed value, and the
 *     radical inverse sums table[block_k] * (b^g)^(n-1-k) over the n blocks of
 *     the largest power (b^g)^n that stays below 2^32, in 32-bit unsigned
 *     arithmetic, then multiplies the float of that sum by
 *     float((1 - 2^-22) / (b^g)^n);
 *   - Halton_enum picks, for pixel (x, y) of a W x H image, the index of the
 *     i-th point whose first two coordinates fall in that pixel: with
 *     2^p2 >= W and 3^p3 >= H minimal, index = (rev2(x) * m_x + rev3(y) * m_y)
 *     mod (2^p2 3^p3) + i * 2^p2 3^p3 (Chinese remainder theorem), and the
 *     camera coordinates are the first two components scaled by 2^p2 and 3^p3.
 * Parity: bit-exact against the library itself (oracle/_ref/halton_ref and
 * the fixture tests/golden/halton_golden.npz it wrote).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "aurora_oracle.h"

#define N_DIMS 256
#define MAX_BASE 1619

typedef struct {
    unsigned base, block;   /* b and b^g */
    unsigned n_blocks;      /* n */
    unsigned top;           /* (b^g)^(n-1) */
    float scale;
    unsigned short *table;  /* block -> permuted reversed value */
} Dim;

static Dim g_dims[N_DIMS];
static int g_ready = 0;

static unsigned short *faure_perm(unsigned short **perms, unsigned b) {
    if (perms[b]) return perms[b];
    unsigned short *p = (unsigned short *)malloc(sizeof(unsigned short) * b);
    if (b <= 3) {
        for (unsigned i = 0; i < b; ++i) p[i] = (unsigned short)i;
    } else if (b % 2 == 0) {
        const unsigned short *h = faure_perm(perms, b / 2);
        for (unsigned i = 0; i < b / 2; ++i) {
            p[i] = (unsigned short)(2 * h[i]);
            p[b / 2 + i] = (unsigned short)(2 * h[i] + 1);
        }
    } else {
        const unsigned c = b / 2;
        const unsigned short *q = faure_perm(perms, b - 1);
        for (unsigned i = 0; i < b - 1; ++i) {
            unsigned v = q[i] + (q[i] >= c ? 1u : 0u);
            p[i < c ? i : i + 1] = (unsigned short)v;
        }
        p[c] = (unsigned short)c;
    }
    perms[b] = p;
    return p;
}

static void halton_init(void) {
    if (g_ready) return;
    unsigned short **perms = (unsigned short **)calloc(MAX_BASE + 1, sizeof(unsigned short *));
    unsigned b = 1;
    for (int d = 0; d < N_DIMS; ++d) {
        /* next prime */
        for (;;) {
            ++b;
            int prime = 1;
            for (unsigned k = 2; k * k <= b; ++k)
                if (b % k == 0) { prime = 0; break; }
            if (prime) break;
        }
        Dim *D = &g_dims[d];
        memset(D, 0, sizeof *D);
        D->base = b;
        if (d == 0) continue;  /* base 2: bit reversal */
        unsigned g = 1, block = b;
        while (block * b <= 500) { block *= b; ++g; }
        uint64_t maxp = block;
        unsigned n = 1;
        while (maxp * block < ((uint64_t)1 << 32)) { maxp *= block; ++n; }
        D->block = block;
        D->n_blocks = n;
        D->top = (unsigned)(maxp / block);
        D->scale = (float)(0x1.fffffcp-1 / (double)maxp);
        const unsigned short *perm = faure_perm(perms, b);
        D->table = (unsigned short *)malloc(sizeof(unsigned short) * block);
        for (unsigned i = 0; i < block; ++i) {
            /* the g digits of i, least significant first, permuted, read back
             * most significant first */
            unsigned v = 0, r = i;
            for (unsigned k = 0; k < g; ++k) {
                v = v * b + perm[r % b];
                r /= b;
            }
            D->table[i] = (unsigned short)v;
        }
    }
    for (unsigned k = 0; k <= MAX_BASE; ++k) free(perms[k]);
    free(perms);
    g_ready = 1;
}

static unsigned reverse_bits(unsigned v) {
    unsigned r = 0;
    for (int k = 0; k < 32; ++k) {
        r = (r << 1) | (v & 1u);
        v >>= 1;
    }
    return r;
}

float orc_halton_sample(unsigned dim, unsigned index) {
    halton_init();
    if (dim >= N_DIMS) return 0.f;
    if (dim == 0) {
        union { unsigned u; float f; } r;
        r.u = 0x3f800000u | (reverse_bits(index) >> 9);
        return r.f - 1.f;
    }
    const Dim *D = &g_dims[dim];
    unsigned sum = 0, p = D->top, r = index;
    for (unsigned k = 0; k < D->n_blocks; ++k) {
        sum += (unsigned)D->table[r % D->block] * p;
        r /= D->block;
        p /= D->block;
    }
    return (float)sum * D->scale;
}

/* Halton_enum for a W x H image */
void orc_halton_enum(unsigned w, unsigned h, unsigned *p2, unsigned *p3, unsigned *mx, unsigned *my, unsigned *inc) {
    unsigned a = 0, pw = 1, b = 0, ph = 1;
    while (pw < w) { pw *= 2; ++a; }
    while (ph < h) { ph *= 3; ++b; }
    /* m_x = ph * (ph^-1 mod pw), m_y = pw * (pw^-1 mod ph), found by search */
    unsigned ix = 0, iy = 0;
    if (pw > 1)
        for (unsigned k = 1; k < pw; ++k)
            if ((uint64_t)ph * k % pw == 1) { ix = k; break; }
    if (ph > 1)
        for (unsigned k = 1; k < ph; ++k)
            if ((uint64_t)pw * k % ph == 1) { iy = k; break; }
    *p2 = a; *p3 = b;
    *mx = ph * ix;
    *my = pw * iy;
    *inc = pw * ph;
}

unsigned orc_halton_index(unsigned w, unsigned h, unsigned i, unsigned x, unsigned y) {
    unsigned p2, p3, mx, my, inc;
    orc_halton_enum(w, h, &p2, &p3, &mx, &my, &inc);
    unsigned hx = p2 ? reverse_bits(x) >> (32 - p2) : 0u;
    unsigned hy = 0, r = y;
    for (unsigned k = 0; k < p3; ++k) {
        hy = hy * 3 + r % 3;
        r /= 3;
    }
    const unsigned off = (unsigned)(((uint64_t)hx * mx + (uint64_t)hy * my) % inc);
    return off + i * inc;
}

/* The value of sample slot `slot` of the path through pixel (x, y), sample
 * (pass) i, under the Halton source: slots 0 and 1 are the camera's offsets
 * within the pixel (camera.cpp:39-45: scale_x(sample(0, index)) - x), slots 2
 * to 255 are dimensions 2 to 255 at the same index. */
float orc_halton_slot(unsigned w, unsigned h, unsigned i, unsigned x, unsigned y, unsigned slot) {
    unsigned p2, p3, mx, my, inc;
    orc_halton_enum(w, h, &p2, &p3, &mx, &my, &inc);
    const unsigned index = orc_halton_index(w, h, i, x, y);
    const float v = orc_halton_sample(slot, index);
    if (slot == 0) return v * (float)(1u << p2) - (float)x;
    if (slot == 1) {
        unsigned ph = 1;
        for (unsigned k = 0; k < p3; ++k) ph *= 3;
        return v * (float)ph - (float)y;
    }
    return v;
}
