// pe_sobol.h — Burley-2020 hash-Owen-scrambled, index-shuffled 2D Sobol values as pure integer ALU.
//
// Same values, bit for bit, as SobolSampler (samplers/sobol_sampler.cpp:92-146): for dimension d of
// sample i in pixel (x, y)
//   pixelHash = hc(hc(mix(seedLo ^ mix(seedHi)), x), y)              (startPixel, :133-140)
//   s         = hc(pixelHash, d)
//   index     = rev(LK(rev(i), s))                                   (get2D, :111-117)
//   x         = rev(LK(index, hc(s, 0)))
//   y         = rev(LK(revSobol1(index), hc(s, 1)))
//   value     = (u32 >> 8) * 2^-24
// The reference looks revSobol1 up in four 256-entry byte tables (:44-56). On gfx950 it is computed
// in registers instead: bit j of rev(sobol1(i)) is the XOR of the index bits k whose binary digits
// contain j's (the Pascal generator matrix mod 2 is the subset relation, Lucas' theorem), i.e. the
// superset-sum transform of the 32 index bits, five shift/mask/xor steps. Bit reversal is one
// v_bfrev_b32. No memory traffic at all.
#pragma once

#include "pe_math.h"

namespace pe {

PE_HD uint32_t rev_bits(uint32_t x) { return __builtin_bitreverse32(x); }

// Laine–Karras permutation with Burley's constants (sobol_sampler.cpp:62-69).
PE_HD uint32_t laine_karras(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

// lowbias32 (sobol_sampler.cpp:72-79).
PE_HD uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// boost::hash_combine pattern followed by mix32 (sobol_sampler.cpp:83-85).
PE_HD uint32_t hash_combine(uint32_t h, uint32_t v) { return mix32(h ^ (v + 0x9e3779b9u + (h << 6) + (h >> 2))); }

// reverseBits(sobolDim1(index)) without tables (see the file comment).
PE_HD uint32_t rev_sobol1(uint32_t x) {
    x ^= (x >> 1) & 0x55555555u;
    x ^= (x >> 2) & 0x33333333u;
    x ^= (x >> 4) & 0x0F0F0F0Fu;
    x ^= (x >> 8) & 0x00FF00FFu;
    x ^= (x >> 16) & 0x0000FFFFu;
    return x;
}

PE_HD float sobol_unit(uint32_t x) { return static_cast<float>(x >> 8) * 0x1p-24f; }

// startPixel(x, y) (sobol_sampler.cpp:133-140).
PE_HD uint32_t sobol_pixel_hash(uint64_t seed, int32_t px, int32_t py) {
    uint32_t h = mix32(static_cast<uint32_t>(seed) ^ mix32(static_cast<uint32_t>(seed >> 32)));
    h = hash_combine(h, static_cast<uint32_t>(px));
    return hash_combine(h, static_cast<uint32_t>(py));
}

// get2D for dimension `dim` of the sample whose reversed index is rev_index; returns both u32 patterns.
PE_HD void sobol_2d_bits(uint32_t pixel_hash, uint32_t rev_index, uint32_t dim, uint32_t& xb, uint32_t& yb) {
    const uint32_t s = hash_combine(pixel_hash, dim);
    const uint32_t index = rev_bits(laine_karras(rev_index, s));
    xb = rev_bits(laine_karras(index, hash_combine(s, 0u)));
    yb = rev_bits(laine_karras(rev_sobol1(index), hash_combine(s, 1u)));
}

// get1D: get2D().x without computing y (sobol_sampler.cpp:120-125).
PE_HD uint32_t sobol_1d_bits(uint32_t pixel_hash, uint32_t rev_index, uint32_t dim) {
    const uint32_t s = hash_combine(pixel_hash, dim);
    const uint32_t index = rev_bits(laine_karras(rev_index, s));
    return rev_bits(laine_karras(index, hash_combine(s, 0u)));
}

// A path's sampler state: what SobolSampler keeps between get calls, in three registers.
struct SobolState {
    uint32_t pixel_hash;
    uint32_t rev_index;
    uint32_t dim;
};

PE_HD float sobol_get1d(SobolState& st) { return sobol_unit(sobol_1d_bits(st.pixel_hash, st.rev_index, st.dim++)); }
PE_HD void sobol_get2d(SobolState& st, float& u, float& v) {
    uint32_t xb, yb;
    sobol_2d_bits(st.pixel_hash, st.rev_index, st.dim++, xb, yb);
    u = sobol_unit(xb);
    v = sobol_unit(yb);
}

}  // namespace pe
