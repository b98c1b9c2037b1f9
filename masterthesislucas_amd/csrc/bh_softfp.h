// Exact integer emulation of the FP64 arithmetic the reference uses to build Morton codes.
//
// The reference quantises a position with double arithmetic (nBodySolverFMM_SYM_MORTON.h:538-552):
//     t = RN64(RN64(p + (-min)) / L);   q = (uint32) min(max(t * 65536, 0), 65535)
// and sizes the QuadTree root with double adds (quadtree.h:45-47). To reproduce those codes bit
// for bit on the GPU without FP64 arithmetic on the per-particle path, this header restates the
// two IEEE-754 binary64 operations involved (round-to-nearest-even addition, and
// floor(2^B * RN64(d / S))) with 64-bit integer arithmetic. The one exception is the quantiser's
// preparation (bh_sfp_quant_prepare, once per build by one thread): it starts from an FP64
// estimate of the reciprocal of S's significand and makes it exact with a 128-bit integer
// remainder, so its result does not depend on how that estimate rounds. Compiles as host C++ and
// as HIP device code; the host build is checked against native double arithmetic on millions of
// random and boundary cases, and the device-prepared quantiser against exact integer division.
//
// Values are carried as raw binary64 bit patterns (uint64_t). Inputs are finite; results of the
// additions used here are zero or normal numbers.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define BH_SFP_HD __host__ __device__ __forceinline__
#else
#define BH_SFP_HD static inline
#endif

#define BH_SFP_SIGN 0x8000000000000000ull
#define BH_SFP_FRAC 0x000FFFFFFFFFFFFFull
#define BH_SFP_HIDDEN 0x0010000000000000ull

BH_SFP_HD int bh_sfp_clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

BH_SFP_HD uint32_t bh_sfp_f32_bits(float f) {
    union {
        float f;
        uint32_t u;
    } v;
    v.f = f;
    return v.u;
}

BH_SFP_HD uint64_t bh_sfp_f64_bits(double d) {
    union {
        double d;
        uint64_t u;
    } v;
    v.d = d;
    return v.u;
}

// binary32 -> binary64 bit pattern (exact; binary32 subnormals become binary64 normals)
BH_SFP_HD uint64_t bh_sfp_from_f32(float f) {
    const uint32_t u = bh_sfp_f32_bits(f);
    const uint64_t sign = (uint64_t)(u >> 31) << 63;
    const uint32_t e = (u >> 23) & 0xFFu;
    const uint32_t m = u & 0x7FFFFFu;
    if (e == 0) {
        if (m == 0) return sign;
        const int p = 31 - __builtin_clz(m);  // value = m * 2^-149 = 1.f * 2^(p-149)
        const uint64_t frac = ((uint64_t)m << (52 - p)) & BH_SFP_FRAC;
        return sign | ((uint64_t)(p - 149 + 1023) << 52) | frac;
    }
    if (e == 0xFFu) return sign | (0x7FFull << 52) | ((uint64_t)m << 29);
    return sign | ((uint64_t)(e - 127 + 1023) << 52) | ((uint64_t)m << 29);
}

// RN64(a + b) for finite a, b whose sum is zero or a normal number.
BH_SFP_HD uint64_t bh_sfp_add(uint64_t a, uint64_t b) {
    if ((a & ~BH_SFP_SIGN) == 0) return ((b & ~BH_SFP_SIGN) == 0) ? (a & b) : b;
    if ((b & ~BH_SFP_SIGN) == 0) return a;
    int ea = (int)((a >> 52) & 0x7FF), eb = (int)((b >> 52) & 0x7FF);
    uint64_t ma = (a & BH_SFP_FRAC) | (ea ? BH_SFP_HIDDEN : 0);
    uint64_t mb = (b & BH_SFP_FRAC) | (eb ? BH_SFP_HIDDEN : 0);
    if (!ea) ea = 1;
    if (!eb) eb = 1;
    uint64_t sa = a & BH_SFP_SIGN, sb = b & BH_SFP_SIGN;
    if (ea < eb || (ea == eb && ma < mb)) {  // make |a| >= |b|
        int te = ea; ea = eb; eb = te;
        uint64_t tm = ma; ma = mb; mb = tm;
        uint64_t ts = sa; sa = sb; sb = ts;
    }
    uint64_t A = ma << 10;  // leading bit at position 62; ten guard bits below the 53 kept
    uint64_t B = mb << 10;
    const int d = ea - eb;
    if (d >= 64) {
        B = B ? 1u : 0u;
    } else if (d > 0) {
        const uint64_t sticky = (B & ((1ull << d) - 1)) ? 1u : 0u;
        B = (B >> d) | sticky;  // jam the shifted-out bits into the lowest bit
    }
    uint64_t S;
    if (sa == sb) {
        S = A + B;
    } else {
        S = A - B;
        if (S == 0) return 0;  // x - x = +0 under round-to-nearest
    }
    const int msb = 63 - bh_sfp_clz64(S);
    uint64_t keep;
    int sh;
    if (msb >= 52) {
        sh = msb - 52;
        keep = S >> sh;
        if (sh > 0) {
            const uint64_t rem = S & ((1ull << sh) - 1);
            const uint64_t half = 1ull << (sh - 1);
            if (rem > half || (rem == half && (keep & 1u))) ++keep;
            if (keep >> 53) {
                keep >>= 1;
                ++sh;
            }
        }
    } else {
        sh = -(52 - msb);
        keep = S << (52 - msb);
    }
    const int biased = ea + sh - 10;
    return sa | ((uint64_t)biased << 52) | (keep & BH_SFP_FRAC);
}

BH_SFP_HD uint64_t bh_sfp_neg(uint64_t a) { return a ^ BH_SFP_SIGN; }

// binary64 -> binary32 with round-to-nearest-even, for values inside the normal float range.
BH_SFP_HD float bh_sfp_to_f32(uint64_t a) {
    union {
        uint32_t u;
        float f;
    } v;
    const uint32_t sign = (uint32_t)(a >> 63) << 31;
    const int e = (int)((a >> 52) & 0x7FF);
    if (e == 0) {
        v.u = sign;
        return v.f;
    }
    uint64_t m = (a & BH_SFP_FRAC) | BH_SFP_HIDDEN;  // 53 bits
    uint64_t keep = m >> 29;                          // 24 bits
    const uint64_t rem = m & ((1ull << 29) - 1);
    const uint64_t half = 1ull << 28;
    int ef = e - 1023 + 127;
    if (rem > half || (rem == half && (keep & 1u))) ++keep;
    if (keep >> 24) {
        keep >>= 1;
        ++ef;
    }
    v.u = sign | ((uint32_t)ef << 23) | ((uint32_t)keep & 0x7FFFFFu);
    return v.f;
}

// max of two non-negative doubles, with std::max's rule (a < b ? b : a)
BH_SFP_HD uint64_t bh_sfp_max_nonneg(uint64_t a, uint64_t b) { return (a < b) ? b : a; }

// Quantiser for a fixed positive normal divisor S = M * 2^es, M in [2^52, 2^53).
struct bh_sfp_quant {
    uint64_t M;    // significand of S with the hidden bit
    int32_t es;    // S = M * 2^es
    int32_t pad;
    uint64_t R;    // floor((2^116 - 1) / M), in [2^63, 2^64)
};

BH_SFP_HD bh_sfp_quant bh_sfp_quant_prepare(uint64_t s_bits) {
    bh_sfp_quant q;
    const int e = (int)((s_bits >> 52) & 0x7FF);
    q.M = (s_bits & BH_SFP_FRAC) | BH_SFP_HIDDEN;
    q.es = e - 1075;
    q.pad = 0;
    if (q.M == BH_SFP_HIDDEN) {  // (2^116 - 1) / 2^52 = 2^64 - 2^-52
        q.R = ~0ull;
        return q;
    }
    // An FP64 quotient is within 2^12 of the true one (relative error 2^-53, quotient below
    // 2^64 since M > 2^52); the exact remainder N - R M (|.| < 2^66, two's complement in 128
    // bits) then gives a second FP64 correction, and at most a step or two finishes it.
    const __int128 M = (__int128)q.M;
    const __int128 N = ((__int128)1 << 116) - 1;
    uint64_t r = (uint64_t)(0x1p116 / (double)q.M);
    __int128 rem = N - (__int128)r * M;
    const int64_t adj = (int64_t)floor((double)(int64_t)(rem >> 11) * 2048.0 / (double)q.M);
    r += (uint64_t)adj;
    rem -= (__int128)adj * M;
    while (rem < 0) {
        --r;
        rem += M;
    }
    while (rem >= M) {
        ++r;
        rem -= M;
    }
    q.R = r;
    return q;
}

BH_SFP_HD uint64_t bh_sfp_umulhi(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// min(floor(2^B * RN64(d / S)), 2^B - 1) for d = d_bits >= 0 (negative d gives 0), B <= 32.
BH_SFP_HD uint64_t bh_sfp_quantize(uint64_t d_bits, const bh_sfp_quant& Q, int B) {
    const uint64_t qmax = (1ull << B) - 1;
    if ((d_bits & ~BH_SFP_SIGN) == 0 || (d_bits & BH_SFP_SIGN)) return 0;
    const int eb = (int)((d_bits >> 52) & 0x7FF);
    const uint64_t D = (d_bits & BH_SFP_FRAC) | BH_SFP_HIDDEN;  // d is normal
    const int s = (eb - 1075) - Q.es + B;  // 2^B d / S = (D / M) 2^s, D/M in (1/2, 2)
    if (s <= -2) return 0;
    if (s > B + 1) return qmax;
    const uint64_t P = bh_sfp_umulhi(D, Q.R);  // floor(D 2^52 / M) - {0, 1}
    const int sh = 52 - s;
    uint64_t q = (sh >= 64) ? 0 : (P >> sh);  // floor(D 2^s / M) - {0, 1}
    uint64_t den, r;
    if (s >= 0) {
        den = Q.M;
        r = (D << s) - q * Q.M;  // exact modulo 2^64: the true remainder is below 2^55
    } else {
        den = Q.M << 1;
        r = D - q * den;
    }
    if (r >= den) {
        ++q;
        r -= den;
    }
    // RN64 rounds the quotient up to (q+1)/2^B iff it lies within half a gap below it; the
    // gap below (q+1)/2^B is 2^(E-B-52), or half that when q+1 = 2^E. Ties go to (q+1)/2^B,
    // whose significand is even.
    const uint64_t q1 = q + 1;
    const int E = 63 - bh_sfp_clz64(q1);
    const int k = 53 - E + (((q1 & (q1 - 1)) == 0) ? 1 : 0);
    const uint64_t gap = den - r;
    const uint64_t lim = (k >= 64) ? 0 : (den >> k);
    if (gap <= lim) q = q1;
    return q > qmax ? qmax : q;
}

// Root square of the reference QuadTree from the bounding box (quadtree.h:45-47):
// largest = max(maxx - minx, maxy - miny); side = ((minx + largest) + 1e-4) - minx.
BH_SFP_HD uint64_t bh_sfp_quadtree_side(uint64_t minx, uint64_t miny, uint64_t maxx, uint64_t maxy) {
    const uint64_t lx = bh_sfp_add(maxx, bh_sfp_neg(minx));
    const uint64_t ly = bh_sfp_add(maxy, bh_sfp_neg(miny));
    const uint64_t largest = bh_sfp_max_nonneg(lx, ly);
    const uint64_t pad = 0x3F1A36E2EB1C432Dull;  // 1.0e-4
    const uint64_t hx = bh_sfp_add(bh_sfp_add(minx, largest), pad);
    return bh_sfp_add(hx, bh_sfp_neg(minx));
}

// Largest extent of the Morton solvers (nBodySolverFMM_SYM_MORTON.h:84-85):
// compMax(maxPos + (-minPos)).
BH_SFP_HD uint64_t bh_sfp_morton_extent(uint64_t minx, uint64_t miny, uint64_t maxx, uint64_t maxy) {
    const uint64_t lx = bh_sfp_add(maxx, bh_sfp_neg(minx));
    const uint64_t ly = bh_sfp_add(maxy, bh_sfp_neg(miny));
    return bh_sfp_max_nonneg(lx, ly);
}

// Spread the low 16 / 32 bits of x into the even bit positions.
BH_SFP_HD uint32_t bh_spread16(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

BH_SFP_HD uint64_t bh_spread32(uint64_t x) {
    x &= 0xFFFFFFFFull;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}
