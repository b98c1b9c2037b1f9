// Host check of masterthesislucas_amd/csrc/bh_softfp.h against native binary64 arithmetic.
// Built and run by tests/test_softfp.py. Prints "OK <cases>" or the first mismatch and exits 1.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../masterthesislucas_amd/csrc/bh_softfp.h"

static uint64_t s_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {  // splitmix64
    uint64_t z = (s_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double uni() { return (next_u64() >> 11) * (1.0 / 9007199254740992.0); }
static double bits2d(uint64_t b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
}
static float bits2f(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static float rand_float() {
    switch (next_u64() % 6) {
        case 0: return (float)(uni() * 2.0 - 1.0);
        case 1: return (float)((uni() * 2.0 - 1.0) * std::pow(10.0, (int)(next_u64() % 12) - 6));
        case 2: {  // random finite bit pattern, subnormals included
            uint32_t b;
            do { b = (uint32_t)next_u64(); } while (((b >> 23) & 0xFF) == 0xFF);
            return bits2f(b);
        }
        case 3: return (float)((int)(next_u64() % 2001) - 1000) / 1024.0f;
        case 4: return (float)(uni() * 300.0 - 150.0);
        default: return bits2f((uint32_t)(next_u64() % 0x800000u));  // subnormal
    }
}

static double ref_quantize(double d, double S, int B) {
    const double scale = std::ldexp(1.0, B);
    double t = d / S;
    t = t * scale;
    t = (t < 0.0) ? 0.0 : t;
    t = (scale - 1.0 < t) ? scale - 1.0 : t;
    return std::floor(t);
}

int main(int argc, char** argv) {
    long n = argc > 1 ? atol(argv[1]) : 2000000;
    long cases = 0;
    // 0. the quantiser's reciprocal R = floor((2^116 - 1) / M) against integer division, for
    //    the significand range's ends and random significands
    {
        const unsigned __int128 N = ((unsigned __int128)1 << 116) - 1;
        const uint64_t lo = 1ull << 52, hi = (1ull << 53) - 1;
        for (long i = 0; i < n; ++i) {
            uint64_t M;
            if (i < 64) M = (i & 1) ? hi - (uint64_t)(i >> 1) : lo + (uint64_t)(i >> 1);
            else M = lo | (next_u64() & (lo - 1));
            const uint64_t s = (M & ~lo) | (1023ull << 52);  // 1.f with significand M
            const uint64_t got = bh_sfp_quant_prepare(s).R;
            const uint64_t want = (uint64_t)(N / M);
            if (got != want) {
                printf("FAIL reciprocal M=%llu got %llu want %llu\n", (unsigned long long)M,
                       (unsigned long long)got, (unsigned long long)want);
                return 1;
            }
            ++cases;
        }
    }
    // 1. float -> double conversion
    for (long i = 0; i < n; ++i) {
        uint32_t b;
        do { b = (uint32_t)next_u64(); } while (((b >> 23) & 0xFF) == 0xFF);
        float f = bits2f(b);
        uint64_t got = bh_sfp_from_f32(f);
        uint64_t want = bh_sfp_f64_bits((double)f);
        if (got != want) {
            printf("FAIL from_f32 %08x got %016llx want %016llx\n", b, (unsigned long long)got,
                   (unsigned long long)want);
            return 1;
        }
        ++cases;
    }
    // 2. addition of float-derived and general doubles (signs, cancellation, wide exponent gaps)
    for (long i = 0; i < n; ++i) {
        double a, b;
        switch (next_u64() % 4) {
            case 0: a = rand_float(); b = rand_float(); break;
            case 1: a = rand_float(); b = -(double)rand_float(); break;
            case 2: {
                a = (uni() * 2 - 1) * std::pow(2.0, (int)(next_u64() % 80) - 40);
                b = (uni() * 2 - 1) * std::pow(2.0, (int)(next_u64() % 80) - 40);
                break;
            }
            default: {  // near cancellation
                a = (uni() * 2 - 1) * 8.0;
                uint64_t ab = bh_sfp_f64_bits(a) + (next_u64() % 5) - 2;
                b = -bits2d(ab);
                break;
            }
        }
        uint64_t got = bh_sfp_add(bh_sfp_f64_bits(a), bh_sfp_f64_bits(b));
        uint64_t want = bh_sfp_f64_bits(a + b);
        if (got != want) {
            printf("FAIL add %.17g + %.17g got %.17g want %.17g\n", a, b, bits2d(got), a + b);
            return 1;
        }
        ++cases;
    }
    // 3. quantiser: floor(2^B RN64(d/S)) clamped, random and adversarial boundary cases
    const int Bs[4] = {16, 20, 24, 32};
    for (long i = 0; i < n; ++i) {
        const int B = Bs[next_u64() % 4];
        double S = std::ldexp(0.5 + 0.5 * uni(), (int)(next_u64() % 40) - 20);
        if (next_u64() % 4 == 0) S = (double)(float)S;
        double d;
        switch (next_u64() % 4) {
            case 0: d = S * uni(); break;
            case 1: {  // just around a cell boundary k/2^B
                const double k = std::floor(uni() * std::ldexp(1.0, B));
                double c = (k / std::ldexp(1.0, B)) * S;
                uint64_t cb = bh_sfp_f64_bits(c) + (next_u64() % 9) - 4;
                d = bits2d(cb);
                if (d < 0 || d > S) d = c;
                break;
            }
            case 2: d = S * std::ldexp(uni(), -(int)(next_u64() % 60)); break;
            default: {  // d as an exact float difference p - min
                float mn = rand_float(), p = rand_float();
                if (p < mn) { float t = p; p = mn; mn = t; }
                d = (double)p - (double)mn;
                S = d * (1.0 + uni());
                if (S == 0.0) S = 1.0;
                break;
            }
        }
        const bh_sfp_quant Q = bh_sfp_quant_prepare(bh_sfp_f64_bits(S));
        const uint64_t got = bh_sfp_quantize(bh_sfp_f64_bits(d), Q, B);
        const double want = ref_quantize(d, S, B);
        if ((double)got != want) {
            printf("FAIL quantize B=%d d=%.17g S=%.17g got %llu want %.17g\n", B, d, S,
                   (unsigned long long)got, want);
            return 1;
        }
        ++cases;
    }
    // 4. the two root-square rules
    for (long i = 0; i < n / 4; ++i) {
        float a = rand_float(), b = rand_float(), c = rand_float(), e = rand_float();
        if (std::fabs(a) > 1e30f || std::fabs(b) > 1e30f || std::fabs(c) > 1e30f ||
            std::fabs(e) > 1e30f)
            continue;
        double minx = std::fmin(a, b), maxx = std::fmax(a, b), miny = std::fmin(c, e),
               maxy = std::fmax(c, e);
        double largest = std::max(maxx - minx, maxy - miny);
        double want_side = ((minx + largest) + 0.0001) - minx;
        double want_ext = std::max(maxx + (-minx), maxy + (-miny));
        uint64_t s = bh_sfp_quadtree_side(bh_sfp_f64_bits(minx), bh_sfp_f64_bits(miny),
                                          bh_sfp_f64_bits(maxx), bh_sfp_f64_bits(maxy));
        uint64_t x = bh_sfp_morton_extent(bh_sfp_f64_bits(minx), bh_sfp_f64_bits(miny),
                                          bh_sfp_f64_bits(maxx), bh_sfp_f64_bits(maxy));
        if (s != bh_sfp_f64_bits(want_side) || x != bh_sfp_f64_bits(want_ext)) {
            printf("FAIL side %.17g %.17g %.17g %.17g\n", minx, miny, maxx, maxy);
            return 1;
        }
        ++cases;
    }
    // 5. binary64 -> binary32 rounding (normal float range)
    for (long i = 0; i < n; ++i) {
        double d = (uni() * 2 - 1) * std::ldexp(1.0, (int)(next_u64() % 200) - 100);
        if (next_u64() % 3 == 0) {  // halfway and near-halfway patterns
            uint64_t b = bh_sfp_f64_bits(d);
            b = (b & ~((1ull << 29) - 1)) | (1ull << 28) | (next_u64() % 3 == 0 ? 1u : 0u);
            d = bits2d(b);
        }
        if (bh_sfp_f32_bits(bh_sfp_to_f32(bh_sfp_f64_bits(d))) != bh_sfp_f32_bits((float)d)) {
            printf("FAIL to_f32 %.17g\n", d);
            return 1;
        }
        ++cases;
    }
    printf("OK %ld\n", cases);
    return 0;
}
