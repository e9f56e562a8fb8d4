// avx_pin.cpp -- TEST INFRASTRUCTURE ONLY (pins the oracle's distance order).
//
// The squared-L2 kernel of src/metric.cpp:45-78 has this structure: 256-bit
// accumulator, D = dim rounded up to 8, the D % 32 tail handled first by a
// falling-through switch (+16, +8, +0), then 32-float blocks of four
// load/sub/mul/add steps, then the 8 lanes stored and summed left to right.
// This file writes that structure with the same intrinsics so that g++ -O3
// -mavx2 -mfma (what -march=native means on the reference's Xeon) decides
// the fusing exactly as it does for the reference; tests compare it bit for
// bit with the oracle's scalar restatement.  Built twice: as the reference
// is (contraction on) and with -ffp-contract=off (to show the test can tell
// the two apart).
#include <immintrin.h>

extern "C" float pin_l2sqr_avx(const float *t1, const float *t2, unsigned dim) {
    __m256 sum, a, b;
    const unsigned D = (dim + 7) & ~7U;
    const unsigned DR = D % 32;
    const unsigned DD = D - DR;
    const float *l = t1, *r = t2;
    const float *el = l + DD, *er = r + DD;
    float unpack[8] __attribute__((aligned(32))) = {0, 0, 0, 0, 0, 0, 0, 0};
    sum = _mm256_load_ps(unpack);
#define PIN_STEP(pa, pb)              \
    a = _mm256_loadu_ps(pa);          \
    b = _mm256_loadu_ps(pb);          \
    a = _mm256_sub_ps(a, b);          \
    a = _mm256_mul_ps(a, a);          \
    sum = _mm256_add_ps(sum, a);
    switch (DR) {
        case 24: PIN_STEP(el + 16, er + 16)  // fall through
        case 16: PIN_STEP(el + 8, er + 8)    // fall through
        case 8: PIN_STEP(el, er)
    }
    for (unsigned i = 0; i < DD; i += 32, l += 32, r += 32) {
        PIN_STEP(l, r)
        PIN_STEP(l + 8, r + 8)
        PIN_STEP(l + 16, r + 16)
        PIN_STEP(l + 24, r + 24)
    }
#undef PIN_STEP
    _mm256_storeu_ps(unpack, sum);
    return unpack[0] + unpack[1] + unpack[2] + unpack[3] + unpack[4] + unpack[5] + unpack[6] + unpack[7];
}
