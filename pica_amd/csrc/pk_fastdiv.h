// pk_fastdiv.h -- division of non-negative 31-bit integers by a divisor fixed for a launch.
//
// Round-up multiplier method of Granlund & Montgomery ("Division by invariant integers using
// multiplication", PLDI 1994, section 4): with s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1, the quotient n / d is (mulhi(m, n) + n) >> s for every
// 0 <= n < 2^31 (the sum then fits in 32 bits).  Used for the linear cell index of the FP32
// split-position particles, which a kernel decodes once per particle.
#pragma once

#include <stdint.h>

struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    f.s = 0;
    while (((uint64_t)1 << f.s) < d) f.s++;
    f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << f.s) - d)) / d + 1);
    return f;
}

#if defined(__HIPCC__)
#define PK_FASTDIV_HD __host__ __device__ __forceinline__
#else
#define PK_FASTDIV_HD inline
#endif

// the high half of the 64-bit product compiles to one v_mul_hi_u32 on the device
PK_FASTDIV_HD uint32_t fastdiv(uint32_t n, const FastDiv& f)
{
    return (uint32_t)((((uint64_t)f.m * n) >> 32) + n) >> f.s;
}
