/* exact_arith_check.c — CPU proof of the division-free exact arithmetic in the
 * HIP kernels: each fast sequence is compared, for every input of its range,
 * with the IEEE operation it replaces (glibc's sqrtf and float division are
 * correctly rounded).  The hardware approximations (v_sqrt_f32, one ulp) are
 * modelled by every value within one ulp of the exact result.
 *
 *   sqrt_rn     (kfx_common.h): all x in [2^-100, 2^100), 3 starting values
 *   tsdf_decode (kfx_common.h): all 65536 int16 values, v / 32767
 *   div_cell    (kfx_raycast.hip): p / c for every float |p| in [2^-20, 2^8)
 *               and each cell size c of the test and benchmark volumes
 *   rcp_weight  (tsdf_average_quotient, kfx_common.h): RN(1/b) for b in
 *               [1, 129] from v_rcp_f32 (3 starting values) and one Newton step
 *   quotient    (ICP projection, kfx_icp.hip): a / b as Markstein's
 *               correction from RN(1/b), 2*10^8 random pairs (b in
 *               [2^-8, 2^8), |a| in [2^-30, 2^14)); correct rounding for all
 *               pairs is Markstein's theorem, this samples it
 *   inv_depth   (kfx_denoise.hip): 1000 / d for every d in [1, 65535]
 *   avg_div     (tsdf_average_quotient): a / b for b in [2, 129] and every
 *               float a in [2^-100, 129] (the quotient is odd in a, so
 *               positive a suffice)
 *
 * Prints one line per check with its mismatch count; exit status 1 if any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float sqrt_rn(float x, float s) {
    const float sdn = fb(bf(s) - 1u), sup = fb(bf(s) + 1u);
    const float rdn = fmaf(-sdn, s, x), rup = fmaf(-sup, s, x);
    float r = s;
    if (rdn <= 0.0f) r = sdn;
    if (rup > 0.0f) r = sup;
    return r;
}

int main(void) {
    long bad_sqrt = 0, n_sqrt = 0;
    const uint32_t lo = 27u << 23, hi = 227u << 23; /* [2^-100, 2^100) */
#pragma omp parallel for reduction(+ : bad_sqrt, n_sqrt) schedule(static, 1 << 16)
    for (long u = lo; u < (long)hi; ++u) {
        const float x = fb((uint32_t)u), ref = sqrtf(x);
        for (int d = -1; d <= 1; ++d) {
            bad_sqrt += sqrt_rn(x, fb(bf(ref) + (uint32_t)d)) != ref;
            ++n_sqrt;
        }
    }
    printf("sqrt_rn: %ld cases, %ld mismatches\n", n_sqrt, bad_sqrt);

    long bad_dec = 0;
    const float y = 1.0f / 32767.0f;
    for (int v = -32768; v <= 32767; ++v) {
        const float a = (float)v, q0 = a * y, r = fmaf(-q0, 32767.0f, a), q = fmaf(r, y, q0);
        bad_dec += q != a / 32767.0f;
    }
    printf("tsdf_decode: 65536 cases, %ld mismatches\n", bad_dec);

    const float cells[] = {3.072f / 512.0f, 3.0f / 512.0f, 3.0f / 256.0f, 3.0f / 128.0f, 3.0f / 64.0f,
                           1.0f / 64.0f,    1.0f / 32.0f,  1.28f / 64.0f, 1.2f / 64.0f,  1.2f / 32.0f,
                           6.144f / 1024.0f};
    long bad_div = 0, n_div = 0;
    for (unsigned ci = 0; ci < sizeof(cells) / sizeof(cells[0]); ++ci) {
        const float c = cells[ci], ic = 1.0f / c;
        const uint32_t plo = 0x35800000u, phi = 0x43800000u; /* [2^-20, 2^8) */
#pragma omp parallel for reduction(+ : bad_div, n_div) schedule(static, 1 << 16)
        for (long u = plo; u < (long)phi; ++u) {
            for (int s = 0; s < 2; ++s) {
                const float p = s ? -fb((uint32_t)u) : fb((uint32_t)u);
                const float q0 = p * ic, r = fmaf(-q0, c, p), q = fmaf(r, ic, q0);
                bad_div += q != p / c;
                ++n_div;
            }
        }
    }
    printf("div_cell: %ld cases, %ld mismatches\n", n_div, bad_div);

    long bad_rcp = 0;
    for (int b = 1; b <= 129; ++b) {
        const float x = (float)b, ref = 1.0f / x;
        for (int d = -1; d <= 1; ++d) {
            const float y0 = fb(bf(ref) + (uint32_t)d);
            bad_rcp += fmaf(fmaf(-x, y0, 1.0f), y0, y0) != ref;
        }
    }
    printf("rcp_weight: 387 cases, %ld mismatches\n", bad_rcp);

    long bad_avg = 0, n_avg = 0;
    for (int b = 2; b <= 129; ++b) {
        const float fbv = (float)b, y = 1.0f / fbv;
        const uint32_t alo = 27u << 23, ahi = 0x43010001u; /* [2^-100, 129] */
#pragma omp parallel for reduction(+ : bad_avg, n_avg) schedule(static, 1 << 16)
        for (long u = alo; u < (long)ahi; ++u) {
            const float a = fb((uint32_t)u);
            const float q0 = a * y, q = fmaf(fmaf(-q0, fbv, a), y, q0);
            bad_avg += q != a / fbv;
            ++n_avg;
        }
    }
    printf("avg_div: %ld cases, %ld mismatches\n", n_avg, bad_avg);

    long bad_q = 0;
    uint64_t st = 0x9E3779B97F4A7C15ull;
    for (long i = 0; i < 200000000L; ++i) {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        const float b = fb(((uint32_t)(119 + (st % 16)) << 23) | (uint32_t)((st >> 8) & 0x7FFFFF));
        const float a = fb(((uint32_t)(st >> 63) << 31) | ((uint32_t)(97 + ((st >> 32) % 44)) << 23) |
                           (uint32_t)((st >> 40) & 0x7FFFFF));
        const float y = 1.0f / b, q0 = a * y, q = fmaf(fmaf(-q0, b, a), y, q0);
        bad_q += q != a / b;
    }
    printf("quotient: 200000000 cases, %ld mismatches\n", bad_q);

    long bad_inv = 0;
    for (int d = 1; d <= 65535; ++d) {
        const float b = (float)d, y = 1.0f / b, q0 = 1000.0f * y;
        bad_inv += fmaf(fmaf(-q0, b, 1000.0f), y, q0) != 1000.0f / b;
    }
    printf("inv_depth: 65535 cases, %ld mismatches\n", bad_inv);
    return (bad_sqrt || bad_dec || bad_div || bad_rcp || bad_avg || bad_q || bad_inv) ? 1 : 0;
}
