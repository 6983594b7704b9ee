/*
 * cbrt_exact.c -- correctly rounded double cube root for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see lulesh_oracle.h).
 *
 * Method: |x| = M * 2^E with M a 53-bit integer.  Pick t with
 * k = E - 3t in {104,105,106}, so cbrt(|x|) = cbrt(M * 2^k) * 2^t and
 * cbrt(M * 2^k) lies in [2^52, 2^53].  The integer Y nearest to that cube
 * root is found from a long-double estimate and then fixed by the exact
 * midpoint test  (2Y-1)^3 < 8 M 2^k < (2Y+1)^3  in 192-bit integer
 * arithmetic.  A tie is impossible (the left side is odd, the right side
 * even), so Y * 2^t is the correctly rounded result.
 */
#include <math.h>
#include <stdint.h>

typedef unsigned __int128 u128;

/* little-endian three-limb unsigned integer */
typedef struct { uint64_t w[3]; } u192;

static u192 cube_u64(uint64_t a)
{
    u128 a2 = (u128)a * a;                  /* < 2^110 */
    uint64_t l = (uint64_t)a2, h = (uint64_t)(a2 >> 64);
    u128 pl = (u128)l * a;
    u128 ph = (u128)h * a;
    u192 r;
    r.w[0] = (uint64_t)pl;
    u128 mid = (pl >> 64) + (uint64_t)ph;
    r.w[1] = (uint64_t)mid;
    r.w[2] = (uint64_t)(ph >> 64) + (uint64_t)(mid >> 64);
    return r;
}

static u192 shl_u64(uint64_t m, int s)  /* 64 <= s < 128 */
{
    u192 r;
    r.w[0] = 0;
    r.w[1] = m << (s - 64);
    r.w[2] = (s - 64) ? (m >> (128 - s)) : 0;
    return r;
}

static int cmp192(u192 a, u192 b)
{
    for (int i = 2; i >= 0; --i) {
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    }
    return 0;
}

double lo_cbrt_exact(double x)
{
    if (x == 0.0 || !isfinite(x)) return x;
    int neg = x < 0.0;
    double ax = neg ? -x : x;
    int e;
    double m = frexp(ax, &e);               /* ax = m 2^e, m in [0.5,1) */
    uint64_t M = (uint64_t)ldexp(m, 53);    /* exact: 53-bit integer   */
    int E = e - 53;
    int t = (E - 104) >= 0 ? (E - 104) / 3 : -((104 - E + 2) / 3);
    int k = E - 3 * t;                      /* 104..106 */
    u192 n8 = shl_u64(M, k + 3);            /* 8 * M * 2^k */

    long double est = cbrtl(ldexpl((long double)M, k));
    uint64_t Y = (uint64_t)llroundl(est);
    for (int it = 0; it < 8; ++it) {
        if (cmp192(n8, cube_u64(2 * Y + 1)) > 0) { ++Y; continue; }
        if (cmp192(n8, cube_u64(2 * Y - 1)) < 0) { --Y; continue; }
        break;
    }
    double r = ldexp((double)Y, t);
    return neg ? -r : r;
}
