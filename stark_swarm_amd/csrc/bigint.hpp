// bigint.hpp -- host-side multi-word integer division for the trace
// builders (the witness quotients of the RSA-2048 and P-256 AIRs).
#pragma once
#include <stdint.h>
#include <string.h>

namespace sg {

// u (ul words) = q * v + r for v of vl significant words (v[vl-1] != 0),
// ul, vl <= MAXW * 2 and MAXW: long division with 32-bit digits and a
// normalised divisor (Knuth, TAOCP vol. 2, 4.3.1, algorithm D). q gets
// ul - vl + 1 words, r gets vl words.
template <int MAXW>
inline void divmod_words(const uint32_t *u, int ul, const uint32_t *v, int vl, uint32_t *q, uint32_t *r) {
    if (ul < vl) {
        memset(q, 0, 4);
        for (int i = 0; i < vl; i++) r[i] = i < ul ? u[i] : 0;
        return;
    }
    if (vl == 1) {
        uint64_t rem = 0;
        for (int j = ul - 1; j >= 0; j--) {
            const uint64_t cur = rem << 32 | u[j];
            q[j] = (uint32_t)(cur / v[0]);
            rem = cur % v[0];
        }
        r[0] = (uint32_t)rem;
        return;
    }
    const int s = __builtin_clz(v[vl - 1]);
    uint32_t vn[MAXW], un[2 * MAXW + 1];
    for (int i = vl - 1; i > 0; i--) vn[i] = (uint32_t)(((uint64_t)v[i] << s) | ((uint64_t)v[i - 1] >> (32 - s)));
    vn[0] = v[0] << s;
    un[ul] = (uint32_t)((uint64_t)u[ul - 1] >> (32 - s));
    for (int i = ul - 1; i > 0; i--) un[i] = (uint32_t)(((uint64_t)u[i] << s) | ((uint64_t)u[i - 1] >> (32 - s)));
    un[0] = u[0] << s;
    const uint64_t B = 1ULL << 32;
    for (int j = ul - vl; j >= 0; j--) {
        const uint64_t num = (uint64_t)un[j + vl] << 32 | un[j + vl - 1];
        uint64_t qh = num / vn[vl - 1], rh = num % vn[vl - 1];
        while (qh >= B || qh * vn[vl - 2] > (rh << 32 | un[j + vl - 2])) {
            qh--;
            rh += vn[vl - 1];
            if (rh >= B) break;
        }
        int64_t borrow = 0, t;
        for (int i = 0; i < vl; i++) {
            const uint64_t p = qh * vn[i];
            t = (int64_t)un[i + j] - borrow - (int64_t)(p & 0xFFFFFFFFu);
            un[i + j] = (uint32_t)t;
            borrow = (int64_t)(p >> 32) - (t >> 32);
        }
        t = (int64_t)un[j + vl] - borrow;
        un[j + vl] = (uint32_t)t;
        q[j] = (uint32_t)qh;
        if (t < 0) {  // qh was one too large: add v back
            q[j]--;
            uint64_t carry = 0;
            for (int i = 0; i < vl; i++) {
                const uint64_t sum = (uint64_t)un[i + j] + vn[i] + carry;
                un[i + j] = (uint32_t)sum;
                carry = sum >> 32;
            }
            un[j + vl] += (uint32_t)carry;
        }
    }
    for (int i = 0; i < vl; i++) r[i] = (uint32_t)(((uint64_t)un[i] >> s) | ((uint64_t)un[i + 1] << (32 - s)));
}

}  // namespace sg
