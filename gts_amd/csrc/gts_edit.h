// Exact Levenshtein distance (unit costs), usable on host and device.
//
// The reference computes it with a full (len+1)^2 Wagner-Fischer table per
// thread (tree.cuh:128-154, search_v2.cuh:388-414), 47.5 KB of local memory.
// Here the shorter string (<= 64 bytes) is the pattern of Hyyro's bit-vector
// formulation of Myers' algorithm (one 64-bit word holds a DP column as vertical
// +1/-1 deltas), so a pair costs O(len) 64-bit integer operations.  Pattern
// match masks are built per text byte with a SWAR zero-byte test on the pattern
// held in registers.  Longer pairs use a rolling-row DP.
#pragma once
#include <stdint.h>

#ifndef GTS_HD
#define GTS_HD __host__ __device__
#endif

#define GTS_EDIT_MAXLEN 255

// Match mask of byte c against the pattern words (4 bytes per word, m bytes).
GTS_HD inline uint64_t gts_eq_mask(const uint32_t *pw, int nw, uint64_t mmask, unsigned c)
{
    uint32_t rep = c * 0x01010101u;
    uint64_t eq = 0;
#pragma unroll
    for (int w = 0; w < 16; w++) {
        if (w < nw) {
            uint32_t x = pw[w] ^ rep;
            // high bit of each byte set where that byte of x is zero
            uint32_t z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
            uint32_t b4 = ((z >> 7) & 1u) | ((z >> 14) & 2u) | ((z >> 21) & 4u) | ((z >> 28) & 8u);
            eq |= (uint64_t)b4 << (4 * w);
        }
    }
    return eq & mmask;
}

// Levenshtein distance of a (length la) and b (length lb), both <= GTS_EDIT_MAXLEN.
GTS_HD inline int gts_edit_distance(const uint8_t *a, int la, const uint8_t *b, int lb)
{
    if (la == 0) return lb;
    if (lb == 0) return la;
    const uint8_t *p = a, *t = b;
    int m = la, n = lb;
    if (m > n) { p = b; t = a; m = lb; n = la; }
    if (m <= 64) {
        uint32_t pw[16];
        int nw = (m + 3) >> 2;
#pragma unroll
        for (int w = 0; w < 16; w++) {
            uint32_t v = 0;
            if (w < nw) {
                for (int i = 0; i < 4; i++) {
                    int idx = 4 * w + i;
                    if (idx < m) v |= (uint32_t)p[idx] << (8 * i);
                }
            }
            pw[w] = v;
        }
        const uint64_t mmask = (m == 64) ? ~0ull : ((1ull << m) - 1ull);
        const uint64_t hb = 1ull << (m - 1);
        uint64_t Pv = ~0ull, Mv = 0ull;
        int score = m;
        for (int j = 0; j < n; j++) {
            uint64_t Eq = gts_eq_mask(pw, nw, mmask, t[j]);
            uint64_t Xv = Eq | Mv;
            uint64_t Xh = (((Eq & Pv) + Pv) ^ Pv) | Eq;
            uint64_t Ph = Mv | ~(Xh | Pv);
            uint64_t Mh = Pv & Xh;
            if (Ph & hb) score++;
            else if (Mh & hb) score--;
            Ph = (Ph << 1) | 1ull;
            Mh = Mh << 1;
            Pv = Mh | ~(Xv | Ph);
            Mv = Ph & Xv;
        }
        return score;
    }
    // rolling-row DP for long pairs (m > 64): row over the shorter string
    int row[GTS_EDIT_MAXLEN + 1];
    for (int j = 0; j <= m; j++) row[j] = j;
    for (int i = 1; i <= n; i++) {
        int diag = row[0];
        row[0] = i;
        for (int j = 1; j <= m; j++) {
            int up = row[j];
            int v = up + 1;
            if (row[j - 1] + 1 < v) v = row[j - 1] + 1;
            int c = diag + (t[i - 1] == p[j - 1] ? 0 : 1);
            if (c < v) v = c;
            row[j] = v;
            diag = up;
        }
    }
    return row[m];
}
