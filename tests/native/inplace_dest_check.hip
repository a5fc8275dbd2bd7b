// Host-side check of hp::inplace_dest (the closed form of fast3tree's
// in-place partition, src/fast3tree.c:762-773) against that loop, restated
// here, on random class patterns: every destination must be the loop's.
// Built and run by tests/test_inplace_order.py; no GPU code runs.
#include "hp_internal.h"
#include <cstdio>
#include <random>
#include <vector>
int main() {
    std::mt19937 rng(7);
    for (int trial = 0; trial < 300000; trial++) {
        int n = 1 + rng() % 70;
        double q = (rng() % 1000) / 1000.0;
        std::vector<int> big(n);
        for (auto &b : big) b = (rng() % 1000) / 1000.0 < q;
        // reference loop
        std::vector<int> a(n);
        for (int i = 0; i < n; i++) a[i] = i;
        int i = 0, j = n - 1;
        for (; i < j; i++)
            if (big[a[i]]) { std::swap(a[i], a[j]); i--; j--; }
        // closed form
        std::vector<int> st;
        int ns = 0;
        for (int k = 0; k < n; k++) if (!big[k]) { st.push_back(k); ns++; }
        for (int k = 0; k < n; k++) if (big[k]) st.push_back(k);
        std::vector<int> out(n, -1);
        int rl = 0, rr = 0;
        for (int p = 0; p < n; p++) {
            bool left = !big[p];
            int rank = left ? rl++ : rr++;
            int d = hp::inplace_dest(p, left, rank, n, ns, [&](int k) { return st[k]; });
            if (d < 0 || d >= n || out[d] != -1) { printf("bad dest n=%d p=%d d=%d\n", n, p, d); return 1; }
            out[d] = p;
        }
        if (out != a) { printf("mismatch n=%d\n", n); return 1; }
    }
    // select_bit against a scan of the bits
    for (int trial = 0; trial < 200000; trial++) {
        uint64_t m = ((uint64_t)rng() << 32) ^ rng();
        if (trial & 1) m &= ((uint64_t)rng() << 32) ^ rng();
        int k = 0;
        for (int b = 0; b < 64; b++)
            if (m >> b & 1) {
                if (hp::select_bit(m, k) != b) { printf("select_bit wrong\n"); return 1; }
                k++;
            }
    }
    // the lookups as the wavefront subtree makes them: a node of c points at
    // lanes [lf, lf + c), its k-th left / right point by select over masks
    for (int trial = 0; trial < 100000; trial++) {
        int c = 1 + rng() % 64, lf = rng() % (65 - c);
        uint64_t mN = (c == 64 ? ~0ull : ((1ull << c) - 1)) << lf, mL = 0;
        for (int b = lf; b < lf + c; b++) if (rng() & 1) mL |= 1ull << b;
        mL |= ((uint64_t)rng() << 32 ^ rng()) & ~mN;  // other nodes' lanes
        const uint64_t mLn = mL & mN, mRn = ~mL & mN;
        const int nl = __builtin_popcountll(mLn);
        std::vector<int> a(c);
        for (int i = 0; i < c; i++) a[i] = i;
        int i = 0, j = c - 1;
        for (; i < j; i++)
            if (!(mL >> (lf + a[i]) & 1)) { std::swap(a[i], a[j]); i--; j--; }
        std::vector<int> out(c, -1);
        for (int lane = lf; lane < lf + c; lane++) {
            const bool gl = mL >> lane & 1;
            const uint64_t below = mN & ((1ull << lane) - 1);
            const int rank = gl ? __builtin_popcountll(mL & below) : __builtin_popcountll(~mL & below);
            const int d = hp::inplace_dest(lane - lf, gl, rank, c, nl, [&](int k) {
                return (k < nl ? hp::select_bit(mLn, k) : hp::select_bit(mRn, k - nl)) - lf;
            });
            if (d < 0 || d >= c || out[d] != -1) { printf("bad mask dest\n"); return 1; }
            out[d] = lane - lf;
        }
        if (out != a) { printf("mask form mismatch c=%d\n", c); return 1; }
    }
    printf("inplace_dest: 300000 random partitions match the swap loop; select_bit and the "
           "mask form too\n");
    return 0;
}
