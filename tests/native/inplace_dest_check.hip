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
    printf("inplace_dest: 300000 random partitions match the swap loop\n");
    return 0;
}
