// Host check of the far schedule of the direct-staging kernels
// (Tables::sched, built by stencil.cpp): walks every entry as the kernels do
// -- the image records from the staging formula of m2l_item_direct, each
// class's reads from sched_fara / sched_farb / sched_half -- and counts how
// often each far (delta, k, j) pair of the tables is evaluated: exactly once
// at every theta, with 196 lockstep steps per cell instead of 208 at 0.5.
// Also checks the lane masks the kernels form from the transposed step rows
// (spread_groups, repeat_bytes, step_lanes) against the per-lane rule.
#include "octofmm_internal.hpp"

#include <cstdio>
#include <vector>

using namespace octofmm;

int main() {
    int fails = 0;
    // the lane-mask helpers, exhaustively: byte g of spread_groups(e) is all
    // ones exactly where bit g of e is set, and every byte of repeat_bytes(b) is b
    for (unsigned e = 0; e < 256; ++e) {
        const uint64_t sg = spread_groups(e), rb = repeat_bytes(e);
        for (int g = 0; g < 8; ++g) {
            if (((sg >> (8 * g)) & 0xff) != (((e >> g) & 1) ? 0xffu : 0u)) ++fails;
            if (((rb >> (8 * g)) & 0xff) != e) ++fails;
        }
    }
    if (fails) printf("lane-mask helpers wrong: %d\n", fails);
    for (double th : {0.5, 0.49, 0.45, 0.4, 0.34}) {
        static Tables t; char err[256];
        if (!build_tables(th, t, err, sizeof err)) { printf("fail %s\n", err); return 1; }
        std::vector<int> cov(kMaxDelta * 64, 0);
        int steps = 0;
        for (int q = 0; q < t.nsched; ++q) {
            const int e = t.sched[q], a = e & 0xff, b = (e >> 8) & 0xff, kind = (e >> 16) & 0xff, axis = e >> 24;
            // image records 0..7: (delta, child)
            int rd[8], rj[8];
            if (kind == 0) { for (int r = 0; r < 8; ++r) { rd[r] = a; rj[r] = r; } }
            else {
                const int va = kind == 1 ? 1 : 0;
                for (int lo = 0; lo < 48; lo += 12) {   // the staging formula for lanes 0..47, first chunk of each record
                    const int r = lo / 12, low = (1 << axis) - 1, jr = ((r & ~low) << 1) | (r & low);
                    rd[r] = a; rj[r] = jr | (va << axis);
                    rd[4 + r] = b; rj[4 + r] = jr | ((1 - va) << axis);
                }
            }
            int maxs = 0;
            for (int k = 0; k < 8; ++k) {
                const unsigned m = (unsigned)((t.sched_fara[q] >> (8 * k)) | (t.sched_farb[q] >> (8 * k))) & 0xff;
                const int off = ((t.sched_half[q] >> k) & 1) * 4;
                for (int s = 0; s < 8; ++s) if ((m >> s) & 1) { cov[rd[s + off] * 64 + k * 8 + rj[s + off]]++; if (s + 1 > maxs) maxs = s + 1; }
                // partner-emptiness halves: a-mask bits must come from a's records, b-mask bits from b's
                for (int s = 0; s < 8; ++s) {
                    if ((t.sched_fara[q] >> (8 * k + s)) & 1) if (rd[s + off] != a) ++fails;
                    if ((t.sched_farb[q] >> (8 * k + s)) & 1) if (rd[s + off] != b || kind == 0) ++fails;
                    // the step masks the kernels read: the same bits by step, then class
                    if (((t.sched_rowa[q] >> (8 * s + k)) & 1) != ((t.sched_fara[q] >> (8 * k + s)) & 1)) ++fails;
                    if (((t.sched_rowb[q] >> (8 * s + k)) & 1) != ((t.sched_farb[q] >> (8 * k + s)) & 1)) ++fails;
                }
                // the lane mask of each step (step_lanes) against the per-lane
                // rule: lane (group g, class k) evaluates step s when class k
                // reads record s from a and g's partner at a is there, or from b
                // and g's partner at b is there; over a sample of empty-group sets
                static const unsigned kEmpty[] = {0x00, 0xff, 0x01, 0x80, 0x55, 0xaa, 0x0f, 0xf0, 0x25, 0xc8};
                for (unsigned ea : kEmpty)
                    for (unsigned eb : kEmpty)
                        for (int s = 0; s < 8; ++s) {
                            const uint64_t m = step_lanes(t.sched_rowa[q], t.sched_rowb[q], spread_groups(ea),
                                                          spread_groups(eb), s);
                            for (int g = 0; g < 8; ++g) {
                                const bool fa = ((t.sched_fara[q] >> (8 * k + s)) & 1) && !((ea >> g) & 1);
                                const bool fb = ((t.sched_farb[q] >> (8 * k + s)) & 1) && !((eb >> g) & 1);
                                if ((((m >> (8 * g + k)) & 1) != 0) != (fa || fb)) ++fails;
                            }
                        }
            }
            steps += maxs;
        }
        // the p2m kernel's step rows: far or near bits transposed
        for (int d = 0; d < t.ndelta; ++d)
            for (int k = 0; k < 8; ++k)
                for (int j = 0; j < 8; ++j)
                    if (((t.allrows[d] >> (8 * j + k)) & 1) != (((t.farbits[d] | t.nearbits[d]) >> (8 * k + j)) & 1)) ++fails;
        int bad = 0, nfar = 0;
        for (int d = 0; d < t.ndelta; ++d) for (int bit = 0; bit < 64; ++bit) {
            const int f = (t.farbits[d] >> bit) & 1; nfar += f;
            if (cov[d * 64 + bit] != f) ++bad;
        }
        fails += bad;
        printf("theta %.2f entries %d steps %d plain %d far %d mismatches %d\n", th, t.nsched, steps,
               8 * t.nfar_delta, nfar, bad);
    }
    printf(fails ? "FAIL\n" : "OK\n");
    return fails != 0;
}
