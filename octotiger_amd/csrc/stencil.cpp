// stencil.cpp -- interaction tables from the opening angle theta.
//
// Follows the reference's stencil rule: for each of the 8 parity positions of
// a cell inside its parent, every offset whose coarse (parent-level) distance
// passes the opening criterion 1/|dc| > theta enters the stencil; the offset is
// a multipole ("far") partner when its fine distance satisfies 1/|d| <= theta
// and a monopole-only ("near") partner otherwise; the union over the 8 parity
// positions is the stencil mask, and it must be symmetric and no wider than the
// stencil box compiled for theta >= 0.34
// (src/multipole_interactions/util/calculate_stencil.cpp:23-161,
//  src/monopole_interactions/util/calculate_stencil.cpp:19-177,
//  octotiger/common_kernel/helper.hpp:20-31).  The kernels then apply the
// per-pair coarse test dc^2 < (1/theta)^2 in double precision
// (octotiger/multipole_interactions/kernel/kokkos_kernel.hpp:202-256).
#include "octofmm_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace octofmm {
namespace {

double recip_dist(long long a0, long long a1, long long a2, long long b0, long long b1, long long b2) {
    const double t = static_cast<double>((a0 - b0) * (a0 - b0) + (a1 - b1) * (a1 - b1) + (a2 - b2) * (a2 - b2));
    return t > 0.0 ? 1.0 / std::sqrt(t) : 1.0e10;
}

inline int box_index(int x, int y, int z) {
    return (x + 5) * kStencilInx * kStencilInx + (y + 5) * kStencilInx + (z + 5);
}

// An 8 x 8 bit matrix held as 8 bytes, transposed: bit j of byte k becomes
// bit k of byte j.
uint64_t transpose_bytes(uint64_t m) {
    uint64_t r = 0;
    for (int k = 0; k < 8; ++k)
        for (int j = 0; j < 8; ++j) r |= ((m >> (8 * k + j)) & 1) << (8 * j + k);
    return r;
}

int find_delta(const Tables& t, int x, int y, int z) {
    for (int i = 0; i < t.ndelta; ++i)
        if (t.delta[i][0] == x && t.delta[i][1] == y && t.delta[i][2] == z) return i;
    return -1;
}

// The far schedule (Tables::sched): the split images of each pair of
// opposite face offsets whose far bits have the split form, then the other
// far offsets in delta order.  The split images come first as
// [2x, 2y, 1x, 1y, 2z, 1z] (kind, axis): the three kind-2 entries (4 steps
// each) sit at positions 0, 1 and 4, which fall to the waves with the most
// entries when 2, 4 or 8 waves take every 2nd, 4th or 8th entry of the 26 at
// theta = 0.5, and the two images of a pair, which read the same partner
// records, are staged together or one after the other, so the second read
// of those records finds them in the L2 cache.
void build_far_schedule(Tables& t) {
    std::vector<int> paired(kMaxDelta, 0);
    std::vector<int32_t> split1, split2;
    for (int axis = 2; axis >= 0; --axis) {
        int e[3] = {0, 0, 0};
        e[2 - axis] = 1;   // child-index bit 2 is x, 1 is y, 0 is z
        const int a = find_delta(t, e[0], e[1], e[2]), b = find_delta(t, -e[0], -e[1], -e[2]);
        if (a < 0 || b < 0) continue;
        bool split = true;
        for (int k = 0; k < 8; ++k)
            for (int j = 0; j < 8; ++j) {
                const int jb = (j >> axis) & 1, kb = (k >> axis) & 1;
                const bool fa = (t.farbits[a] >> (8 * k + j)) & 1, fb = (t.farbits[b] >> (8 * k + j)) & 1;
                split = split && fa == (jb == 1 || kb == 0) && fb == (jb == 0 || kb == 1);
            }
        if (!split) continue;
        paired[a] = paired[b] = 1;
        split1.push_back(a | b << 8 | 1 << 16 | axis << 24);
        split2.push_back(a | b << 8 | 2 << 16 | axis << 24);
    }
    std::vector<int32_t> order;
    const size_t np = split1.size();
    if (np > 0) order.push_back(split2[0]);
    if (np > 1) order.push_back(split2[1]);
    if (np > 0) order.push_back(split1[0]);
    if (np > 1) order.push_back(split1[1]);
    if (np > 2) {
        order.push_back(split2[2]);
        order.push_back(split1[2]);
    }
    for (int q = 0; q < t.nfar_delta; ++q)
        if (!paired[t.far_delta[q]]) order.push_back(t.far_delta[q]);
    const int n = (int)order.size();
    for (int q = 0; q < n; ++q) {
        const int32_t e = order[q];
        const int kind = (e >> 16) & 0xff, axis = e >> 24;
        uint64_t fa = 0, fb = 0;
        uint32_t half = 0;
        for (int k = 0; k < 8; ++k) {
            const int kb = (k >> axis) & 1;
            uint64_t ma = 0, mb = 0;
            if (kind == 0) {
                ma = (t.farbits[e & 0xff] >> (8 * k)) & 0xff;
            } else if (kind == 1) {
                ma = 0x0f;
                mb = 0xf0;
            } else {
                ma = kb ? 0 : 0x0f;
                mb = kb ? 0x0f : 0;
                half |= (uint32_t)kb << k;
            }
            fa |= ma << (8 * k);
            fb |= mb << (8 * k);
        }
        t.sched[q] = e;
        t.sched_fara[q] = fa;
        t.sched_farb[q] = fb;
        t.sched_rowa[q] = transpose_bytes(fa);
        t.sched_rowb[q] = transpose_bytes(fb);
        t.sched_half[q] = half;
    }
    t.nsched = n;
}

}   // namespace

bool build_tables(double theta, Tables& t, char* err, size_t errlen) {
    std::memset(&t, 0, sizeof(t));
    if (!(theta >= kThetaFloor) || !(theta < 1.0)) {
        std::snprintf(err, errlen, "theta %g outside [%g, 1)", theta, kThetaFloor);
        return false;
    }
    // stencil mask and phase over the 11^3 box; first parity position that
    // contributes an offset decides its phase (it is a function of the offset)
    std::vector<int> seen(kStencilBox, 0);
    for (int i = 0; i < kStencilBox; ++i) t.phase[i] = 1;
    int mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int p = 0; p < 8; ++p) {
        const int i0 = (p >> 2) & 1, i1 = (p >> 1) & 1, i2 = p & 1;
        for (int j0 = i0 - kInx; j0 < i0 + kInx; ++j0)
            for (int j1 = i1 - kInx; j1 < i1 + kInx; ++j1)
                for (int j2 = i2 - kInx; j2 < i2 + kInx; ++j2) {
                    if (i0 == j0 && i1 == j1 && i2 == j2) continue;
                    const int c0 = (i0 + kInx) / 2 - kInx / 2, c1 = (i1 + kInx) / 2 - kInx / 2,
                              c2 = (i2 + kInx) / 2 - kInx / 2;
                    const int e0 = (j0 + kInx) / 2 - kInx / 2, e1 = (j1 + kInx) / 2 - kInx / 2,
                              e2 = (j2 + kInx) / 2 - kInx / 2;
                    const double tf = recip_dist(i0, i1, i2, j0, j1, j2);
                    const double tc = recip_dist(c0, c1, c2, e0, e1, e2);
                    if (!(tc > theta)) continue;
                    const int d[3] = {j0 - i0, j1 - i1, j2 - i2};
                    for (int k = 0; k < 3; ++k) {
                        mn[k] = d[k] < mn[k] ? d[k] : mn[k];
                        mx[k] = d[k] > mx[k] ? d[k] : mx[k];
                    }
                    if (d[0] < -5 || d[0] > 5 || d[1] < -5 || d[1] > 5 || d[2] < -5 || d[2] > 5) continue;
                    const int bi = box_index(d[0], d[1], d[2]);
                    if (!seen[bi]) {
                        seen[bi] = 1;
                        t.mask[bi] = 1;
                        t.phase[bi] = (tf <= theta) ? 1 : 0;
                    }
                }
    }
    if (mx[2] != mx[1] || mx[1] != mx[0] || mn[2] != mn[1] || mn[1] != mn[0] || mn[2] != -mx[2]) {
        std::snprintf(err, errlen, "stencil for theta %g is not symmetric", theta);
        return false;
    }
    if (mx[0] > 5) {
        std::snprintf(err, errlen, "stencil for theta %g wider than the compiled maximum 5", theta);
        return false;
    }
    // p2p constants and root exclusions over the box
    for (int x = -5; x <= 5; ++x)
        for (int y = -5; y <= 5; ++y)
            for (int z = -5; z <= 5; ++z) {
                const int bi = box_index(x, y, z);
                if (t.mask[bi]) {
                    const double fx = x, fy = y, fz = z;
                    const double r = std::sqrt(fx * fx + fy * fy + fz * fz);
                    const double r3 = r * r * r;
                    t.four[bi][0] = -1.0 / r;
                    t.four[bi][1] = fx / r3;
                    t.four[bi][2] = fy / r3;
                    t.four[bi][3] = fz / r3;
                }
                t.root_excluded[bi] = (!t.phase[bi] || (x == 0 && y == 0 && z == 0)) ? 1 : 0;
            }
    // coarse offsets passing the kernels' double-precision coarse test
    const double th2 = (1.0 / theta) * (1.0 / theta);
    int n = 0, maxd = 0;
    for (int ax = -2; ax <= 2; ++ax)
        for (int ay = -2; ay <= 2; ++ay)
            for (int az = -2; az <= 2; ++az) {
                if (!(static_cast<double>(ax * ax + ay * ay + az * az) < th2)) continue;
                uint64_t fb = 0, nb = 0;
                for (int k = 0; k < 8; ++k)
                    for (int j = 0; j < 8; ++j) {
                        const int dx = 2 * ax + ((j >> 2) & 1) - ((k >> 2) & 1);
                        const int dy = 2 * ay + ((j >> 1) & 1) - ((k >> 1) & 1);
                        const int dz = 2 * az + (j & 1) - (k & 1);
                        if (dx == 0 && dy == 0 && dz == 0) continue;
                        if (dx < -5 || dx > 5 || dy < -5 || dy > 5 || dz < -5 || dz > 5) continue;
                        const int bi = box_index(dx, dy, dz);
                        if (!t.mask[bi]) continue;
                        if (t.phase[bi])
                            fb |= uint64_t(1) << (8 * k + j);
                        else
                            nb |= uint64_t(1) << (8 * k + j);
                    }
                if (!fb && !nb) continue;
                t.delta[n][0] = (int8_t)ax;
                t.delta[n][1] = (int8_t)ay;
                t.delta[n][2] = (int8_t)az;
                t.farbits[n] = fb;
                t.nearbits[n] = nb;
                const int m = std::max(std::abs(ax), std::max(std::abs(ay), std::abs(az)));
                maxd = m > maxd ? m : maxd;
                ++n;
            }
    t.ndelta = n;
    t.max_delta = maxd;
    t.nfar_delta = 0;
    for (int i = 0; i < n; ++i) {
        if (t.farbits[i]) t.far_delta[t.nfar_delta++] = i;
        t.allrows[i] = transpose_bytes(t.farbits[i] | t.nearbits[i]);
    }
    build_far_schedule(t);
    int nf = 0, nn = 0;
    for (int i = 0; i < n; ++i) {
        nf += __builtin_popcountll(t.farbits[i] & 0xffull);
        nn += __builtin_popcountll(t.nearbits[i] & 0xffull);
    }
    t.nfar_per_cell = nf;
    t.nnear_per_cell = nn;
    for (int k = 0; k < 8; ++k) {
        int a = 0, b = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < 8; ++j) {
                if ((t.farbits[i] >> (8 * k + j)) & 1) t.farlist[k][a++] = (uint16_t)(i * 8 + j);
                if ((t.nearbits[i] >> (8 * k + j)) & 1) t.nearlist[k][b++] = (uint16_t)(i * 8 + j);
            }
        t.nfar_k[k] = a;
        t.nnear_k[k] = b;
    }
    const int h = 2 * maxd + 1;
    const int W = kInx + 2 * h;
    t.p2p_halo = h;
    std::vector<int> slot(kStencilBox, -1);
    t.np2pconst = 0;
    for (int k = 0; k < 8; ++k) {
        int q = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < 8; ++j) {
                if (!(((t.farbits[i] | t.nearbits[i]) >> (8 * k + j)) & 1)) continue;
                const int dx = 2 * t.delta[i][0] + ((j >> 2) & 1) - ((k >> 2) & 1);
                const int dy = 2 * t.delta[i][1] + ((j >> 1) & 1) - ((k >> 1) & 1);
                const int dz = 2 * t.delta[i][2] + (j & 1) - (k & 1);
                const int woff = (dx * W + dy) * W + dz;
                const int bi = box_index(dx, dy, dz);
                if (slot[bi] < 0) {
                    slot[bi] = t.np2pconst++;
                    for (int c = 0; c < 4; ++c) t.p2pconst[slot[bi]][c] = t.four[bi][c];
                }
                t.p2plist[k][q++] = (slot[bi] * 32) | ((woff + 2048) << 16);
            }
        t.np2p_k[k] = q;
    }
    // with max |delta| = 1 every position 2 delta + j of the 6^3 block lies in
    // the box, so 215 pairs per class means the block minus the target itself
    t.p2p_cube = (maxd == 1 && h == 3) ? 1 : 0;
    for (int k = 0; k < 8; ++k) t.p2p_cube &= (t.np2p_k[k] == 215) ? 1 : 0;
    return true;
}

}   // namespace octofmm
