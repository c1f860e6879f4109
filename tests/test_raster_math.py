"""CPU checks of the raster's device arithmetic (blitzremesher_amd/csrc/
raster_device.hpp): the whole-rectangle coverage classifier used for large
faces and the lane paths' bitmaps (two samples per step; one row of up to
four samples per step) must agree with the per-sample predicate (src/evaluate.cpp:166-168) on every sample. The
functions are taken verbatim from the header, compiled for the host with
g++ -ffp-contract=off and checked by brute force over seeded triangles,
including corners on sample centres and strict (r = 0) edges."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "blitzremesher_amd" / "csrc" / "raster_device.hpp"

PRELUDE = """#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <cstdlib>
using std::fabs; using std::isfinite; using std::isnan; using std::fmax; using std::fmin; using std::max; using std::min;
using std::floor; using std::ceil;
#define __device__
#define __forceinline__ inline
// The device's approximate reciprocal only seeds a search; any value works.
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
struct double2 {
    double x, y;
};
struct Setup {
    double ax, ay, bx, by, cx, cy, r0, r1, r2;
    int x0, x1, y0, y1;
};
"""

CHECK = r"""
// The reference's predicate as written (src/evaluate.cpp:166-168), the
// ground truth for every device routine below.
static bool covers_ref(const Setup& s, int x, int y) {
    const double px = x + .5, py = y + .5;
    const double u = (s.cx - s.bx) * (py - s.by) - (s.cy - s.by) * (px - s.bx);
    const double v = (s.ax - s.cx) * (py - s.cy) - (s.ay - s.cy) * (px - s.cx);
    const double w = (s.bx - s.ax) * (py - s.ay) - (s.by - s.ay) * (px - s.ax);
    return !(u + s.r0 < 0 || v + s.r1 < 0 || w + s.r2 < 0);
}

int main(int argc, char** argv) {
    const int iterations = atoi(argv[1]);
    std::mt19937_64 rng(11);
    long rects = 0, full = 0, empty = 0, bad = 0, spans_mixed = 0;
    for (int it = 0; it < iterations; ++it) {
        const int mode = it % 4;
        const double span = it % 3 == 0 ? 24 : 80;
        auto coord = [&]() {
            double v = std::uniform_real_distribution<double>(0, span)(rng);
            if (mode == 1) v = std::round(v * 2) / 2;
            if (mode == 2) v = std::round(v * 8) / 8;
            if (mode == 3) v = std::round(v) + .5 + std::uniform_int_distribution<int>(-2, 2)(rng) * 1e-13;
            return v;
        };
        Setup s;
        s.ax = coord(); s.ay = coord(); s.bx = coord(); s.by = coord(); s.cx = coord(); s.cy = coord();
        if (it % 5 == 0) s.bx = s.ax;
        if (it % 7 == 0) s.cy = s.by;
        s.r0 = .5 * (fabs(s.cx - s.bx) + fabs(s.cy - s.by));
        s.r1 = .5 * (fabs(s.ax - s.cx) + fabs(s.ay - s.cy));
        s.r2 = .5 * (fabs(s.bx - s.ax) + fabs(s.by - s.ay));
        if (it % 11 == 0) s.r0 = s.r1 = s.r2 = 0;
        s.x0 = int(std::floor(std::min({s.ax, s.bx, s.cx})));
        s.x1 = int(std::floor(std::max({s.ax, s.bx, s.cx})));
        s.y0 = int(std::floor(std::min({s.ay, s.by, s.cy})));
        s.y1 = int(std::floor(std::max({s.ay, s.by, s.cy})));
        for (int ty = s.y0 >> 3; ty <= (s.y1 >> 3); ++ty)
            for (int tx = s.x0 >> 3; tx <= (s.x1 >> 3); ++tx) {
                const int xa = max(tx << 3, s.x0), xb = min((tx << 3) | 7, s.x1);
                const int ya = max(ty << 3, s.y0), yb = min((ty << 3) | 7, s.y1);
                const int kind = classify_rect(s, xa, xb, ya, yb);
                int n = 0, in = 0;
                uint64_t box = 0;
                for (int y = ya; y <= yb; ++y)
                    for (int x = xa; x <= xb; ++x) {
                        ++n;
                        in += covers_ref(s, x, y);
                        bad += covers(s, x, y) != covers_ref(s, x, y);
                        box |= 1ull << (8 * (y & 7) + (x & 7));
                    }
                ++rects;
                bad += rect_word(xa & 7, xb & 7, ya & 7, yb & 7) != box;
                if (kind == 0) { ++empty; bad += in != 0; }
                if (kind == 1) { ++full; bad += in != n; }
            }
    }
    // Lane path: the bitmap of a box of at most 8x8 samples, rows [ya, yb].
    for (int it = 0; it < iterations; ++it) {
        Setup s;
        const double ox = std::uniform_int_distribution<int>(0, 40)(rng), oy = std::uniform_int_distribution<int>(0, 40)(rng);
        auto c = [&](double o) {
            double v = o + std::uniform_real_distribution<double>(0, 7.99)(rng);
            if (it % 3 == 1) v = std::round(v * 2) / 2;
            return v;
        };
        s.ax = c(ox); s.ay = c(oy); s.bx = c(ox); s.by = c(oy); s.cx = c(ox); s.cy = c(oy);
        s.r0 = .5 * (fabs(s.cx - s.bx) + fabs(s.cy - s.by));
        s.r1 = .5 * (fabs(s.ax - s.cx) + fabs(s.ay - s.cy));
        s.r2 = .5 * (fabs(s.bx - s.ax) + fabs(s.by - s.ay));
        if (it % 7 == 0) s.r0 = s.r1 = s.r2 = 0;
        s.x0 = int(std::floor(std::min({s.ax, s.bx, s.cx})));
        s.x1 = int(std::floor(std::max({s.ax, s.bx, s.cx})));
        s.y0 = int(std::floor(std::min({s.ay, s.by, s.cy})));
        s.y1 = int(std::floor(std::max({s.ay, s.by, s.cy})));
        if (s.x1 - s.x0 > 7 || s.y1 - s.y0 > 7) continue;
        const int ya = s.y0 + (it % 5 == 0), yb = s.y1 - (it % 4 == 0);
        if (ya > yb) continue;
        uint64_t want = 0;
        for (int y = ya; y <= yb; ++y)
            for (int x = s.x0; x <= s.x1; ++x)
                if (covers_ref(s, x, y)) want |= 1ull << (8 * (y - s.y0) + (x - s.x0));
        ++rects;
        bad += lane_bitmap(s, ya, yb) != want;
        if (s.x1 - s.x0 < 4) bad += lane_bitmap4(s, ya, yb) != want;
    }
    // fl(d + r) < 0 against d < -r over special and extreme values (r >= 0
    // or NaN, as conservative radii are).
    {
        const double specials[] = {0.0, -0.0, 1.0, -1.0, 0.5, 4.9e-324, -4.9e-324, 2.2e-308, -2.2e-308,
                                   1e300, -1e300, 1.7976931348623157e308, -1.7976931348623157e308,
                                   INFINITY, -INFINITY, NAN, 3.0000000000000004, -3.0};
        std::uniform_int_distribution<int> pick(0, int(sizeof specials / sizeof specials[0]) - 1);
        for (int it = 0; it < iterations * 4; ++it) {
            double d = specials[pick(rng)], r = std::fabs(specials[pick(rng)]);
            if (it % 3 == 0) d = std::ldexp(std::uniform_real_distribution<double>(-1, 1)(rng), int(rng() % 2100) - 1075);
            if (it % 5 == 0) r = std::ldexp(std::uniform_real_distribution<double>(0, 1)(rng), int(rng() % 2100) - 1075);
            if (it % 7 == 0) d = -r * (1 + std::uniform_int_distribution<int>(-3, 3)(rng) * 2.2e-16);
            bad += ((d + r) < 0) != (d < -r);
        }
    }
    // quad_word over random quads of boxes up to 24 x 24 samples, against
    // the per-sample predicate.
    {
        for (int it = 0; it < iterations; ++it) {
            const double span = it % 3 == 0 ? 8 : 24;
            auto coord = [&]() {
                double v = std::uniform_real_distribution<double>(0, span)(rng) + 40;
                if (it % 4 == 1) v = std::round(v * 2) / 2;
                if (it % 4 == 2) v = std::round(v) + .5;
                return v;
            };
            Setup s;
            s.ax = coord(); s.ay = coord(); s.bx = coord(); s.by = coord(); s.cx = coord(); s.cy = coord();
            s.r0 = .5 * (fabs(s.cx - s.bx) + fabs(s.cy - s.by));
            s.r1 = .5 * (fabs(s.ax - s.cx) + fabs(s.ay - s.cy));
            s.r2 = .5 * (fabs(s.bx - s.ax) + fabs(s.by - s.ay));
            if (it % 7 == 0) s.r0 = s.r1 = s.r2 = 0;
            s.x0 = int(std::floor(std::min({s.ax, s.bx, s.cx})));
            s.x1 = int(std::floor(std::max({s.ax, s.bx, s.cx})));
            s.y0 = int(std::floor(std::min({s.ay, s.by, s.cy})));
            s.y1 = int(std::floor(std::max({s.ay, s.by, s.cy})));
            for (int q = s.x0 & ~3; q <= s.x1; q += 4)
                for (int ty = s.y0 >> 3; ty <= s.y1 >> 3; ++ty) {
                    const int xa = std::max(q, s.x0), xb = std::min(q + 3, s.x1);
                    int ya = std::max(ty * 8, s.y0), yb = std::min(ty * 8 + 7, s.y1);
                    if (it % 5 == 0 && ya < yb) ++ya;  // band-clipped rows
                    uint64_t want = 0;
                    for (int y = ya; y <= yb; ++y)
                        for (int x = xa; x <= xb; ++x)
                            if (covers_ref(s, x, y)) want |= 1ull << (8 * (y & 7) + (x & 7));
                    bad += quad_word(s, q, xa, xb, ya, yb) != want;
                }
        }
    }
    // scatter_bitmap against a per-bit move of 4- and 8-wide bitmaps.
    {
        std::uniform_int_distribution<int> origin(0, 4095);
        for (int it = 0; it < iterations; ++it) {
            const int cols = it % 2 ? 4 : 8;
            uint64_t local = rng();
            for (int r = 0; r < 8; ++r)
                for (int c = cols; c < 8; ++c) local &= ~(1ull << (8 * r + c));
            const int x0 = origin(rng), y0 = origin(rng);
            uint64_t want[4] = {0, 0, 0, 0}, got[4];
            for (int r = 0; r < 8; ++r)
                for (int c = 0; c < 8; ++c)
                    if (local >> (8 * r + c) & 1) {
                        const int x = (x0 & 7) + c, y = (y0 & 7) + r;
                        want[(y >= 8) * 2 + (x >= 8)] |= 1ull << (8 * (y & 7) + (x & 7));
                    }
            scatter_bitmap(local, x0, y0, got);
            for (int k = 0; k < 4; ++k) bad += got[k] != want[k];
        }
    }
    // box_in_frame against the reference's ordered std::min / std::max,
    // isfinite and frame tests (src/evaluate.cpp:155-160), corners swapped
    // for back faces as the setup does, over special coordinates.
    {
        const double size = 1024;
        const double specials[] = {0.0, -0.0, 0.25, 3.5, 511.75, 1023.999, 1024.0, 1024.0000000000002, -1e-300,
                                   1e-300, 1e308, -1e308, INFINITY, -INFINITY, NAN, -NAN, 7.0, 1000.5};
        std::uniform_int_distribution<int> pick(0, int(sizeof specials / sizeof specials[0]) - 1);
        std::uniform_real_distribution<double> any(-8, 1032);
        const auto ref_min = [](double a, double b, double c) { return std::min({a, b, c}); };
        const auto ref_max = [](double a, double b, double c) { return std::max({a, b, c}); };
        for (int it = 0; it < iterations * 4; ++it) {
            double v[6];
            for (double& x : v) x = it % 3 == 0 ? any(rng) : specials[pick(rng)];
            if (it % 5 == 1)
                for (double& x : v) x = any(rng);
            double2 a{v[0], v[1]}, d{v[2], v[3]}, e{v[4], v[5]};
            const double area = edge(a.x, a.y, d.x, d.y, e.x, e.y);
            if (area < 0) std::swap(d, e);
            const double rminx = ref_min(a.x, d.x, e.x), rmaxx = ref_max(a.x, d.x, e.x);
            const double rminy = ref_min(a.y, d.y, e.y), rmaxy = ref_max(a.y, d.y, e.y);
            const bool want = !(!isfinite(rminx) || !isfinite(rmaxx) || !isfinite(rminy) || !isfinite(rmaxy) ||
                                rminx < 0 || rminy < 0 || rmaxx > size || rmaxy > size);
            double minx, maxx, miny, maxy;
            const bool got = box_in_frame(a, d, e, area, size, minx, maxx, miny, maxy);
            bad += got != want;
            if (got && want)
                bad += std::floor(minx) != std::floor(rminx) || std::floor(maxx) != std::floor(rmaxx) ||
                       std::floor(miny) != std::floor(rminy) || std::floor(maxy) != std::floor(rmaxy);
        }
    }
    // span_words over 64-sample spans of faces up to 600 samples across
    // (and thin slivers), full and interval forms, against the per-sample
    // predicate; rows clipped as a band clips them.
    {
        for (int it = 0; it < iterations / 4; ++it) {
            const double span = it % 4 == 0 ? 600 : it % 4 == 1 ? 140 : it % 4 == 2 ? 40 : 12;
            const double o = 8 + (it % 7) * 13.25;
            auto coord = [&]() {
                double v = o + std::uniform_real_distribution<double>(0, span)(rng);
                if (it % 5 == 1) v = std::round(v * 2) / 2;
                if (it % 5 == 2) v = std::round(v) + .5;
                return v;
            };
            Setup s;
            s.ax = coord(); s.ay = coord(); s.bx = coord(); s.by = coord(); s.cx = coord(); s.cy = coord();
            if (it % 6 == 3) s.by = s.ay;  // a horizontal edge (b == 0)
            if (it % 9 == 4) s.cx = s.bx + 1e-9, s.cy = s.by + 300;  // a nearly vertical sliver
            s.r0 = .5 * (fabs(s.cx - s.bx) + fabs(s.cy - s.by));
            s.r1 = .5 * (fabs(s.ax - s.cx) + fabs(s.ay - s.cy));
            s.r2 = .5 * (fabs(s.bx - s.ax) + fabs(s.by - s.ay));
            if (it % 7 == 0) s.r0 = s.r1 = s.r2 = 0;
            s.x0 = int(std::floor(std::min({s.ax, s.bx, s.cx})));
            s.x1 = int(std::floor(std::max({s.ax, s.bx, s.cx})));
            s.y0 = int(std::floor(std::min({s.ay, s.by, s.cy})));
            s.y1 = int(std::floor(std::max({s.ay, s.by, s.cy})));
            for (int sx = s.x0 >> 6; sx <= s.x1 >> 6; ++sx)
                for (int ty = s.y0 >> 3; ty <= s.y1 >> 3; ++ty) {
                    const int xa = std::max(64 * sx, s.x0), xb = std::min(64 * sx + 63, s.x1);
                    int ya = std::max(ty * 8, s.y0), yb = std::min(ty * 8 + 7, s.y1);
                    if (it % 3 == 0 && ya < yb) ++ya;  // band-clipped rows
                    if (it % 4 == 0 && ya < yb) --yb;
                    uint64_t want[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    int in = 0, n = 0;
                    for (int y = ya; y <= yb; ++y)
                        for (int x = xa; x <= xb; ++x) {
                            ++n;
                            if (covers_ref(s, x, y)) {
                                ++in;
                                want[(x - 64 * sx) >> 3] |= 1ull << (8 * (y & 7) + (x & 7));
                            }
                        }
                    const int kind = classify_rect(s, xa, xb, ya, yb);
                    // The two halves of the tile row, as the raster takes them.
                    auto words = [&](bool whole, uint64_t got[8]) {
                        for (int k = 0; k < 8; ++k) got[k] = 0;
                        for (int half = 0; half < 2; ++half) {
                            const int ra = std::max(ya, 8 * ty + 4 * half), rb = std::min(yb, 8 * ty + 4 * half + 3);
                            if (ra > rb) continue;
                            uint32_t w[8];
                            span_half_words(s, 64 * sx, xa, xb, ra, rb, whole, w);
                            for (int k = 0; k < 8; ++k) got[k] |= uint64_t(w[k]) << (32 * half);
                        }
                    };
                    uint64_t got[8];
                    words(false, got);
                    for (int k = 0; k < 8; ++k) bad += got[k] != want[k];
                    if (kind == 1) {
                        words(true, got);
                        for (int k = 0; k < 8; ++k) bad += got[k] != want[k];
                        ++full;
                    }
                    ++rects;
                    if (in && in < n) ++spans_mixed;
                }
        }
    }
    printf("%ld %ld %ld %ld %ld\n", rects, full, empty, bad, spans_mixed);
    return 0;
}
"""


def _function(text: str, signature: str) -> str:
    i = text.index(signature)
    i = text.rfind("\n", 0, i) + 1
    depth, j = 0, text.index("{", i)
    for k in range(j, len(text)):
        depth += {"{": 1, "}": -1}.get(text[k], 0)
        if depth == 0:
            return text[i:k + 1]
    raise AssertionError(signature)


def _struct(text: str, opening: str) -> str:
    i = text.index(opening)
    return text[i:text.index("};", i) + 2]


def _run(tmp_path: Path, source: str, iterations: int) -> tuple[int, ...]:
    src = tmp_path / "check.cpp"
    exe = tmp_path / "check"
    src.write_text(source)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe), str(iterations)], check=True, capture_output=True, text=True).stdout.split()
    return tuple(int(v) for v in out)


def _source(plant: tuple[str, str] | None = None) -> str:
    """The check program; `plant` = (text, replacement) plants a mistake in
    the extracted copy of the header."""
    text = HEADER.read_text()
    if plant:
        assert plant[0] in text
        text = text.replace(plant[0], plant[1])
    parts = [_function(text, "double edge("), _function(text, "double min3("), _function(text, "double max3("),
             _function(text, "bool box_in_frame("), _function(text, "bool covers("),
             _function(text, "int classify_rect("), _function(text, "uint64_t rect_word("),
             _function(text, "uint64_t lane_bitmap("), _function(text, "uint64_t lane_bitmap4("),
             _function(text, "uint64_t quad_word("), _function(text, "void scatter_bitmap("),
             _struct(text, "struct RowEdge {"), _function(text, "bool row_pass("), _function(text, "void clip_row("),
             _function(text, "void transpose4("), _function(text, "void span_half_words(")]
    return PRELUDE + "\n\n".join(parts) + CHECK


needs_gxx = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


@needs_gxx
def test_device_raster_math_agrees_with_every_sample(tmp_path):
    rects, full, empty, bad, spans_mixed = _run(tmp_path, _source(), 40000)
    assert bad == 0
    # Every class occurs, so each branch of the classifier is exercised,
    # and many spans are partly covered (the row intervals decide them).
    assert full > 1000 and empty > 1000 and rects - full - empty > 1000 and spans_mixed > 1000


@needs_gxx
@pytest.mark.parametrize("plant", [
    # Choosing the row end by the wrong sign of A claims coverage the
    # per-sample predicate denies.
    ("a >= 0 ? t1a : t1b, t1max = a >= 0 ? t1b : t1a", "a >= 0 ? t1b : t1a, t1max = a >= 0 ? t1a : t1b"),
    # The second sample of a step past the end of the row.
    ("const bool in1 = (x < s.x1) &", "const bool in1 = (x <= s.x1) &"),
    # The second sample tested at the first one's position.
    ("qx = px + 1.0;", "qx = px;"),
    # Columns past the box's width kept in the four-wide row.
    ("const bool in = (c < w) &", "const bool in = (c <= w) &"),
    # A row's bits shifted by the band row instead of the box row.
    ("| lo) << (8 * (ya - s.y0));", "| lo);"),
    # Row centres not stepped between rows of the four-wide path.
    ("        py += 1.0;\n    }\n    return ((uint64_t(hi)", "    }\n    return ((uint64_t(hi)"),
    # The right tile's columns taken without the complementary mask.
    ("const uint64_t right = (local >> (8 - ox)) & ~keep64;", "const uint64_t right = (local >> (8 - ox));"),
    # A quad's columns masked from the wrong end.
    ("const uint32_t columns = (0xFu >> (3 - (xb - q))) & (0xFu << (xa - q));",
     "const uint32_t columns = (0xFu >> (xb - q)) & (0xFu << (xa - q));"),
    # The quad's right half placed in the tile's left half.
    ("    const int shift = q & 7;\n    uint64_t word = 0;", "    const int shift = 0;\n    uint64_t word = 0;"),
    # NaN corners sent down the unordered min/max path.
    ("    if (!isnan(area)) {\n        minx = fmin", "    if (true) {\n        minx = fmin"),
    # The frame test's upper bound taken as inclusive.
    ("return !(minx < 0 || miny < 0 || maxx > size || maxy > size);",
     "return !(minx < 0 || miny < 0 || maxx >= size || maxy >= size);"),
    # A row's last passing sample overshot by one (the walk tests x, not x + 1).
    ("while (x < r && row_pass(e, x + 1)) ++x;", "while (x < r && row_pass(e, x)) ++x;"),
    # A suffix edge's first sample left one short.
    ("while (x > l && row_pass(e, x - 1)) --x;", "while (x > l && row_pass(e, x)) --x;"),
    # Prefix and suffix edges exchanged.
    ("    if (e.b > 0) {\n        // Last passing", "    if (e.b < 0) {\n        // Last passing"),
    # A constant edge (b == 0) taken as passing everywhere.
    ("        if (!row_pass(e, l)) r = l - 1;\n", ""),
    # A row's mask one column too wide.
    ("(~0ull >> (63 - (r - base)))", "(~0ull >> (62 - (r - base)))"),
    # The byte round of the transpose moving 2-byte blocks.
    ("    t = ((a0 >> 8) ^ a1) & 0x00FF00FFu;", "    t = ((a0 >> 8) ^ a1) & 0x0000FFFFu;"),
    # Tiles 4-7 taken from the low half of the row masks.
    ("    w[4] = uint32_t(m0 >> 32);", "    w[4] = uint32_t(m0);"),
    # A half's rows placed by their tile row instead of their half.
    ("        const int row = y & 3;", "        const int row = y & 7;"),
    # Row terms not stepped between rows of a span.
    ("        py += 1.0;\n    }\n    w[0] = uint32_t(m0);", "    }\n    w[0] = uint32_t(m0);"),
])
def test_check_catches_planted_mistakes(tmp_path, plant):
    _, _, _, bad, _ = _run(tmp_path, _source(plant), 4000)
    assert bad > 0
