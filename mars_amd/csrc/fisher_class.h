// Fisher classification of one block from its four quadrant statistics.
//
// Shared by the device kernels (fisher.hip) and a host-compiled check of the
// same code (the tests build it with a C++ compiler, MARS_HD empty), so the
// function that runs on the GPU is also checked on the CPU.
//
// The reference classifies a block (crates/mars-search/src/classify.rs) by
//   newclass        classify.rs:94-100: the descending order of the four
//                   quadrant sums (TL, TR, BL, BR) under the exact
//                   three-pass bubble sort of classify.rs:63-75, looked up in
//                   ORDERING (tables.rs:13-38) -> (isometry, class);
//   flips           classify.rs:154-164: the block mapped by that isometry;
//   variance_class  classify.rs:105-115: the same sort over the flipped
//                   block's quadrant variances sum2/n - (sum/n)^2
//                   (classify.rs:79-92) -> one of 24 orders.
// Every isometry maps whole quadrants onto quadrants, so the flipped block's
// quadrant q holds the values of the original quadrant kFlipQuad[iso][q],
// in another order.  A quadrant's sum and sum of squares are exact integers
// in binary64 (at most 256 values of at most 1020), so they do not depend
// on the summation order, and the variances and every comparison come out
// as the reference's bit for bit from the integer sums alone.
#pragma once

#include <stdint.h>

#ifndef MARS_HD
#define MARS_HD __host__ __device__
#endif

namespace mars {
namespace fisher {

constexpr uint32_t kBuckets = 72; // 3 shape classes x 24 variance orders

// tables.rs:13-38 ORDERING, one row per descending quadrant order
// (a, b, c, d): the order packed two bits per entry (a lowest), then the
// isometry and the class newclass reads off the row.
constexpr uint8_t kOrderPerm[24] = {
    0x1b, 0x1e, 0x36, 0x27, 0x2d, 0x39, 0x4b, 0x4e, 0xc6, 0x87, 0x8d, 0xc9,
    0x63, 0x72, 0xd2, 0x93, 0xe1, 0xb1, 0x6c, 0x78, 0xd8, 0x9c, 0xb4, 0xe4,
};
constexpr uint8_t kOrderIso[24] = {3, 5, 5, 7, 1, 1, 3, 5, 2, 7, 1, 4,
                                   3, 2, 2, 7, 4, 4, 6, 6, 6, 0, 0, 0};
constexpr uint8_t kOrderClass[24] = {0, 1, 2, 0, 1, 2, 1, 0, 2, 1, 0, 2,
                                     2, 0, 1, 2, 1, 0, 2, 1, 0, 2, 1, 0};

// tables.rs:42-51 MAPPING: the isometry applied to a candidate is
// kMapping[range isometry][domain isometry].
constexpr uint8_t kMapping[8][8] = {
    {0, 1, 2, 3, 4, 5, 6, 7}, {2, 0, 3, 1, 7, 6, 4, 5}, {1, 3, 0, 2, 6, 7, 5, 4},
    {3, 2, 1, 0, 5, 4, 7, 6}, {4, 7, 6, 5, 0, 3, 2, 1}, {5, 6, 7, 4, 3, 0, 1, 2},
    {6, 4, 5, 7, 1, 2, 0, 3}, {7, 5, 4, 6, 2, 1, 3, 0},
};

// Source quadrant of each quadrant of flips(block, iso) (classify.rs:140-152
// evaluated on the quadrant grid; quadrants numbered TL, TR, BL, BR).
constexpr uint8_t kFlipQuad[8][4] = {
    {0, 1, 2, 3}, {1, 3, 0, 2}, {2, 0, 3, 1}, {3, 2, 1, 0},
    {1, 0, 3, 2}, {2, 3, 0, 1}, {0, 2, 1, 3}, {3, 1, 2, 0},
};

// classify.rs:63-75 three-pass descending bubble sort, then tables.rs:57-63
// match_order: the ORDERING row of the resulting order.
MARS_HD inline uint32_t order_row(double a0, double a1, double a2, double a3)
{
    double a[4] = {a0, a1, a2, a3};
    uint32_t o[4] = {0, 1, 2, 3};
    for (int i = 2; i >= 0; i--)
        for (int j = 0; j <= i; j++)
            if (a[j] < a[j + 1]) {
                const double t = a[j];
                a[j] = a[j + 1];
                a[j + 1] = t;
                const uint32_t u = o[j];
                o[j] = o[j + 1];
                o[j + 1] = u;
            }
    const uint32_t code = o[0] | (o[1] << 2) | (o[2] << 4) | (o[3] << 6);
    uint32_t row = 0;
    for (uint32_t j = 0; j < 24; j++)
        if (kOrderPerm[j] == code) row = j;
    return row;
}

// (bucket * 8 + isometry) of a block whose h x h quadrants (TL, TR, BL, BR)
// have sums sum[q] and sums of squares sq[q]; bucket = class * 24 +
// variance class (fisher.rs:35-40, 59-61).
MARS_HD inline uint32_t classify(const int64_t (&sum)[4], const int64_t (&sq)[4], uint32_t h)
{
    const uint32_t row =
        order_row((double)sum[0], (double)sum[1], (double)sum[2], (double)sum[3]);
    const uint32_t iso = kOrderIso[row], clas = kOrderClass[row];
    const double n = (double)(h * h);
    double v[4];
    for (int q = 0; q < 4; q++) {
        const uint32_t s = kFlipQuad[iso][q];
        const double m = (double)sum[s] / n;
        v[q] = (double)sq[s] / n - m * m;
    }
    const uint32_t vc = order_row(v[0], v[1], v[2], v[3]);
    return (clas * 24 + vc) * 8 + iso;
}

} // namespace fisher
} // namespace mars
