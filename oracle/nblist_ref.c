/* FP32 neighborlist predicate and fixed-point rounding, restated in C so the
 * arithmetic is plain IEEE single precision (compiled with -ffp-contract=off).
 * See oracle/nblist.py for the definition; mirrors the predicate of
 * kernels/k_neighborlist.cuh:374-384 (min-image branch) of the reference. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static float min_image(float d, float b, float ib) { return d - b * rintf(d * ib); }

/* mask[r * N + j] = 1 iff j >= r * block and some row atom i of block r has
 * d2(i, j) < c * c in FP32. Coordinates and box are rounded to FP32 first. */
int oracle_fp32_ixn_mask(int N, const double *x, const double *box, double cutoff, int block, uint8_t *mask) {
    const float bx = (float)box[0], by = (float)box[4], bz = (float)box[8];
    const float ibx = 1.0f / bx, iby = 1.0f / by, ibz = 1.0f / bz;
    const float c = (float)cutoff;
    const float c2 = c * c;
    const int nb = (N + block - 1) / block;
    memset(mask, 0, (size_t)nb * N);
    for (int r = 0; r < nb; r++) {
        const int r0 = r * block;
        const int r1 = r0 + block < N ? r0 + block : N;
        for (int j = r0; j < N; j++) {
            const float xj = (float)x[3 * j], yj = (float)x[3 * j + 1], zj = (float)x[3 * j + 2];
            for (int i = r0; i < r1; i++) {
                const float dx = min_image((float)x[3 * i] - xj, bx, ibx);
                const float dy = min_image((float)x[3 * i + 1] - yj, by, iby);
                const float dz = min_image((float)x[3 * i + 2] - zj, bz, ibz);
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (d2 < c2) {
                    mask[(size_t)r * N + j] = 1;
                    break;
                }
            }
        }
    }
    return 0;
}

/* Row subset (neighborlist.cu:215-243): row block r holds rows[r * block ..
 * r * block + block - 1] in the order given; mask[r * N + j] = 1 iff j is not a
 * row atom and some row atom of block r has d2 < c * c in FP32. */
int oracle_fp32_ixn_mask_rows(int N, const double *x, const double *box, double cutoff, int block, int NR,
                              const int *rows, uint8_t *mask) {
    const float bx = (float)box[0], by = (float)box[4], bz = (float)box[8];
    const float ibx = 1.0f / bx, iby = 1.0f / by, ibz = 1.0f / bz;
    const float c = (float)cutoff;
    const float c2 = c * c;
    const int nb = (NR + block - 1) / block;
    memset(mask, 0, (size_t)nb * N);
    uint8_t *is_row = (uint8_t *)calloc((size_t)N, 1);
    if (!is_row) {
        return -1;
    }
    for (int q = 0; q < NR; q++) {
        is_row[rows[q]] = 1;
    }
    for (int r = 0; r < nb; r++) {
        const int r0 = r * block;
        const int r1 = r0 + block < NR ? r0 + block : NR;
        for (int j = 0; j < N; j++) {
            if (is_row[j]) {
                continue;
            }
            const float xj = (float)x[3 * j], yj = (float)x[3 * j + 1], zj = (float)x[3 * j + 2];
            for (int q = r0; q < r1; q++) {
                const int i = rows[q];
                const float dx = min_image((float)x[3 * i] - xj, bx, ibx);
                const float dy = min_image((float)x[3 * i + 1] - yj, by, iby);
                const float dz = min_image((float)x[3 * i + 2] - zj, bz, ibz);
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (d2 < c2) {
                    mask[(size_t)r * N + j] = 1;
                    break;
                }
            }
        }
    }
    free(is_row);
    return 0;
}

/* llrint with the out-of-range / NaN value of x86 (0x8000000000000000). */
uint64_t oracle_round_to_fixed(double v) {
    if (!(v > -9223372036854775808.0 && v < 9223372036854775808.0)) {
        return 0x8000000000000000ull;
    }
    return (uint64_t)llrint(v);
}
