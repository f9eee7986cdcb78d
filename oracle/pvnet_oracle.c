/*
 * pvnet_oracle.c — TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the warm-start network operators of include/pvnet.h, straight from
 * their definitions (the reference has no network: its solver starts from rest,
 * Simulation.cpp:13-15).  The input gradient is computed as the adjoint of the forward
 * convolution (a scatter over taps), not through the flipped weight layout the device path
 * uses, so that the device's weight packing is checked rather than restated.  Only tests/ load
 * this library; the product path never does.
 *
 * Tensors: double, channels-last, row length = channel count, voxel order (b, z, y, x).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

static void tap_off(int ntaps, int tap, int* dz, int* dy, int* dx) {
    if (ntaps == 27) {
        *dz = tap / 9 - 1;
        *dy = (tap / 3) % 3 - 1;
        *dx = tap % 3 - 1;
    } else {
        *dz = *dy = *dx = 0;
    }
}

/* neighbour of voxel (b, z, y, x) at the tap's offset, or -1 outside the box */
static int64_t nbr(const int32_t s[4], int64_t b, int z, int y, int x, int ntaps, int tap) {
    int dz, dy, dx;
    tap_off(ntaps, tap, &dz, &dy, &dx);
    z += dz;
    y += dy;
    x += dx;
    if (z < 0 || z >= s[1] || y < 0 || y >= s[2] || x < 0 || x >= s[3]) return -1;
    return ((b * s[1] + z) * s[2] + y) * (int64_t)s[3] + x;
}

#define FOR_VOXELS(s)                                \
    for (int64_t b = 0; b < (s)[0]; b++)             \
        for (int iz = 0; iz < (s)[1]; iz++)          \
            for (int iy = 0; iy < (s)[2]; iy++)      \
                for (int ix = 0; ix < (s)[3]; ix++)

#define VOX(s) (((b * (s)[1] + iz) * (s)[2] + iy) * (int64_t)(s)[3] + ix)

/* y[m][n] = act(bias[n] + sum_{tap,ci} x[m+off(tap)][ci] w[n][tap*cin+ci]); s = {nb, d, h, w} */
void onet_conv_fwd(const int32_t s[4], int ntaps, int cin, int cout, const double* x, const double* w,
                   const double* bias, int relu, double* y) {
    const int K = ntaps * cin;
    FOR_VOXELS(s) {
        const int64_t m = VOX(s);
        for (int n = 0; n < cout; n++) {
            double acc = bias ? bias[n] : 0.0;
            for (int tap = 0; tap < ntaps; tap++) {
                const int64_t q = nbr(s, b, iz, iy, ix, ntaps, tap);
                if (q < 0) continue;
                for (int ci = 0; ci < cin; ci++) acc += x[q * cin + ci] * w[(int64_t)n * K + tap * cin + ci];
            }
            y[m * cout + n] = (relu && acc < 0.0) ? 0.0 : acc;
        }
    }
}

/* dx = adjoint of the forward convolution applied to dy: dx[m+off][ci] += dy[m][n] w[n][tap*cin+ci] */
void onet_conv_dgrad(const int32_t s[4], int ntaps, int cin, int cout, const double* dy, const double* w,
                     double* dx) {
    const int K = ntaps * cin;
    const int64_t M = (int64_t)s[0] * s[1] * s[2] * s[3];
    memset(dx, 0, sizeof(double) * (size_t)(M * cin));
    FOR_VOXELS(s) {
        const int64_t m = VOX(s);
        for (int tap = 0; tap < ntaps; tap++) {
            const int64_t q = nbr(s, b, iz, iy, ix, ntaps, tap);
            if (q < 0) continue;
            for (int ci = 0; ci < cin; ci++) {
                double acc = 0.0;
                for (int n = 0; n < cout; n++) acc += dy[m * cout + n] * w[(int64_t)n * K + tap * cin + ci];
                dx[q * cin + ci] += acc;
            }
        }
    }
}

/* dw[n][tap*cin+ci] = sum_m dy[m][n] x[m+off][ci];  db[n] = sum_m dy[m][n] */
void onet_conv_wgrad(const int32_t s[4], int ntaps, int cin, int cout, const double* x, const double* dy,
                     double* dw, double* db) {
    const int K = ntaps * cin;
    memset(dw, 0, sizeof(double) * (size_t)cout * K);
    if (db) memset(db, 0, sizeof(double) * (size_t)cout);
    FOR_VOXELS(s) {
        const int64_t m = VOX(s);
        for (int n = 0; n < cout; n++) {
            const double g = dy[m * cout + n];
            if (db) db[n] += g;
            for (int tap = 0; tap < ntaps; tap++) {
                const int64_t q = nbr(s, b, iz, iy, ix, ntaps, tap);
                if (q < 0) continue;
                for (int ci = 0; ci < cin; ci++) dw[(int64_t)n * K + tap * cin + ci] += g * x[q * cin + ci];
            }
        }
    }
}

/* 2x2x2 average pooling; s is the fine shape */
void onet_pool2_fwd(const int32_t s[4], int c, const double* x, double* y) {
    const int32_t sc[4] = {s[0], s[1] / 2, s[2] / 2, s[3] / 2};
    FOR_VOXELS(sc) {
        const int64_t o = VOX(sc);
        for (int ch = 0; ch < c; ch++) {
            double acc = 0.0;
            for (int k = 0; k < 8; k++) {
                const int64_t f = ((b * s[1] + 2 * iz + (k >> 2)) * s[2] + 2 * iy + ((k >> 1) & 1)) * (int64_t)s[3] +
                                  2 * ix + (k & 1);
                acc += x[f * c + ch];
            }
            y[o * c + ch] = acc / 8.0;
        }
    }
}

void onet_pool2_bwd(const int32_t s[4], int c, const double* dy, double* dx) {
    const int32_t sc[4] = {s[0], s[1] / 2, s[2] / 2, s[3] / 2};
    FOR_VOXELS(sc) {
        const int64_t o = VOX(sc);
        for (int k = 0; k < 8; k++) {
            const int64_t f = ((b * s[1] + 2 * iz + (k >> 2)) * s[2] + 2 * iy + ((k >> 1) & 1)) * (int64_t)s[3] +
                              2 * ix + (k & 1);
            for (int ch = 0; ch < c; ch++) dx[f * c + ch] = dy[o * c + ch] / 8.0;
        }
    }
}

/* nearest 2x upsampling; s is the fine shape */
void onet_up2_fwd(const int32_t s[4], int c, const double* x, double* y) {
    FOR_VOXELS(s) {
        const int64_t f = VOX(s);
        const int64_t o = ((b * (s[1] / 2) + iz / 2) * (s[2] / 2) + iy / 2) * (int64_t)(s[3] / 2) + ix / 2;
        for (int ch = 0; ch < c; ch++) y[f * c + ch] = x[o * c + ch];
    }
}

void onet_up2_bwd(const int32_t s[4], int c, const double* dy, double* dx) {
    const int64_t Mc = (int64_t)s[0] * (s[1] / 2) * (s[2] / 2) * (s[3] / 2);
    memset(dx, 0, sizeof(double) * (size_t)(Mc * c));
    FOR_VOXELS(s) {
        const int64_t f = VOX(s);
        const int64_t o = ((b * (s[1] / 2) + iz / 2) * (s[2] / 2) + iy / 2) * (int64_t)(s[3] / 2) + ix / 2;
        for (int ch = 0; ch < c; ch++) dx[o * c + ch] += dy[f * c + ch];
    }
}

/* distance from each voxel to the nearest solid voxel (code 0) of its box, by brute force over
 * every solid voxel of the box; 255 when the box has none */
void onet_wall_distance(const int32_t s[4], const uint8_t* code, double* dist) {
    const int64_t per = (int64_t)s[1] * s[2] * s[3];
    FOR_VOXELS(s) {
        const int64_t m = VOX(s);
        double best = 1e300;
        for (int64_t q = 0; q < per; q++) {
            if (code[b * per + q] != 0) continue;
            const int qz = (int)(q / ((int64_t)s[2] * s[3])), qy = (int)((q / s[3]) % s[2]), qx = (int)(q % s[3]);
            const double d2 = (double)(qz - iz) * (qz - iz) + (double)(qy - iy) * (qy - iy) + (double)(qx - ix) * (qx - ix);
            if (d2 < best) best = d2;
        }
        dist[m] = best > 255.0 * 255.0 ? 255.0 : sqrt(best);
    }
}

/* the eight input features of a geometry code volume (see pvnet_features) */
void onet_features(const int32_t s[4], const uint8_t* code, const double* prm, const double* dist, double* out) {
    FOR_VOXELS(s) {
        const int64_t m = VOX(s);
        const int cd = code[m];
        const double fl = cd > 0 ? 1.0 : 0.0;
        double* f = out + m * 8;
        f[0] = fl;
        f[1] = cd == 2 ? 1.0 : 0.0;
        f[2] = cd == 3 ? 1.0 : 0.0;
        f[3] = fl * (prm[3 * b + 1] - 0.5) / 0.4;
        f[4] = fl * prm[3 * b] / 0.05;
        f[5] = fl * ((double)ix / (double)(s[3] > 1 ? s[3] - 1 : 1) - 0.5);
        f[6] = fl * dist[m] / 16.0;
        f[7] = fl * prm[3 * b + 2] / (3.14159265358979323846 * 256.0);
    }
}

/* masked MSE: L = sum_{fluid m, c<nc} wc[c] (y[m][c] - t[m][c])^2 / (n_fluid nc); dy = dL/dy
 * (rows of length ldy, zero beyond nc and on solid voxels).  Returns L. */
double onet_mse_loss(int64_t M, int nc, const double* y, int ldy, const double* t, const uint8_t* code,
                     const double* wc, double* dy) {
    double e = 0.0, cnt = 0.0;
    for (int64_t m = 0; m < M; m++) {
        if (code[m] == 0) continue;
        cnt += 1.0;
        for (int c = 0; c < nc; c++) {
            const double r = y[m * ldy + c] - t[m * nc + c];
            e += wc[c] * r * r;
        }
    }
    for (int64_t m = 0; m < M; m++)
        for (int c = 0; c < ldy; c++)
            dy[m * ldy + c] = (code[m] && c < nc && cnt > 0) ? 2.0 * wc[c] * (y[m * ldy + c] - t[m * nc + c]) / (cnt * nc) : 0.0;
    return cnt > 0 ? e / (cnt * nc) : 0.0;
}

/* Adam, bias-corrected (Kingma & Ba 2015, Algorithm 1) */
void onet_adam(int64_t n, double* p, const double* g, double* m, double* v, double lr, double b1, double b2,
               double eps, int t) {
    const double c1 = 1.0 - pow(b1, t), c2 = 1.0 - pow(b2, t);
    for (int64_t i = 0; i < n; i++) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        p[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
    }
}
